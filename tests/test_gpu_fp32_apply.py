"""FP32 blocks applied in FP32 arithmetic (CYMG_OPT_FP32_APPLY): the operand
rounded to FP32 and FP32 accumulation, in Meca::precondition and in the solve
(four-launch iteration), against an FP32 product in numpy and the FP64 oracle."""
import numpy as np
import pytest

from conftest import oracle_from_scene, rel_err
from test_gpu_parity import _step_pair, gpu_from_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim", [2, 3])
def test_fp32_apply_precondition(pkg, scenes, dim):
    """Each fiber's block applied in FP32: equal to numpy's float32 product of
    the stored FP32 inverse and the FP32-rounded operand within FP32 rounding,
    and different (at FP32 rounding level) from the FP64-accumulated apply of
    the same blocks."""
    sc = scenes.mini(dim=dim, seed=31, nfib=10, nlinks=40)
    gpu = gpu_from_scene(pkg, sc)
    gpu.setup(sc.gauss(0))
    gpu.compute_preconditioner(pkg.PRECOND_FP32)
    x = np.random.default_rng(7).standard_normal(sc.size)
    y64 = gpu.precondition(x)
    gpu.set_option(pkg.OPT_FP32_APPLY, 1)
    assert gpu.get_option(pkg.OPT_FP32_APPLY) == 1
    y32 = gpu.precondition(x)
    ref = np.empty(sc.size)
    o = 0
    for f in range(sc.nfib):
        bs = dim * int(sc.nb_points[f])
        Bi = gpu.get_block(f, inverse=True).astype(np.float32)
        ref[o:o + bs] = (Bi @ x[o:o + bs].astype(np.float32)).astype(np.float64)
        o += bs
    assert rel_err(y32, ref) < 1e-5
    assert rel_err(y32, y64) < 1e-5
    assert np.abs(y32 - y64).max() > 0.0, "the FP32 apply must differ from the FP64-accumulated one"
    gpu.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_fp32_apply_steps(pkg, scenes, oracle, dim):
    """Two crosslinked steps with FP32 blocks applied in FP32 arithmetic (the
    four-launch iteration): converged, positions within 1e-6 of the FP64
    oracle's at a solver tolerance of 1e-9."""
    sc = scenes.mini(dim=dim, seed=32, nfib=12, nlinks=40, radius=2.0)
    gpu = gpu_from_scene(pkg, sc)
    gpu.set_option(pkg.OPT_FP32_APPLY, 1)
    orc = oracle_from_scene(oracle, sc)
    from conftest import oracle_assemble
    for s in range(2):
        g = sc.gauss(s)
        r1 = gpu.solve(1e-9, g, tol_mode=pkg.TOL_RELATIVE, precondition=pkg.PRECOND_FP32)
        r2 = orc.solve(1e-12, g, tol_mode=1)
        print(f"fp32 apply dim={dim} step {s}: iterations {r1['iterations']} (oracle {r2['iterations']})", flush=True)
        assert r1["converged"] and r2["converged"]
        assert rel_err(gpu.fiber_points(), orc.fiber_points()) < 1e-6
        if s == 0:
            gpu.prepare(None)
            pkg.assemble(gpu, sc, upload_couples=False)
            orc.prepare(orc.fiber_points())
            oracle_assemble(orc, sc)
    gpu.close()
