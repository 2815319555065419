"""Trajectory frames packed on the device (cymg_write_frame: Simul::writeObjects
in binary mode, simul_file.cc:402-489) against the oracle's frame: byte for byte
in both forms (floats as the reference writes them, and the lossless doubles),
in 2D and 3D, with fibers, Solids, long reference forms and sections of other
objects around them; after solved steps (the oracle given the device's points);
after a re-registration in a new order; and the lossless frame read back by
cymg_read_frame gives the device's points bit for bit."""
import numpy as np
import pytest

from test_oracle_frame import DATE

pytestmark = pytest.mark.gpu

PARAMS = dict(time_step=0.005, kT=0.0042, viscosity=0.02)
BEFORE = b"\n#section space\ne\x01\x01\x00"
AFTER = b"\n#section couple\nc\x01\x02\x00"


def _scene(dim, seed, nfib=12, nsol=3):
    rng = np.random.default_rng(seed)
    fibers, solids, recs = [], [], []
    for f in range(nfib):
        n = int(rng.integers(2, 20))
        cut = float(rng.uniform(0.1, 0.5))
        d = rng.standard_normal(dim)
        d /= np.linalg.norm(d)
        pts = [rng.uniform(-2, 2, dim)]
        for _ in range(n - 1):
            d = d + 0.3 * rng.standard_normal(dim)
            d /= np.linalg.norm(d)
            pts.append(pts[-1] + cut * d)
        fibers.append((np.array(pts), cut))
        recs.append((int(rng.integers(1, 70000)), int(rng.choice([0, 0, 0, -1, 5])), int(rng.integers(0, 300)),
                     int(rng.integers(0, 2**32)), float(rng.uniform(0.1, 1)), float(rng.uniform(-5, 5))))
    for k in range(nsol):
        n = int(rng.integers(2, 9))
        rad = rng.uniform(0.05, 0.3, n)
        p = rng.uniform(-2, 2, dim) + 0.4 * rng.standard_normal((n, dim))
        solids.append((rad, p))
        recs.append((int(rng.integers(1, 70000)), 0, int(rng.integers(0, 3)), 0, 0.0, 0.0))
    return fibers, solids, recs


def _load(m, fibers, solids):
    for p, cut in fibers:
        m.add_fiber(len(p), cut, 20.0)
    for rad, p in solids:
        m.add_solid(rad, p.ravel())
    return np.concatenate([p.ravel() for p, _ in fibers] + [p.ravel() for _, p in solids])


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("rb", [4, 8])
def test_frame_bytes_equal_the_oracle(pkg, oracle, dim, rb):
    fibers, solids, recs = _scene(dim, 7 + dim)
    gpu = pkg.Meca(dim, 0)
    orc = oracle.OracleMeca(dim, PARAMS)
    x = _load(gpu, fibers, solids)
    _load(orc, fibers, solids)
    gpu.set_params(**PARAMS)
    gpu.prepare(x)
    orc.prepare(x)
    gpu.set_frame_records(recs)
    print(f"frame dim={dim} rb={rb}", flush=True)
    a = gpu.write_frame(2.5, DATE, rb, BEFORE, AFTER)
    b = orc.write_frame(recs, 2.5, DATE, rb, BEFORE, AFTER)
    assert a == b
    # switching forms rebuilds the packed layout
    other = 12 - rb
    assert gpu.write_frame(2.5, DATE, other) == orc.write_frame(recs, 2.5, DATE, other)
    assert gpu.write_frame(2.5, DATE, rb, BEFORE, AFTER) == b
    gpu.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_frames_after_solved_steps(pkg, oracle, dim):
    fibers, solids, recs = _scene(dim, 20 + dim, nfib=30, nsol=4)
    gpu = pkg.Meca(dim, 0)
    orc = oracle.OracleMeca(dim, PARAMS)
    x = _load(gpu, fibers, solids)
    _load(orc, fibers, solids)
    gpu.set_params(**PARAMS)
    gpu.set_frame_records(recs)
    rng = np.random.default_rng(dim)
    for step in range(3):
        gpu.prepare(x if step == 0 else None)
        gpu.inter_link_ee(0, len(fibers[0][0]) + 1, 50.0)
        gpu.solve(1e-8, rng.standard_normal(x.size), tol_mode=pkg.TOL_RELATIVE)
        pts = gpu.fiber_points()
        print(f"frame after step {step} dim={dim}", flush=True)
        lossless = gpu.write_frame(0.005 * (step + 1), DATE, pkg.FRAME_DOUBLE)
        got = pkg.read_frame(lossless, dim)
        assert np.array_equal(got["points"].ravel(), pts)
        assert got["time"] == 0.005 * (step + 1)
        orc.prepare(pts)
        assert gpu.write_frame(1.0, DATE, pkg.FRAME_FLOAT) == orc.write_frame(recs, 1.0, DATE, 4)
        assert lossless == orc.write_frame(recs, 0.005 * (step + 1), DATE, 8)
    gpu.close()


def test_frame_records_follow_the_registration(pkg, oracle):
    dim = 3
    fibers, solids, recs = _scene(dim, 31)
    gpu = pkg.Meca(dim, 0)
    gpu.set_params(**PARAMS)
    x = _load(gpu, fibers, solids)
    gpu.prepare(x)
    gpu.set_frame_records(recs)
    gpu.write_frame(0.0, DATE)
    # a new registration: fibers reversed, one fewer
    order = list(range(len(fibers) - 1))[::-1]
    f2 = [fibers[k] for k in order]
    r2 = [recs[k] for k in order] + recs[len(fibers):]
    gpu.clear()
    x2 = _load(gpu, f2, solids)
    gpu.prepare(x2)
    with pytest.raises(pkg.CymgError, match="frame records"):
        gpu.write_frame(0.0, DATE)
    gpu.set_frame_records(r2)
    orc = oracle.OracleMeca(dim, PARAMS)
    _load(orc, f2, solids)
    orc.prepare(x2)
    assert gpu.write_frame(0.0, DATE) == orc.write_frame(r2, 0.0, DATE, 4)
    gpu.close()
