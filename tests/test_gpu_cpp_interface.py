"""The C++ Meca interface (cytosim-2017_amd/host/meca.hpp), driven by a
compiled C++ program the way Simul::solve drives the reference's Meca (every
step: clear, add fibers then Solids, prepare, interactions, solve), gives the
oracle's fiber and Solid positions."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, oracle_assemble, oracle_from_scene, rel_err
from frame_reader import read_frame

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "cytosim-2017_amd", "_build", "drive_meca")


def write_scene(path, sc, nsteps, tolerance, tol_mode):
    with open(path, "w") as f:
        f.write(f"{sc.dim} {sc.nfib} {sc.nlinks} {nsteps} {sc.nsol} {sc.naster_links}\n")
        f.write(f"{float(sc.time_step)!r} {float(sc.kT)!r} {float(sc.viscosity)!r} {float(tolerance)!r} "
                f"{tol_mode} {float(sc.confine_radius)!r} {float(sc.confine_stiffness)!r}\n")
        P = sc.points
        for fi in range(sc.nfib):
            o, n = sc.dim * int(sc.mat_index[fi]), sc.dim * int(sc.nb_points[fi])
            f.write(f"{int(sc.nb_points[fi])} {float(sc.segmentation[fi])!r} {float(sc.rigidity[fi])!r}\n")
            f.write(" ".join(repr(float(x)) for x in P[o:o + n]) + "\n")
        for k in range(sc.nsol):
            f.write(f"{sc.sol_radius[k].size} {int(sc.sol_timer[k])}\n")
            f.write(" ".join(repr(float(x)) for x in sc.sol_radius[k]) + "\n")
            f.write(" ".join(repr(float(x)) for x in sc.solid_points(k)) + "\n")
        for k in range(sc.nlinks):
            f.write(f"{int(sc.link_fiber_a[k])} {float(sc.link_abs_a[k])!r} {int(sc.link_fiber_b[k])} "
                    f"{float(sc.link_abs_b[k])!r} {float(sc.link_stiffness[k])!r}\n")
        for k in range(sc.naster_links):
            f.write(f"{int(sc.aster_kind[k])} {int(sc.aster_s1[k])} {int(sc.aster_s2[k])} {int(sc.aster_fiber[k])} "
                    f"{float(sc.aster_abs[k])!r} {float(sc.aster_w[k])!r}\n")
        for s in range(nsteps):
            f.write(" ".join(repr(float(x)) for x in sc.gauss(s)) + "\n")


def test_cpp_driver_matches_oracle(scenes, oracle, tmp_path):
    assert os.path.exists(DRIVER), "build the package first"
    sc = scenes.mini(dim=3, seed=41, nfib=10, nlinks=30, radius=2.0)
    nsteps, tol = 2, 1e-11
    write_scene(tmp_path / "scene.txt", sc, nsteps, tol, 1)
    out = tmp_path / "out.txt"
    r = subprocess.run([DRIVER, str(tmp_path / "scene.txt"), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    vals = open(out).read().split()
    pts = np.array([float(v) for v in vals[2 * nsteps:]])
    orc = oracle_from_scene(oracle, sc)
    for s in range(nsteps):
        if s:
            orc.prepare(orc.fiber_points())
            oracle_assemble(orc, sc)
        orc.solve(tol, sc.gauss(s), tol_mode=1)
    assert rel_err(pts, orc.fiber_points()) < 1e-9


@pytest.mark.parametrize("short", [False, True])
def test_cpp_driver_asters_match_oracle(scenes, oracle, tmp_path, short):
    """Solids through the C++ interface, re-registered every step as
    Simul::solve does (their reshape timers kept by the host objects): 9 steps
    so that every Solid is reshaped at least once.  With `short`, some fibers
    are shorter than the clamp distance and are held by their plus end
    (aster.cc:77-84)."""
    assert os.path.exists(DRIVER), "build the package first"
    sc = scenes.mini_asters(dim=3, seed=42, lmin=0.3, lmax=1.4, core=0.8) if short else \
        scenes.mini_asters(dim=3, seed=42)
    if short:
        assert (sc.aster_kind == 2).any() and (sc.aster_kind == 1).any()
    nsteps, tol = 9, 1e-11
    write_scene(tmp_path / "scene.txt", sc, nsteps, tol, 1)
    out = tmp_path / "out.txt"
    frame = tmp_path / "frame.cmo"
    r = subprocess.run([DRIVER, str(tmp_path / "scene.txt"), str(out), str(frame)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    vals = open(out).read().split()
    pts = np.array([float(v) for v in vals[2 * nsteps:]])
    orc = oracle_from_scene(oracle, sc)
    for s in range(nsteps):
        if s:
            orc.prepare(orc.fiber_points())
            oracle_assemble(orc, sc)
        orc.solve(tol, sc.gauss(s), tol_mode=1)
    assert rel_err(pts, orc.fiber_points()) < 1e-9
    # the lossless frame the driver wrote through the C++ interface holds the same points
    fr = read_frame(open(frame, "rb").read(), 3)
    assert np.array_equal(np.concatenate([o["points"].ravel() for o in fr.objects]), pts)
    assert [o["tag"] for o in fr.objects].count("d") == sc.nsol
