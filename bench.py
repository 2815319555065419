"""Benchmark of the MI355X implicit mechanics step (one GPU).

Workload: 3D, 1000 fibers with lengths uniform in [2, 15] um (segmentation 0.5,
5-31 points), rigidity 20, confined in a sphere of radius 10 (stiffness 200),
20000 crosslinkers (stiffness 100) bound at crossings within 0.05 um,
viscosity 0.02, time_step 0.005, kT 0.0042.  A step is one implicit step:
Meca::prepare + device assembly (confinement, couples) + Meca::solve at
relative residual |r|_inf/|b|_inf < 1e-6 + apply (update, forces, reshape).
Fiber points and the per-step N(0,1) Brownian draws are resident in HBM before
the timed region.

Modes: fp64 (zero seed, FP64 preconditioner blocks: `value`), fp32 (FP32
blocks) and learned-fp32 (FP32 blocks, seeded by the per-fiber network trained
on the device from an untimed rollout of FP64 steps; rollout and training time
are reported beside it, not inside the timed steps).

Prints one JSON line.  --gpus N with N > 1 exits: every box has one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_matvec(sc, kb_full, kc_full):
    """Compulsory HBM bytes of one fused operator launch (Meca::multiply):
    x in and y out (16n), CSR value+column of every stored entry of the full
    symmetric mB/mC rows (12 per entry), the gathered coupled x (8n upper
    bound: each point's coordinates once), per segment rfDiff (8*DIM), pttrf
    factors (16) and P' coefficient (8)."""
    n = sc.size
    S = sc.npts - sc.nfib
    return 16 * n + 12 * (kb_full + kc_full) + 8 * n + (8 * sc.dim + 24) * S


def algorithmic_bytes_iteration_launch(sc, kb_full, kc_full, block_bytes):
    """Compulsory HBM bytes of one launch of the two-launch BiCGStab
    iteration, the mean of its two launches P and Q:
      vectors: P reads r,t,shat,w,y,phat,z,x,rt and writes x,r,y,phat',v,z'
      (15 n doubles), Q reads r,v,x,phat,y,z,rt and writes r,x,shat,t,w (12 n);
      operator: CSR column+value of the full mB/mC rows (12 per entry), row
      extents (4 per point row and per dof row), per segment rfDiff (8*DIM),
      JJ' diagonal and off-diagonal (16) and P' coefficient (8);
      block inverse: sum over fibers of bs^2 entries of the stored type."""
    n = sc.size
    S = sc.npts - sc.nfib
    vectors = 8 * n * (15 + 12) / 2
    operator = 12 * (kb_full + kc_full) + 4 * (sc.npts + n) + (8 * sc.dim + 24) * S
    return vectors + operator + block_bytes


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nfib", type=int, default=1000)
    ap.add_argument("--nlinks", type=int, default=20000)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--cpu-steps", type=int, default=6, help="oracle steps timed for cpu_baseline (0: skip)")
    ap.add_argument("--modes", default="fp64,fp32,learned-fp32",
                    help="fp64 / fp32: zero-seeded BCGSP with FP64 / FP32 preconditioner blocks; "
                         "learned-fp32: seeded by the trained per-fiber network, FP32 blocks")
    ap.add_argument("--autotune", type=int, default=1,
                    help="run the library's autotuner (tile shape, block-inverse method) before the warmup")
    ap.add_argument("--guess-rollout", type=int, default=32,
                    help="untimed FP64 steps recorded as training samples for the learned guess")
    ap.add_argument("--guess-train-steps", type=int, default=1000)
    ap.add_argument("--guess-batch", type=int, default=512)
    ap.add_argument("--guess-lr", type=float, default=1e-3)
    ap.add_argument("--timed-launches", type=int, default=40,
                    help="iteration launches timed with HIP events for the roofline")
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per iteration launch measured with rocprofv3 counters (reported as is)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run a plain device-to-device copy of this many MiB (counter calibration)")
    ap.add_argument("--batched-steps", type=int, default=3,
                    help="steps of the 64-system batched workload (0: skip)")
    ap.add_argument("--grid-places", type=int, default=50000,
                    help="places of the timed FiberGrid query (0: skip)")
    ap.add_argument("--aster-steps", type=int, default=5,
                    help="timed steps of the aster workload (10 Solids + 2000 fibers, 50000 motors; 0: skip)")
    return ap.parse_args()


def train_guess(pkg, m, sc, args, tol, torch):
    """Record a rollout of zero-seeded FP64 steps (Brownian draws of their own,
    distinct from the timed steps'), then train the network on the device."""
    t0 = time.perf_counter()
    m.guess_record(args.guess_rollout * sc.nfib)
    for s in range(args.guess_rollout):
        if s > 0:
            m.prepare(None)
            pkg.assemble(m, sc, upload_couples=False)
        m.solve(tol, sc.gauss(1000 + s), tol_mode=pkg.TOL_RELATIVE, precondition=pkg.PRECOND_FP64)
    m.guess_record(0)
    t1 = time.perf_counter()
    m.guess_init(1)
    first = m.guess_train(1, batch=args.guess_batch, lr=args.guess_lr, seed=1)
    last = m.guess_train(args.guess_train_steps, batch=args.guess_batch, lr=args.guess_lr, seed=2)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return {"rollout_steps": args.guess_rollout, "samples": args.guess_rollout * sc.nfib,
            "rollout_s": round(t1 - t0, 3), "train_steps": args.guess_train_steps, "batch": args.guess_batch,
            "train_s": round(t2 - t1, 3), "loss_first": first, "loss_last": last}


def run_mode(pkg, sc, precond, steps, warmup, tol, gauss_dev, torch, guess=0, train=None, autotune=True):
    m = pkg.Meca(sc.dim, 0)
    pkg.load_scene(m, sc)
    iters = []
    tuned = None
    if autotune:
        # untimed: the autotuner picks the tile shape and block-inverse method on this system
        m.setup(gauss_device_ptr=gauss_dev[0].data_ptr())
        m.compute_preconditioner(precond if precond else pkg.PRECOND_FP64)
        tuned = m.autotune(reps=20)
    info = train(m) if train else None
    if train:
        # the timed steps start from the scene's initial state, like the other modes
        m.prepare(sc.points)
        pkg.assemble(m, sc, upload_couples=False)

    def one(step):
        if step > 0:
            m.prepare(None)
            pkg.assemble(m, sc, upload_couples=False)
        rep = m.solve(tol, tol_mode=pkg.TOL_RELATIVE, precondition=precond, guess=guess,
                      gauss_device_ptr=gauss_dev[step % len(gauss_dev)].data_ptr())
        if guess and rep["guess_used"] != guess:
            raise RuntimeError("learned guess not used")
        iters.append(rep["iterations"])
        return rep

    for s in range(warmup):
        one(s)
    torch.cuda.synchronize()
    iters.clear()
    t0 = time.perf_counter()
    for s in range(steps):
        one(warmup + s)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    if tuned:
        info = dict(info or {}, autotune=tuned)
    return m, dt, iters, info


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one GPU, so --gpus {args.gpus} "
              "is not run (the multi-GPU sharding is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    pkg = __graft_entry__.load_package()
    import importlib
    scenes = importlib.import_module(__graft_entry__.PKG_NAME + ".scene")
    print(f"bench: generating scene nfib={args.nfib} nlinks={args.nlinks}", file=sys.stderr, flush=True)
    sc = scenes.config2(seed=2, nfib=args.nfib, nlinks=args.nlinks)
    nsteps_total = args.warmup + args.steps
    gauss_dev = [torch.from_numpy(sc.gauss(s)).to("cuda:0") for s in range(min(nsteps_total, 8))]
    torch.cuda.synchronize()

    results = {}
    for mode in args.modes.split(","):
        precond = {"fp64": pkg.PRECOND_FP64, "fp32": pkg.PRECOND_FP32, "learned-fp32": pkg.PRECOND_FP32}[mode]
        learned = mode.startswith("learned")
        print(f"bench: mode {mode} warmup={args.warmup} steps={args.steps}", file=sys.stderr, flush=True)
        m, dt, iters, info = run_mode(pkg, sc, precond, args.steps, args.warmup, args.tol, gauss_dev, torch,
                                      guess=pkg.GUESS_LEARNED if learned else pkg.GUESS_ZERO,
                                      train=(lambda mm: train_guess(pkg, mm, sc, args, args.tol, torch))
                                      if learned else None, autotune=bool(args.autotune))
        results[mode] = dict(seconds=dt, steps_per_s=args.steps / dt, ms_per_step=1e3 * dt / args.steps,
                             mean_iters=float(np.mean(iters)), meca=m, training=info)
        print(f"bench: {mode} {args.steps / dt:.2f} steps/s, mean iterations {np.mean(iters):.1f}"
              + (f", training {info}" if info else ""), file=sys.stderr, flush=True)
        if mode != "fp64":
            m.close()

    # dominant-kernel timing on the final state of the FP64 run
    m = results["fp64"]["meca"]
    m.setup(gauss_device_ptr=gauss_dev[0].data_ptr())
    mv_ms = m.time_kernel(0, reps=200)
    m.compute_preconditioner(pkg.PRECOND_FP64)
    pc_ms = m.time_kernel(1, reps=200)
    build_ms = m.time_kernel(2, reps=5)
    build_only_ms = m.time_kernel(4, reps=5)
    build_lu_ms = m.time_kernel(5, reps=5)
    rB, cB, vB = m.matrix_entries(0)
    rC, cC, vC = m.matrix_entries(1)
    kb_full = 2 * int((rB != cB).sum()) + int((rB == cB).sum())
    kc_full = 2 * int((rC != cC).sum()) + int((rC == cC).sum())
    mv_bytes = algorithmic_bytes_matvec(sc, kb_full, kc_full)
    mv_achieved = mv_bytes / (mv_ms * 1e-3) / 1e9
    nblk = int(((sc.dim * sc.nb_points.astype(np.int64)) ** 2).sum())
    m.compute_preconditioner(pkg.PRECOND_FP64)
    it_ms = m.time_kernel(6, reps=args.timed_launches)
    it_bytes = algorithmic_bytes_iteration_launch(sc, kb_full, kc_full, 8 * nblk)
    achieved = it_bytes / (it_ms * 1e-3) / 1e9
    it32_ms = it32_achieved = None
    if "fp32" in results:
        m.compute_preconditioner(pkg.PRECOND_FP32)
        it32_ms = m.time_kernel(6, reps=args.timed_launches)
        it32_achieved = algorithmic_bytes_iteration_launch(sc, kb_full, kc_full, 4 * nblk) / (it32_ms * 1e-3) / 1e9

    if args.copy_calibration_mb > 0:
        nb = args.copy_calibration_mb << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    batched = None
    if args.batched_steps > 0:
        print("bench: generating 64 independent systems", file=sys.stderr, flush=True)
        batch = scenes.config5(seed=5)
        big = batch.scene
        gb = [torch.from_numpy(big.gauss(s)).to("cuda:0") for s in range(2)]
        mb = pkg.Meca(3, 0)
        pkg.load_batch(mb, batch)
        its = []
        btune = None
        for s in range(1 + args.batched_steps):
            if s == 1:
                if args.autotune:
                    # tile shape and block-inverse method for this workload (untimed)
                    btune = mb.autotune(reps=6)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
            if s > 0:
                mb.prepare(None)
                pkg.assemble(mb, big, upload_couples=False)
            rep = mb.solve(args.tol, tol_mode=pkg.TOL_RELATIVE, precondition=pkg.PRECOND_FP64,
                           gauss_device_ptr=gb[s % 2].data_ptr())
            if s > 0:
                its.append(rep["iterations"])
        torch.cuda.synchronize()
        bdt = time.perf_counter() - t0
        # full-work iteration launch over all 64 systems (every tile iterating)
        mb.setup(gauss_device_ptr=gb[0].data_ptr())
        mb.compute_preconditioner(pkg.PRECOND_FP64)
        b_ms = mb.time_kernel(6, reps=10)
        rB, cB, _ = mb.matrix_entries(0)
        rC, cC, _ = mb.matrix_entries(1)
        b_kb = 2 * int((rB != cB).sum()) + int((rB == cB).sum())
        b_kc = 2 * int((rC != cC).sum()) + int((rC == cC).sum())
        del rB, cB, rC, cC
        b_blk = int(((big.dim * big.nb_points.astype(np.int64)) ** 2).sum())
        b_bytes = algorithmic_bytes_iteration_launch(big, b_kb, b_kc, 8 * b_blk)
        b_gbs = b_bytes / (b_ms * 1e-3) / 1e9
        batched = {"workload": f"batch-64-systems-{big.nfib}-fibers-{big.npts}-points-{big.nlinks}-links",
                   "systems": batch.nsys, "fibers": big.nfib, "unknowns": big.size,
                   "steps_per_s": round(args.batched_steps / bdt, 3),
                   "system_steps_per_s": round(batch.nsys * args.batched_steps / bdt, 2),
                   "max_bicgstab_iters": float(np.mean(its)),
                   "untimed_setup": {"autotune": btune},
                   "iteration_launch": {"avg_ms": round(b_ms, 4), "bytes_per_launch": int(b_bytes),
                                        "achieved_gbs": round(b_gbs, 1), "frac_of_hbm_peak": round(b_gbs / 8000.0, 4)}}
        print(f"bench: batched {batched}", file=sys.stderr, flush=True)
        mb.close()

    asters = None
    if args.aster_steps > 0:
        print("bench: generating the aster scene", file=sys.stderr, flush=True)
        ast = scenes.config3(seed=3)
        ga = [torch.from_numpy(ast.gauss(s)).to("cuda:0") for s in range(2)]
        ma = pkg.Meca(3, 0)
        pkg.load_scene(ma, ast)
        its = []
        atune = None
        for s in range(2 + args.aster_steps):
            if s == 2:
                if args.autotune:
                    atune = ma.autotune(reps=6)      # untimed
                torch.cuda.synchronize()
                t0 = time.perf_counter()
            if s > 0:
                ma.prepare(None)
                pkg.assemble(ma, ast, upload_couples=False)
            rep = ma.solve(args.tol, tol_mode=pkg.TOL_RELATIVE, precondition=pkg.PRECOND_FP64,
                           gauss_device_ptr=ga[s % 2].data_ptr())
            # the motors walk under their loads between solves (Couple::stepAA)
            ma.step_couples(ast.motor_max_speed, ast.motor_stall_force)
            if s >= 2:
                its.append(rep["iterations"])
        torch.cuda.synchronize()
        adt = time.perf_counter() - t0
        _, _, cst = ma.get_couples()
        asters = {"workload": f"asters-10x200-fibers-{ast.nfib}-solids-{ast.nsol}-points-{ast.npts}-motors-{ast.nlinks}",
                  "unknowns": ast.size, "steps_per_s": round(args.aster_steps / adt, 3),
                  "ms_per_step": round(1e3 * adt / args.aster_steps, 3),
                  "mean_bicgstab_iters": float(np.mean(its)),
                  "iteration": "two launches per iteration (a tile per Solid)",
                  "untimed_setup": {"autotune": atune},
                  "motors": f"stepped on the device between solves (speed {ast.motor_max_speed} um/s, "
                            f"stall {ast.motor_stall_force} pN); {int(((cst & 3) == 3).sum())} of "
                            f"{cst.size} couples still bridging"}
        print(f"bench: asters {asters}", file=sys.stderr, flush=True)
        ma.close()

    grid = None
    if args.grid_places > 0:
        # FiberGrid query for many places at once (cymg_nearby_segments) on the
        # 1000-fiber system: places near the fibers, tryToAttach's range test;
        # host arrays in and out are included in the time
        mg = pkg.Meca(3, 0)
        pkg.load_scene(mg, sc)
        rng = np.random.default_rng(11)
        P = sc.points.reshape(-1, 3)
        W = P[rng.integers(0, P.shape[0], args.grid_places)] + rng.uniform(-0.1, 0.1, (args.grid_places, 3))
        mg.nearby_segments(W, 0.05, inclusive=True)          # untimed: buffers
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            counts, out = mg.nearby_segments(W, 0.05, inclusive=True)
        gdt = (time.perf_counter() - t0) / reps
        grid = {"workload": f"{args.grid_places} places near the fibers of the 1000-fiber system, range 0.05",
                "ms_per_query": round(1e3 * gdt, 3), "places_per_s": round(args.grid_places / gdt, 1),
                "segments_found": int(counts.sum()),
                "note": "one cymg_nearby_segments call, host arrays in and out included"}
        print(f"bench: grid {grid}", file=sys.stderr, flush=True)
        mg.close()

    cpu = None
    if args.cpu_steps > 0:
        from oracle import pyoracle
        print(f"bench: cpu baseline, {args.cpu_steps} oracle steps", file=sys.stderr, flush=True)
        orc = pyoracle.from_scene(sc)
        t0 = time.perf_counter()
        its = []
        for s in range(args.cpu_steps):
            rep = orc.solve(args.tol, sc.gauss(s), tol_mode=1)
            its.append(rep["iterations"])
            orc.prepare(orc.fiber_points())
            pyoracle.assemble_scene(orc, sc)
        cdt = time.perf_counter() - t0
        cpu = {"value": args.cpu_steps / cdt, "unit": "implicit steps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} steps of the same 1000-fiber workload, FP64 block-LU BCGSP, "
                         f"mean {np.mean(its):.1f} iterations, single thread"}

    traffic = args.traffic
    if traffic is None:
        # HBM bytes per iteration launch measured with rocprofv3 counters on this
        # workload (FETCH_SIZE and WRITE_SIZE passes, calibrated by a plain copy)
        tpath = os.path.join(ROOT, "profiles", "iteration_launch_traffic.json")
        wl = f"fibers-3d-{args.nfib}-crosslinks-{args.nlinks}-sphere10"
        if os.path.exists(tpath):
            with open(tpath) as fh:
                t = json.load(fh)
            if t.get("workload") == wl:
                traffic = t.get("traffic_bytes")

    best = results["fp64"]
    line = {
        "metric": "implicit steps/s @ rel.res 1e-6: plain FP64 GPU vs learned+mixed; BiCGStab iters",
        "value": round(best["steps_per_s"], 3),
        "unit": "implicit steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(best["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"fibers-3d-{args.nfib}-crosslinks-{args.nlinks}-sphere10",
                   "fibers": args.nfib, "points": sc.npts, "unknowns": sc.size, "links": args.nlinks,
                   "tolerance": f"|r|inf/|b|inf<{args.tol:g}"},
        "modes": {k: {"steps_per_s": round(v["steps_per_s"], 3), "ms_per_step": round(v["ms_per_step"], 3),
                      "mean_bicgstab_iters": round(v["mean_iters"], 2),
                      **({"untimed_setup": v["training"]} if v["training"] else {})} for k, v in results.items()},
        "learned_mixed_vs_plain_fp64": (round(results["learned-fp32"]["steps_per_s"] / best["steps_per_s"], 3)
                                        if "learned-fp32" in results else None),
        "roofline": {"bound": "hbm", "kernel": "k_tp/k_tq (one launch of the FP64 BiCGStab iteration)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_launch": int(it_bytes), "avg_launch_ms": round(it_ms, 5)},
        "kernels_ms": {"iteration_launch_fp64": it_ms, "iteration_launch_fp32": it32_ms,
                       "iteration_launch_fp32_gbs": it32_achieved and round(it32_achieved, 1),
                       "matvec": mv_ms, "matvec_gbs": round(mv_achieved, 1), "precond_apply": pc_ms,
                       "precond_build_fp64": build_ms, "precond_assemble_only": build_only_ms,
                       "precond_build_getrf_getri": build_lu_ms},
        "cpu_baseline": cpu,
        "batched": batched,
        "asters": asters,
        "nearby_segments": grid,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
