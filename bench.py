"""Benchmark of the MI355X implicit mechanics step (one GPU).

Workload: 3D, 1000 fibers with lengths uniform in [2, 15] um (segmentation 0.5,
5-31 points), rigidity 20, confined in a sphere of radius 10 (stiffness 200),
20000 crosslinkers (stiffness 100) bound at crossings within 0.05 um,
viscosity 0.02, time_step 0.005, kT 0.0042.  A step is one implicit step:
Meca::prepare + device assembly (confinement, couples) + Meca::solve at
relative residual |r|_inf/|b|_inf < 1e-6 + apply (update, forces, reshape) +
the export of the points and forces to the host (Meca::solve's getPoints /
getForces, meca.cc:826-835).  Fiber points and the per-step N(0,1) Brownian
draws are resident in HBM before the timed region.

Modes (each on its own context, from the scene's initial state, same
Brownian draws): fp64 (zero seed, FP64 preconditioner blocks: `value`),
fp64-previous (seeded by the previous step's solution, the initial guess the
learned one is compared with), fp32 (FP32 blocks), learned-fp64 and
learned-fp32 (seeded by the per-fiber network, trained on the device from
(input, solution) pairs of an untimed rollout of zero-seeded FP64 steps with
draws of their own; rollout and training times are reported beside the steps,
not inside them), and on request fp32-arith (FP32 blocks applied in FP32
arithmetic, the operand rounded and FP32 accumulation, which the four-launch
iteration solves).  Every mode's first step is compared with the oracle's FP64 step
solved to 1e-12 on the same input.

Beside it: the 64-system batched workload (FP64, FP32 and learned + FP32), the
aster workload (FP64 and FP32, first step against the FP64 oracle), the
FiberGrid query, and the CPU oracle on the same steps as the FP64 mode.

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
MODES = {   # name: (preconditioner, initial guess)
    "fp64": ("PRECOND_FP64", "GUESS_ZERO"),
    "fp64-previous": ("PRECOND_FP64", "GUESS_PREVIOUS"),
    "fp32": ("PRECOND_FP32", "GUESS_ZERO"),
    "learned-fp64": ("PRECOND_FP64", "GUESS_LEARNED"),
    "learned-fp32": ("PRECOND_FP32", "GUESS_LEARNED"),
    # FP32 blocks applied in FP32 arithmetic (operand rounded, FP32 accumulate):
    # a nonlinear M^-1, solved by the four-launch iteration
    "fp32-arith": ("PRECOND_FP32", "GUESS_ZERO"),
}
MODE_OPTIONS = {"fp32-arith": (("OPT_FP32_APPLY", 1),)}


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


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
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nfib", type=int, default=1000)
    ap.add_argument("--nlinks", type=int, default=20000)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--cpu-steps", type=int, default=24, help="oracle steps timed for cpu_baseline (0: skip)")
    ap.add_argument("--modes", default="fp64,fp64-previous,fp32,learned-fp64,learned-fp32",
                    help="comma-separated subset of " + ", ".join(MODES) + " (fp32-arith: about 2 s per step on "
                    "the 1000-fiber system, off by default)")
    ap.add_argument("--oracle-check", type=int, default=1,
                    help="compare each mode's first step with the oracle's FP64 step solved to 1e-12")
    ap.add_argument("--autotune", type=int, default=1,
                    help="run the library's autotuner (tile shape, block-inverse method) before the warmup")
    ap.add_argument("--guess-rollout", type=int, default=200,
                    help="untimed FP64 steps recorded as training pairs (one per fiber and step)")
    ap.add_argument("--guess-train-steps", type=int, default=4000)
    ap.add_argument("--guess-batch", type=int, default=512)
    ap.add_argument("--guess-lr", type=float, default=1e-3)
    ap.add_argument("--timed-launches", type=int, default=40,
                    help="iteration launches timed with HIP events for the roofline")
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per iteration launch measured with rocprofv3 counters (reported as is)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run a plain device-to-device copy of this many MiB (counter calibration)")
    ap.add_argument("--batched-steps", type=int, default=3,
                    help="timed steps of the 64-system batched workload per mode (0: skip)")
    ap.add_argument("--batched-modes", default="fp64,fp32,learned-fp32")
    ap.add_argument("--batched-rollout", type=int, default=4,
                    help="untimed FP64 steps of the batched workload recorded for its learned guess")
    ap.add_argument("--grid-places", type=int, default=50000,
                    help="places of the timed FiberGrid query (0: skip)")
    ap.add_argument("--aster-steps", type=int, default=5,
                    help="timed steps of the aster workload (10 Solids + 2000 fibers, 50000 motors; 0: skip)")
    ap.add_argument("--tile-threads", type=int, default=0,
                    help="with --autotune 0: threads per tile of the iteration launches (256 or 512; 0: default)")
    ap.add_argument("--tile-points", type=int, default=-1,
                    help="with --autotune 0: target points per tile (-1: default)")
    ap.add_argument("--only", default="all", choices=["all", "c2", "batched", "asters", "grid", "side"],
                    help="run one workload alone (for profiler runs); the JSON line is printed with 'all' or 'c2'")
    return ap.parse_args()


def cpu_description():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return model, os.cpu_count(), usable


def train_guess(pkg, m, sc, rollout, train_steps, batch, lr, tol, torch, assemble):
    """Record a rollout of zero-seeded FP64 steps (Brownian draws of their own,
    distinct from the timed steps'), then train the network on the device;
    returns the training record and the weights."""
    t0 = time.perf_counter()
    m.guess_record(rollout * sc.nfib)
    for s in range(rollout):
        if s > 0:
            m.prepare(None)
            assemble(m)
        m.solve(tol, sc.gauss(1000 + s), tol_mode=pkg.TOL_RELATIVE, precondition=pkg.PRECOND_FP64)
    m.guess_record(0)
    samples = m.guess_samples()
    t1 = time.perf_counter()
    m.guess_init(1)
    first = m.guess_eval(8192, seed=99)
    last_batch = m.guess_train(train_steps, batch=batch, lr=lr, seed=2)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    last = m.guess_eval(8192, seed=99)
    info = {"rollout_steps": rollout, "pairs": samples, "rollout_s": round(t1 - t0, 3),
            "train_steps": train_steps, "batch": batch, "epochs": round(train_steps * batch / max(samples, 1), 2),
            "train_s": round(t2 - t1, 3), "loss_first": first, "loss_last": last,
            "loss": "on the same 8192 recorded samples before and after training (cymg_guess_eval)",
            "last_minibatch_loss": last_batch}
    return info, m.guess_weights()


def errors(pts, sol, ref_pts, ref_sol):
    return {"position_rel": float(np.abs(pts - ref_pts).max() / np.abs(ref_pts).max()),
            "position_rel_2norm": float(np.linalg.norm(pts - ref_pts) / np.linalg.norm(ref_pts)),
            "displacement_rel": float(np.abs(sol - ref_sol).max() / np.abs(ref_sol).max())}


def fixed_tiles(pkg, m, args):
    """The tile shape given on the command line (profiler runs without the autotuner)."""
    if args.tile_threads:
        m.set_option(pkg.OPT_TILE_THREADS, args.tile_threads)
    if args.tile_points >= 0:
        m.set_option(pkg.OPT_TILE_POINTS, args.tile_points)


def copy_calibration(args, torch):
    """A plain device-side copy kernel of known size (an elementwise kernel that
    reads and writes every element once: a memcpy may run on the DMA engine,
    which the counters do not see), for scaling the HBM byte counters
    (FETCH_SIZE, WRITE_SIZE) of a profiler run."""
    if args.copy_calibration_mb <= 0:
        return
    nb = args.copy_calibration_mb << 20
    src = torch.ones(nb // 8, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    del src, dst


def run_c2_mode(pkg, sc, mode, args, gauss_dev, torch, weights, ref):
    precond = getattr(pkg, MODES[mode][0])
    guess = getattr(pkg, MODES[mode][1])
    m = pkg.Meca(sc.dim, 0)
    pkg.load_scene(m, sc)
    fixed_tiles(pkg, m, args)
    for opt, val in MODE_OPTIONS.get(mode, ()):
        m.set_option(getattr(pkg, opt), val)
    tuned = None
    if args.autotune:
        # untimed: the autotuner picks the tile shape and block-inverse method on this system
        m.setup(gauss_device_ptr=gauss_dev[0].data_ptr())
        m.compute_preconditioner(precond)
        tuned = m.autotune(reps=20)
        m.prepare(sc.points)
        pkg.assemble(m, sc, upload_couples=False)
    if guess == pkg.GUESS_LEARNED:
        m.set_guess_weights(weights)
    pts_host = np.empty(sc.size)
    for_host = np.empty(sc.size)
    iters = []
    used = []
    err = err_typ = None

    def one(step):
        if step > 0:
            m.prepare(None)
            pkg.assemble(m, sc, upload_couples=False)
        rep = m.solve(args.tol, tol_mode=pkg.TOL_RELATIVE, precondition=precond, guess=guess,
                      gauss_device_ptr=gauss_dev[step % len(gauss_dev)].data_ptr())
        # the learned guess is replaced by zero when its residual is not below |b|
        if step > 0 and rep["guess_used"] not in (guess, pkg.GUESS_ZERO):
            raise RuntimeError(f"{mode}: initial guess {rep['guess_used']} reported")
        used.append(rep["guess_used"] == guess)
        # Meca::solve's export: getPoints and getForces of every Mecable
        m.export_points(pts_host)
        m.export_vector(pkg.VEC_FOR, for_host)
        iters.append(rep["iterations"])

    for s in range(args.warmup):
        if s == args.warmup - 1 and args.oracle_check and s > 0:
            before = pts_host.copy()          # the points this step starts from
        one(s)
        if s == 0 and ref is not None:
            err = errors(pts_host, m.vector(pkg.VEC_SOL), ref[0], ref[1])
        if s == args.warmup - 1 and args.oracle_check and s > 0:
            # a step of the running simulation (after the relaxation of the
            # generated configuration): the oracle solves it from the same points
            from oracle import pyoracle
            orc = pyoracle.OracleMeca(sc.dim, sc.params())
            pyoracle.register_scene(orc, sc)
            orc.prepare(before)
            pyoracle.assemble_scene(orc, sc)
            orc.solve(1e-12, sc.gauss(s % len(gauss_dev)), tol_mode=1)
            err_typ = dict(errors(pts_host, m.vector(pkg.VEC_SOL), orc.fiber_points(), orc.vector(pyoracle.SOL)),
                           step=s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        one(args.warmup + s)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    timed = iters[args.warmup:]
    return m, {"seconds": dt, "steps_per_s": args.steps / dt, "ms_per_step": 1e3 * dt / args.steps,
               "mean_iters": float(np.mean(timed)), "iters": iters, "error": err, "error_step": err_typ,
               "guess_kept": float(np.mean(used[args.warmup:])), "autotune": tuned}


def c2(args, pkg, scenes, torch):
    log(f"generating scene nfib={args.nfib} nlinks={args.nlinks}")
    sc = scenes.config2(seed=2, nfib=args.nfib, nlinks=args.nlinks)
    gauss_dev = [torch.from_numpy(sc.gauss(s)).to("cuda:0") for s in range(8)]
    torch.cuda.synchronize()
    modes = [x for x in args.modes.split(",") if x]
    for x in modes:
        if x not in MODES:
            raise SystemExit(f"unknown mode {x}")
    if "fp64" not in modes:
        modes.insert(0, "fp64")
    from oracle import pyoracle

    ref = None
    if args.oracle_check:
        t0 = time.perf_counter()
        orc = pyoracle.from_scene(sc)
        orc.solve(1e-12, sc.gauss(0), tol_mode=1)
        ref = (orc.fiber_points(), orc.vector(pyoracle.SOL))
        log(f"oracle FP64 step at 1e-12: {time.perf_counter() - t0:.1f} s")
        del orc

    training = weights = None
    if any(MODES[x][1] == "GUESS_LEARNED" for x in modes):
        mt = pkg.Meca(sc.dim, 0)
        pkg.load_scene(mt, sc)
        training, weights = train_guess(pkg, mt, sc, args.guess_rollout, args.guess_train_steps, args.guess_batch,
                                        args.guess_lr, args.tol, torch,
                                        lambda mm: pkg.assemble(mm, sc, upload_couples=False))
        mt.close()
        log(f"learned guess: {training}")

    results = {}
    keep = None
    for mode in modes:
        log(f"mode {mode} warmup={args.warmup} steps={args.steps}")
        m, r = run_c2_mode(pkg, sc, mode, args, gauss_dev, torch, weights, ref)
        results[mode] = r
        log(f"{mode}: {r['steps_per_s']:.2f} steps/s, mean iterations {r['mean_iters']:.1f}, error {r['error']}")
        if mode == "fp64":
            keep = m
        else:
            m.close()

    # dominant-kernel timing on the final state of the FP64 run
    m = keep
    m.setup(gauss_device_ptr=gauss_dev[0].data_ptr())
    mv_ms = m.time_kernel(0, reps=200)
    m.compute_preconditioner(pkg.PRECOND_FP64)
    pc_ms = m.time_kernel(1, reps=200)
    build_ms = m.time_kernel(2, reps=5)
    build_only_ms = m.time_kernel(4, reps=5)
    build_lu_ms = m.time_kernel(5, reps=5)
    rB, cB, _ = m.matrix_entries(0)
    rC, cC, _ = m.matrix_entries(1)
    kb_full = 2 * int((rB != cB).sum()) + int((rB == cB).sum())
    kc_full = 2 * int((rC != cC).sum()) + int((rC == cC).sum())
    mv_bytes = algorithmic_bytes_matvec(sc, kb_full, kc_full)
    nblk = int(((sc.dim * sc.nb_points.astype(np.int64)) ** 2).sum())
    m.compute_preconditioner(pkg.PRECOND_FP64)
    it_ms = m.time_kernel(6, reps=args.timed_launches)
    it_bytes = algorithmic_bytes_iteration_launch(sc, kb_full, kc_full, 8 * nblk)
    m.compute_preconditioner(pkg.PRECOND_FP32)
    it32_ms = m.time_kernel(6, reps=args.timed_launches)
    it32_bytes = algorithmic_bytes_iteration_launch(sc, kb_full, kc_full, 4 * nblk)
    m.close()
    kern = {"iteration_launch_fp64_ms": it_ms, "iteration_launch_fp64_bytes": int(it_bytes),
            "iteration_launch_fp32_ms": it32_ms, "iteration_launch_fp32_bytes": int(it32_bytes),
            "iteration_launch_fp32_gbs": round(it32_bytes / (it32_ms * 1e-3) / 1e9, 1),
            "matvec_ms": mv_ms, "matvec_gbs": round(mv_bytes / (mv_ms * 1e-3) / 1e9, 1), "precond_apply_ms": pc_ms,
            "precond_build_fp64_ms": build_ms, "precond_assemble_only_ms": build_only_ms,
            "precond_build_getrf_getri_ms": build_lu_ms}
    return sc, results, training, kern, (it_ms, it_bytes)


def batched(args, pkg, scenes, torch):
    log("generating 64 independent systems")
    batch = scenes.config5(seed=5)
    big = batch.scene
    gb = [torch.from_numpy(big.gauss(s)).to("cuda:0") for s in range(2)]
    torch.cuda.synchronize()
    out = {"workload": f"batch-64-systems-{big.nfib}-fibers-{big.npts}-points-{big.nlinks}-links",
           "systems": batch.nsys, "fibers": big.nfib, "unknowns": big.size, "modes": {}}
    weights = None
    for mode in [x for x in args.batched_modes.split(",") if x]:
        precond = getattr(pkg, MODES[mode][0])
        guess = getattr(pkg, MODES[mode][1])
        mb = pkg.Meca(3, 0)
        pkg.load_batch(mb, batch)
        fixed_tiles(pkg, mb, args)
        rec = {}
        if guess == pkg.GUESS_LEARNED:
            if weights is None:
                info, weights = train_guess(pkg, mb, big, args.batched_rollout, args.guess_train_steps,
                                            args.guess_batch, args.guess_lr, args.tol, torch,
                                            lambda mm: pkg.assemble(mm, big, True, False))
                rec["untimed_training"] = info
                mb.prepare(big.points)
                pkg.assemble(mb, big, True, False)
            mb.set_guess_weights(weights)
        if args.autotune:
            # untimed: tile shape and block-inverse method for this workload
            mb.setup(gauss_device_ptr=gb[0].data_ptr())
            mb.compute_preconditioner(precond)
            rec["untimed_autotune"] = mb.autotune(reps=6)
            mb.prepare(big.points)
            pkg.assemble(mb, big, True, False)
        its = []
        t0 = None
        for s in range(1 + args.batched_steps):
            if s == 1:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
            if s > 0:
                mb.prepare(None)
                pkg.assemble(mb, big, True, False)
            rep = mb.solve(args.tol, tol_mode=pkg.TOL_RELATIVE, precondition=precond, guess=guess,
                           gauss_device_ptr=gb[s % 2].data_ptr())
            if s > 0:
                its.append(rep["iterations"])
        torch.cuda.synchronize()
        if t0 is not None:
            bdt = time.perf_counter() - t0
            rec.update({"steps_per_s": round(args.batched_steps / bdt, 3),
                        "system_steps_per_s": round(batch.nsys * args.batched_steps / bdt, 2),
                        "max_bicgstab_iters": float(np.mean(its))})
        if guess == pkg.GUESS_ZERO:
            # full-work iteration launch over all 64 systems (every tile iterating)
            mb.setup(gauss_device_ptr=gb[0].data_ptr())
            mb.compute_preconditioner(precond)
            b_ms = mb.time_kernel(6, reps=10)
            rB, cB, _ = mb.matrix_entries(0)
            rC, cC, _ = mb.matrix_entries(1)
            b_kb = 2 * int((rB != cB).sum()) + int((rB == cB).sum())
            b_kc = 2 * int((rC != cC).sum()) + int((rC == cC).sum())
            del rB, cB, rC, cC
            b_blk = int(((big.dim * big.nb_points.astype(np.int64)) ** 2).sum())
            b_bytes = algorithmic_bytes_iteration_launch(big, b_kb, b_kc, (8 if precond == pkg.PRECOND_FP64 else 4)
                                                         * b_blk)
            b_gbs = b_bytes / (b_ms * 1e-3) / 1e9
            rec["iteration_launch"] = {"avg_ms": round(b_ms, 4), "bytes_per_launch": int(b_bytes),
                                       "achieved_gbs": round(b_gbs, 1), "frac_of_hbm_peak": round(b_gbs / 8000.0, 4)}
        out["modes"][mode] = rec
        log(f"batched {mode}: {rec}")
        mb.close()
    return out


def asters_workload(args, pkg, scenes, torch):
    log("generating the aster scene")
    ast = scenes.config3(seed=3)
    ga = [torch.from_numpy(ast.gauss(s)).to("cuda:0") for s in range(2)]
    torch.cuda.synchronize()
    ref = None
    if args.oracle_check:
        from oracle import pyoracle
        t0 = time.perf_counter()
        orc = pyoracle.from_scene(ast)
        orc.solve(1e-12, ast.gauss(0), tol_mode=1)
        ref = (orc.fiber_points(), orc.vector(pyoracle.SOL))
        log(f"oracle FP64 aster step at 1e-12: {time.perf_counter() - t0:.1f} s")
        del orc
    out = {"workload": f"asters-10x200-fibers-{ast.nfib}-solids-{ast.nsol}-points-{ast.npts}-motors-{ast.nlinks}",
           "unknowns": ast.size, "iteration": "two launches per iteration (a tile per Solid)",
           "motors": f"stepped on the device between solves (speed {ast.motor_max_speed} um/s, "
                     f"stall {ast.motor_stall_force} pN)", "modes": {}}
    for mode in ("fp64", "fp32"):
        precond = getattr(pkg, MODES[mode][0])
        ma = pkg.Meca(3, 0)
        pkg.load_scene(ma, ast)
        fixed_tiles(pkg, ma, args)
        its = []
        rec = {}
        t0 = None
        for s in range(2 + args.aster_steps):
            if s == 2:
                if args.autotune:
                    rec["untimed_autotune"] = ma.autotune(reps=6)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
            if s > 0:
                ma.prepare(None)
                pkg.assemble(ma, ast, upload_couples=False)
            rep = ma.solve(args.tol, tol_mode=pkg.TOL_RELATIVE, precondition=precond,
                           gauss_device_ptr=ga[s % 2].data_ptr())
            if s == 0 and ref is not None:
                rec["first_step_error_vs_fp64_oracle"] = errors(ma.fiber_points(), ma.vector(pkg.VEC_SOL), *ref)
            # the motors walk under their loads between solves (Couple::stepAA)
            ma.step_couples(ast.motor_max_speed, ast.motor_stall_force)
            if s >= 2:
                its.append(rep["iterations"])
        torch.cuda.synchronize()
        adt = time.perf_counter() - t0
        _, _, cst = ma.get_couples()
        rec.update({"steps_per_s": round(args.aster_steps / adt, 3), "ms_per_step": round(1e3 * adt / args.aster_steps, 3),
                    "mean_bicgstab_iters": float(np.mean(its)),
                    "couples_bridging": f"{int(((cst & 3) == 3).sum())} of {cst.size}"})
        out["modes"][mode] = rec
        log(f"asters {mode}: {rec}")
        ma.close()
    return out


def grid_query(args, pkg, scenes, sc=None):
    # FiberGrid query for many places at once on the 1000-fiber system:
    # places near the fibers, tryToAttach's range test; host arrays in and out
    # are included in the time
    if sc is None:
        sc = scenes.config2(seed=2, nfib=args.nfib, nlinks=args.nlinks)
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
    t0 = time.perf_counter()
    for _ in range(reps):
        counts2, _ = mg.nearby_segments(W, 0.05, inclusive=True, noalloc=True)
    ndt = (time.perf_counter() - t0) / reps
    mg.close()
    res = {"workload": f"{args.grid_places} places near the fibers of the 1000-fiber system, range 0.05",
           "ms_per_query": round(1e3 * gdt, 3), "places_per_s": round(args.grid_places / gdt, 1),
           "ms_per_query_noalloc": round(1e3 * ndt, 3), "segments_found": int(counts.sum()),
           "noalloc_same_counts": bool(np.array_equal(counts, counts2)),
           "note": "one cymg_nearby_segments call, host arrays in and out included"}
    log(f"grid {res}")
    return res


def steric_and_frames(args, pkg, scenes, sc=None):
    # the steric grid (Simul::setStericInteractions) and trajectory frames
    # (Simul::writeObjects) on the 1000-fiber system; host timings of one call
    # each, the frame's host bytes included
    if sc is None:
        sc = scenes.config2(seed=2, nfib=args.nfib, nlinks=args.nlinks)
    m = pkg.Meca(3, 0)
    pkg.load_scene(m, sc)
    n = sc.nfib
    m.set_steric([10.0] * 3, 0.0, 100.0, 0.0, np.ones(n, np.int32), np.full(n, 0.025), np.full(n, 0.025))
    st = []
    for rep in range(4):
        m.prepare(None)
        pkg.assemble(m, sc)
        t0 = time.perf_counter()
        m.steric_interactions()
        st.append(time.perf_counter() - t0)
        m.setup(sc.gauss(0))
    m.set_frame_records([(k + 1, 0, 0, 0, 0.5, 0.0) for k in range(n)])
    fr = {}
    for name, rb in (("float", pkg.FRAME_FLOAT), ("double", pkg.FRAME_DOUBLE)):
        data = m.write_frame(0.0, "bench", rb)       # untimed: layout
        reps = 10
        t0 = time.perf_counter()
        for _ in range(reps):
            data = m.write_frame(0.0, "bench", rb)
        dt = (time.perf_counter() - t0) / reps
        fr[name] = {"ms_per_frame": round(1e3 * dt, 3), "bytes": len(data)}
    m.close()
    res = {"steric": {"workload": "every fiber of the 1000-fiber system steric, radius 0.025, range 0.025",
                      "ms_per_call": round(1e3 * float(np.median(st[1:])), 3),
                      "note": "cymg_steric_interactions: items, grid sort, count walk, scans, emit walk"},
           "trajectory_frame": dict(fr, workload="the 1000-fiber system, fiber section",
                                    note="the Python write_frame: a size query, then cymg_write_frame into host memory")}
    log(f"steric/frames {res}")
    return res


def cpu_baseline(args, sc, gpu_iters):
    from oracle import pyoracle
    log(f"cpu baseline, {args.cpu_steps} oracle steps")
    orc = pyoracle.from_scene(sc)
    its = []
    t0 = time.perf_counter()
    for s in range(args.cpu_steps):
        if s > 0:
            orc.prepare(orc.fiber_points())
            pyoracle.assemble_scene(orc, sc)
        rep = orc.solve(args.tol, sc.gauss(s % 8), tol_mode=1)
        its.append(rep["iterations"])
    cdt = time.perf_counter() - t0
    model, nproc, usable = cpu_description()
    k = min(len(its), len(gpu_iters))
    return ({"value": args.cpu_steps / cdt, "unit": "implicit steps/s", "cores": 1, "kind": "port",
             "cpu_model": model, "nproc": nproc, "usable_cpus": usable,
             "sample": f"steps 0-{args.cpu_steps - 1} of the same 1000-fiber workload (same draws as the FP64 GPU "
                       f"mode), FP64 block-LU BCGSP, |r|inf/|b|inf<{args.tol:g}, mean {np.mean(its):.1f} iterations, "
                       f"single thread; assembly of the 20000 couples in one C call per step"},
            {"steps": k, "gpu_fp64": [int(x) for x in gpu_iters[:k]], "oracle_fp64": [int(x) for x in its[:k]]})


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
    copy_calibration(args, torch)
    if args.only == "batched":
        batched(args, pkg, scenes, torch)
        return
    if args.only == "asters":
        asters_workload(args, pkg, scenes, torch)
        return
    if args.only == "grid":
        grid_query(args, pkg, scenes)
        return
    if args.only == "side":
        steric_and_frames(args, pkg, scenes)
        return

    sc, results, training, kern, (it_ms, it_bytes) = c2(args, pkg, scenes, torch)
    achieved = it_bytes / (it_ms * 1e-3) / 1e9
    full = args.only == "all"
    bat = batched(args, pkg, scenes, torch) if full and args.batched_steps > 0 else None
    ast = asters_workload(args, pkg, scenes, torch) if full and args.aster_steps > 0 else None
    grid = grid_query(args, pkg, scenes, sc) if full and args.grid_places > 0 else None
    side = steric_and_frames(args, pkg, scenes, sc) if full else None
    cpu = same = None
    if args.cpu_steps > 0:
        cpu, same = cpu_baseline(args, sc, results["fp64"]["iters"])

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
                      "timed_steps_seeded_by_the_mode_guess": v["guess_kept"],
                      "first_step_error_vs_fp64_oracle": v["error"],
                      "running_step_error_vs_fp64_oracle": v["error_step"],
                      **({"untimed_autotune": v["autotune"]} if v["autotune"] else {})}
                  for k, v in results.items()},
        "error_reference": "oracle FP64 step solved to |r|inf/|b|inf<1e-12 on the same input: the first step "
                           "(a relaxation of the generated configuration, points outside the sphere pulled in) "
                           "and the last warmup step (from the mode's own points); position_rel = "
                           "max|dpos|/max|pos|, position_rel_2norm = |dpos|_2/|pos|_2, displacement_rel = "
                           "max|dx|/max|x| of the step's solution",
        "learned_guess_training": training,
        "learned_mixed_vs_plain_fp64": (round(results["learned-fp32"]["steps_per_s"] / best["steps_per_s"], 3)
                                        if "learned-fp32" in results else None),
        "roofline": {"bound": "hbm", "kernel": "k_tp/k_tq (one launch of the FP64 BiCGStab iteration)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_launch": int(it_bytes), "avg_launch_ms": round(it_ms, 5)},
        "kernels": kern,
        "cpu_baseline": cpu,
        "iterations_same_steps": same,
        "batched": bat,
        "asters": ast,
        "nearby_segments": grid,
        "steric": side and side["steric"],
        "trajectory_frame": side and side["trajectory_frame"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
