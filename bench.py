"""Benchmark: 42-RHS solve time to rel-res 1e-8 on the 0.75 mm frame grid
(256 x 256 x 48 hex8, 9,709,203 DOF), mixed-precision solver vs plain FP64 GMG-PCG.

One step = update(moduli) + Galerkin hierarchy set-up + 42-RHS PCG solve to a
true FP64 relative residual < 1e-8 per column, with the density field, fixture
and the 42 unit-wrench right-hand sides already resident in HBM.

The mixed solver is one block PCG recurrence over the 42 columns in FP32 with reliable
FP64 updates, around an FP32 V-cycle whose level-1 correction is two flexible-CG steps
preconditioned by the level-1 V-cycle, whose own level-2 correction is two flexible-CG
steps again (coarse_krylov 2, krylov_levels 2), with Chebyshev degree 3 on every level
(growth 1.0) and reliable updates after a 3e-3 drop of the FP32 residual
(inner_tolerance); the FP64 solve is the reference's scheme (FP64 V-cycle, degree
3 * 2^level, one PCG recurrence per column), timed with the same warm-up and steps.  Two
more legs separate the gains: the mixed solver with one recurrence per column, and FP64
with the mixed leg's per-column algorithm.

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
[--config frame|frame_small|cantilever] [--no-cpu-baseline] [--no-fp64]
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

CONFIGS = {
    "frame": {"shape": (256, 256, 48), "rhs": 42, "levels": 5},
    "frame_small": {"shape": (96, 96, 16), "rhs": 42, "levels": 4},
    "cantilever": {"shape": (48, 24, 24), "rhs": 1, "levels": 3},
    # 64 seeded uneven frame variants, nx, ny in [48, 160], nz in [12, 32], 6-42 RHS each
    "variants": {"count": 64},
    # the callers either side of the solve: covariance measures and robust projection on the 256 x 256 x 48
    # frame, LOBPCG modes on a 91 x 170 x 48 frame (the 0.75 mm half-domain size), a floating 3-2-1 support
    # on the 96 x 96 x 16 frame
    "rows": {},
}
HBM_PEAK_GBS = 8000.0
TRAFFIC = {1: ROOT / "profiles" / "traffic_k_fine_apply_fp32_frame_256x256x48_42rhs.txt",
           2: ROOT / "profiles" / "traffic_k_fine_mfma_frame_256x256x48_42rhs.txt"}
KERNELS = {1: "k_fine_apply<float, MODE_APPLY> (FP32 vector units)",
           2: "k_fine_mfma<MODE_APPLY> (split-bf16 matrix cores, FP32 vectors)"}


def measured_traffic(config, precision):
    """HBM bytes per launch of a fine operator kernel from its committed rocprofv3 --pmc summary (frame only)."""
    path = TRAFFIC[precision]
    if config != "frame" or not path.exists():
        return None
    vals = {}
    for line in path.read_text().splitlines():
        if line and not line.startswith("#"):
            key, value = line.split()
            vals[key] = float(value)
    return vals["fetch_bytes"] + vals["write_bytes"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="frame", choices=sorted(CONFIGS))
    ap.add_argument("--variants", type=int, default=None, help="variants config: number of variants (default 64)")
    ap.add_argument("--streams", type=int, default=4, help="variants config: concurrent HIP streams")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--scheme", default="refine", choices=["refine", "flexible"],
                    help="mixed solver: FP32 PCG inside FP64 iterative refinement, or FP64 PCG with an FP32 V-cycle")
    ap.add_argument("--inner-tolerance", type=float, default=3e-3,
                    help="refine scheme: FP32 residual drop between reliable FP64 updates")
    ap.add_argument("--apply-reps", type=int, default=20)
    ap.add_argument("--coarse-krylov", type=int, default=2, choices=[0, 1, 2],
                    help="mixed solver: flexible-CG steps of the level-1 solve inside the fine V-cycle (0: V-cycle)")
    ap.add_argument("--krylov-levels", type=int, default=2,
                    help="mixed solver: levels whose coarse-grid correction is the Krylov solve (recursive K-cycle)")
    ap.add_argument("--degree", type=int, default=3, help="mixed solver: Chebyshev degree on the fine level")
    ap.add_argument("--block", dest="block", action="store_true", default=True,
                    help="mixed solver: one block PCG recurrence over the columns (default)")
    ap.add_argument("--per-column", dest="block", action="store_false",
                    help="mixed solver: one PCG recurrence per column")
    ap.add_argument("--chebyshev-kind", type=int, default=1, choices=[1, 4],
                    help="mixed solver: Chebyshev smoother polynomial, 1 the reference's first kind, 4 the fourth kind")
    ap.add_argument("--chebyshev-ratio", type=float, default=20.0,
                    help="mixed solver: the Chebyshev interval is [lambda_max / ratio, lambda_max] (the reference's 20)")
    ap.add_argument("--growth", type=float, default=1.0,
                    help="mixed solver: Chebyshev degree growth per level (degree * growth^level)")
    ap.add_argument("--fine-smoother", default="vector", choices=["matrix", "vector"],
                    help="fine level of the fp32 V-cycle on the matrix cores (split bf16) or the FP32 vector units")
    ap.add_argument("--fp64-only", action="store_true", help="only time the plain FP64 solve (for kernel traces)")
    ap.add_argument("--apply-only", action="store_true",
                    help="build the hierarchy once and only time the fine operator (for counter profiles)")
    ap.add_argument("--learned", default=None,
                    help="mixed solver: trained coarse-correction weights (.npz under the repository) to attach")
    return ap.parse_args()


def _cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _cpu_run(name, threads, wall_cap_s, memory_cap_gb):
    """One child process of oracle.cpu_baseline under a wall-clock and an address-space cap; DNF if it does not
    finish within them."""
    import resource
    import subprocess
    env = {**os.environ, "OMP_NUM_THREADS": str(threads), "OPENBLAS_NUM_THREADS": str(threads),
           "MKL_NUM_THREADS": str(threads)}
    limit = int(memory_cap_gb * 2 ** 30)

    def cap():
        resource.setrlimit(resource.RLIMIT_AS, (limit, limit))

    caps = {"wall_cap_s": wall_cap_s, "address_space_cap_gb": memory_cap_gb}
    try:
        done = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", name, str(threads)], cwd=ROOT, env=env,
                              capture_output=True, text=True, timeout=wall_cap_s, preexec_fn=cap)
    except subprocess.TimeoutExpired:
        return {"workload": name, "threads": threads, "status": "DNF (wall-clock cap)", **caps}
    if done.returncode != 0:
        tail = done.stderr.strip().splitlines()[-1:] if done.stderr else []
        return {"workload": name, "threads": threads, "status": f"DNF (exit {done.returncode})", "error": tail, **caps}
    out = json.loads(done.stdout.strip().splitlines()[-1])
    print(f"# cpu {name} threads {threads}: {out['seconds']:.1f} s", file=sys.stderr, flush=True)
    return {**out, "status": "ok", **caps}


def cpu_baseline():
    """The reference's CPU path (SuperLU with MMD_AT_PLUS_A, assembly to sensitivity; oracle.cpu_baseline) on the
    host cores: the 48 x 24 x 24 cantilever on one thread (the value) and on the host's share of cores, two
    small grids of the frame family with 42 RHS for the time scaling, the 96 x 96 x 16 frame under a cap, and the
    256 x 256 x 48 frame extrapolated from the fitted scaling (SuperLU fill grows about N^2: not attempted)."""
    import numpy as np
    affinity = len(os.sched_getaffinity(0))
    cores = max(1, min(affinity, 16))
    runs = [_cpu_run("cantilever", 1, 300, 48), _cpu_run("cantilever", cores, 300, 48),
            _cpu_run("frame:32,32,6", 1, 120, 48), _cpu_run("frame:48,48,8", 1, 180, 48),
            _cpu_run("frame:96,96,16", 1, 60, 48)]
    small, large = runs[2], runs[3]
    extrapolation = None
    if small.get("status") == "ok" and large.get("status") == "ok":
        alpha = np.log(large["seconds"] / small["seconds"]) / np.log(large["dof"] / small["dof"])
        beta = np.log(large["peak_rss_gb"] / small["peak_rss_gb"]) / np.log(large["dof"] / small["dof"])
        extrapolation = {"fit": "t = t(48x48x8) (N / N(48x48x8))^alpha over the frame family, 42 RHS, 1 thread",
                         "alpha": float(alpha), "rss_exponent": float(beta)}
        for label, dof in (("frame 96x96x16", 479859), ("frame 256x256x48", 9709203)):
            extrapolation[label] = {"dof": dof, "seconds": float(large["seconds"] * (dof / large["dof"]) ** alpha),
                                    "peak_rss_gb": float(large["peak_rss_gb"] * (dof / large["dof"]) ** beta)}
    runs.append({"workload": "frame:256,256,48", "threads": 1, "status": "DNF (not attempted: the extrapolated time "
                 "exceeds any cap of a benchmark run)"})
    first = runs[0]
    return {"value": first.get("seconds"), "unit": "s", "cores": 1, "kind": "port",
            "sample": "48x24x24 cantilever (91,875 DOF, 1 RHS): assembly, SuperLU factor, solve, compliance and "
                      "sensitivity on one thread (the reference's documented thread setting); the other runs are "
                      "listed in runs",
            "host": {"cpu": _cpu_model(), "nproc": os.cpu_count(), "affinity": affinity, "threads_all_cores": cores},
            "runs": runs, "extrapolation": extrapolation}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: --gpus above 1 is not supported (sharding across GPUs is designed, not built)"}))
        sys.exit(2)
    import numpy as np
    import torch
    from ctypes import byref, c_double
    from deep_frame_amd import _native, element, problems
    from deep_frame_amd.multigrid import GeometricMultigrid

    cfg = CONFIGS[args.config]
    if args.config == "variants":
        return run_variants(args, cfg)
    if args.config == "rows":
        return run_rows(args)
    prob = problems.frame(cfg["shape"], seed=0, rhs=cfg["rhs"]) if cfg["rhs"] > 1 else problems.cantilever(seed=0)
    ke = element.stiffness((1, 1, 1), element.isotropic(0.3))
    mod, dmod = problems.moduli(prob["density"])
    dev = torch.device("cuda:0")
    moduli = torch.as_tensor(mod, device=dev)
    rhs = torch.as_tensor(np.ascontiguousarray(prob["forces"].T), device=dev)
    fixed = prob["fixed"]
    ndof = rhs.shape[1]
    nel = moduli.numel()
    print(f"# config {args.config} shape {cfg['shape']} ndof {ndof} rhs {rhs.shape[0]}", file=sys.stderr, flush=True)

    def run(settings, steps, warmup, label, net=None):
        mg = GeometricMultigrid(cfg["shape"], (1, 1, 1), ke, {**settings, "levels": cfg["levels"]})
        if net is not None:
            net.attach(mg)
        times, reports = [], None
        for it in range(warmup + steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            mg.update(moduli)
            x, reports = mg.solve_device(("bench", it), fixed, rhs)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            print(f"# {label} step {it} {'warmup' if it < warmup else 'timed'} {dt:.3f} s, max iterations "
                  f"{max(r[0] for r in reports)}, max rel-res {max(r[1] for r in reports):.2e}", file=sys.stderr,
                  flush=True)
            if not all(r[2] for r in reports):
                raise RuntimeError(f"{label} solve did not converge: {reports}")
            if it >= warmup:
                times.append(dt)
        return mg, times, reports, x

    if args.apply_only:
        mg = GeometricMultigrid(cfg["shape"], (1, 1, 1), ke, {"precision": "float32", "levels": cfg["levels"]})
        mg.update(moduli)
        slot = mg.hierarchy("apply", fixed)["slot"]
        mg.tune_fine("apply", rhs.shape[0])
        for level in (0, 1):
            for precision in ((2, 1, 0) if level == 0 else (1, 0)):
                ms = c_double()
                mg._stream()
                _native.call("dfmi_time_apply", mg.handle, slot, level, precision, rhs.shape[0], args.apply_reps,
                             byref(ms))
                size = 8 if precision == 0 else 4
                nodes = int(np.prod([n // 2 ** level + 1 for n in mg.padded]))
                if level == 0:
                    alg = 2.0 * ndof * rhs.shape[0] * size + nel * size
                else:  # vectors in and out plus the 243-entry stencil per node
                    alg = 2.0 * 3 * nodes * rhs.shape[0] * size + 243.0 * nodes * size
                name = {2: "k_fine_mfma", 1: "k_fine_apply", 0: "k_fine_apply"}[precision] if level == 0 \
                    else "k_stencil_apply"
                print(json.dumps({"kernel": name, "level": level,
                                  "precision": {0: "fp64", 1: "fp32", 2: "bf16x3"}[precision], "rhs": rhs.shape[0],
                                  "ms_per_launch": ms.value, "algorithmic_GBs": alg / (ms.value * 1e-3) / 1e9}))
        mg.close()
        # plain copy of a known size: calibrates the FETCH_SIZE / WRITE_SIZE counters
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": src.numel() * 4}))
        return
    if args.fp64_only:
        _, t64, rep64, _ = run({"precision": "float64"}, args.steps, args.warmup, "fp64")
        print(json.dumps({"fp64_s": float(np.mean(t64)), "iterations": max(r[0] for r in rep64)}))
        return
    mixed = {"precision": "float32", "refine": args.scheme == "refine", "fine_smoother": args.fine_smoother,
             "coarse_krylov": args.coarse_krylov, "krylov_levels": args.krylov_levels, "growth": args.growth,
             "chebyshev_degree": args.degree, "block_krylov": args.block, "chebyshev_kind": args.chebyshev_kind,
             "chebyshev_ratio": args.chebyshev_ratio}
    if args.inner_tolerance:
        mixed["inner_tolerance"] = args.inner_tolerance
    learned = None
    if args.learned:
        from deep_frame_amd.learned import CoarseCorrection
        learned = CoarseCorrection.load(ROOT / args.learned)
    mg, times, reports, x = run(mixed, args.steps, args.warmup, f"mixed-{args.scheme}", learned)
    t_mixed = float(np.mean(times))
    # dominant K.u kernel of the mixed solve: the fine-level smoother / residual operator of the V-cycle
    # (matrix cores unless --fine-smoother vector), on all RHS columns, HIP events on the library stream;
    # the FP32 vector kernel (the Krylov operator) is timed beside it
    key = ("bench", args.warmup + args.steps - 1)
    slot = mg.hierarchies[key]["slot"]
    k = rhs.shape[0]
    tile = mg.tune_fine(key, k) if args.fine_smoother == "matrix" and cfg["shape"][2] <= 64 else None
    alg_bytes = 2.0 * ndof * k * 4 + nel * 4
    timings = {}
    for precision in ((2, 1) if tile else (1,)):
        ms = c_double()
        mg._stream()
        _native.call("dfmi_time_apply", mg.handle, slot, 0, precision, k, args.apply_reps, byref(ms))
        timings[precision] = ms.value
    main_precision = 2 if tile else 1
    achieved = alg_bytes / (timings[main_precision] * 1e-3) / 1e9
    mg.close()
    t_fp64 = t_fp64_same = t_cols = None
    it64 = it64_same = it_cols = None
    if not args.no_fp64:
        # the reference's scheme in FP64, timed like the mixed leg (same warm-up and timed steps)
        mg64, t64, rep64, _ = run({"precision": "float64"}, args.steps, args.warmup, "fp64")
        mg64.close()
        t_fp64, it64 = float(np.mean(t64)), max(r[0] for r in rep64)
        # FP64 with the mixed leg's algorithm (level-1 Krylov solve, degree growth, flexible FP64 PCG): the part
        # of the speed-up that precision alone buys
        same = {"precision": "float64", "coarse_krylov": args.coarse_krylov, "krylov_levels": args.krylov_levels,
                "growth": args.growth, "chebyshev_degree": args.degree, "chebyshev_kind": args.chebyshev_kind,
                "chebyshev_ratio": args.chebyshev_ratio}
        mgs, ts, reps, _ = run(same, min(args.steps, 3), 1, "fp64-mixed-algorithm")
        mgs.close()
        t_fp64_same, it64_same = float(np.mean(ts)), max(r[0] for r in reps)
        if args.block:  # the mixed leg with one recurrence per column: the part of the speed-up block PCG buys
            mgc, tc, repc, _ = run({**mixed, "block_krylov": False}, min(args.steps, 3), 1, "mixed-per-column")
            mgc.close()
            t_cols, it_cols = float(np.mean(tc)), max(r[0] for r in repc)
    baseline = None
    if not args.no_cpu_baseline:
        try:
            baseline = cpu_baseline()
        except Exception as error:  # the baseline is reported, never required
            baseline = {"error": str(error)}
    speedup = t_fp64 / t_mixed if t_fp64 else None
    out = {
        "metric": "42-RHS solve time to rel-res 1e-8 at 9.5M DOF: learned+mixed vs plain FP64",
        "value": speedup, "unit": "x (FP64 time / mixed time)", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_mixed * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 V-cycle / fp64 outer", "data": "synthetic",
        "config": {"workload": f"{args.config} {cfg['shape'][0]}x{cfg['shape'][1]}x{cfg['shape'][2]} hex8, "
                               f"{ndof} DOF, {k} RHS, {cfg['levels']}-level GMG", "scheme": args.scheme,
                   "learned": learned is not None, "learned_weights": args.learned,
                   "mixed_s": t_mixed, "fp64_s": t_fp64, "fp64_iterations": it64,
                   "fp64_mixed_algorithm_s": t_fp64_same, "fp64_mixed_algorithm_iterations": it64_same,
                   "mixed_per_column_s": t_cols, "mixed_per_column_iterations": it_cols,
                   # FP64 -> FP32 with the same per-column algorithm, then per-column -> block PCG
                   "precision_speedup": (t_fp64_same / (t_cols or t_mixed)) if t_fp64_same else None,
                   "block_speedup": (t_cols / t_mixed) if t_cols else None,
                   "iterations": max(r[0] for r in reports),
                   "max_rel_res": max(r[1] for r in reports), "fine_smoother": args.fine_smoother,
                   "coarse_krylov": args.coarse_krylov, "krylov_levels": args.krylov_levels, "growth": args.growth,
                   "chebyshev_degree": args.degree, "chebyshev_kind": args.chebyshev_kind,
                   "chebyshev_ratio": args.chebyshev_ratio,
                   "inner_tolerance": args.inner_tolerance,
                   "block_krylov": args.block, "fine_tile": list(tile[:2]) if tile else None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.config, main_precision),
                     "traffic_source": str(TRAFFIC[main_precision].relative_to(ROOT)),
                     "kernel": KERNELS[main_precision], "ms_per_launch": timings[main_precision],
                     "algorithmic_bytes": alg_bytes,
                     "fp32_vector_kernel": {"ms_per_launch": timings[1],
                                            "achieved_GBs": alg_bytes / (timings[1] * 1e-3) / 1e9}},
        "cpu_baseline": baseline,
    }
    print(json.dumps(out))


def _events_ms(torch, fn, reps=5):
    """Mean milliseconds of fn() between HIP events on the current stream (the library launches there)."""
    fn()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        fn()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / reps


def run_rows(args):
    """One JSON line with a measurement of each caller around the solve (covariance, modal, floating, density)."""
    import numpy as np
    import torch
    from deep_frame_amd import element, problems
    from deep_frame_amd.covariance import InterfaceCovariance
    from deep_frame_amd.density import DensityMap
    from deep_frame_amd.modal import MultigridModal
    from deep_frame_amd.multigrid import GeometricMultigrid
    dev = torch.device("cuda:0")
    ke = element.stiffness((1, 1, 1), element.isotropic(0.3))
    rows = {}

    # interface covariance on the 42 solutions of the 256 x 256 x 48 frame
    shape = CONFIGS["frame"]["shape"]
    prob = problems.frame(shape, seed=0, rhs=42)
    mod, dmod = problems.moduli(prob["density"])
    mg = GeometricMultigrid(shape, (1, 1, 1), ke, {"precision": "float32", "refine": True, "levels": 5})
    mg.update(torch.as_tensor(mod, device=dev))
    f = torch.as_tensor(np.ascontiguousarray(prob["forces"].T), device=dev)
    u, _ = mg.solve_device(b"cov", prob["fixed"], f)
    g = np.random.default_rng(0).normal(size=(42, 42))
    cov = InterfaceCovariance(g @ g.T / 42 + np.eye(42))
    dm = torch.as_tensor(dmod, device=dev)
    gram_ms = _events_ms(torch, lambda: cov.unit_flexibility(u, f))
    limits = {"mean_n_mm": 1.0, "worst_n_mm": 1.0}
    cov.measure(mg, u, f, dm, limits)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        cov.measure(mg, u, f, dm, limits)
    torch.cuda.synchronize()
    gram_bytes = 2.0 * f.numel() * 8
    rows["covariance"] = {"workload": "InterfaceCovariance.measure, 42 unit-wrench solutions, 9,709,203 DOF",
                          "measure_ms": (time.perf_counter() - t0) / 3 * 1e3, "gram_ms": gram_ms,
                          "gram_roofline": {"bound": "hbm", "achieved": gram_bytes / (gram_ms * 1e-3) / 1e9,
                                            "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                            "frac": gram_bytes / (gram_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}}
    print(f"# covariance {rows['covariance']}", file=sys.stderr, flush=True)

    # robust projection and pullback on the same grid
    n = int(np.prod(shape))
    allowed = np.ones(n, bool)
    dmap = DensityMap(shape, (1, 1, 1), allowed, np.zeros(n, bool), ~allowed,
                      {"projection": "robust", "projection_beta": 8.0, "filter_radius_mm": 1.5})
    design = torch.as_tensor(prob["density"].ravel(), device=dev)
    filtered = dmap.filtered(design)
    etas = list(dmap.thresholds.values())
    project_ms = _events_ms(torch, lambda: dmap._project(filtered, etas))
    fields, _ = dmap.fields(design)
    sens = torch.as_tensor(np.random.default_rng(1).normal(size=n), device=dev)
    fields_ms = _events_ms(torch, lambda: dmap.fields(design))
    pull_ms = _events_ms(torch, lambda: dmap.pullback(sens, fields["eroded"][1]))
    project_bytes = n * (8.0 + 1.0 + 2 * 8.0 * len(etas))
    rows["density"] = {"workload": "DensityMap robust fields (3 thresholds) and pullback, 3,145,728 cells, r = 1.5",
                       "fields_ms": fields_ms, "pullback_ms": pull_ms, "project_ms": project_ms,
                       "project_roofline": {"bound": "hbm", "achieved": project_bytes / (project_ms * 1e-3) / 1e9,
                                            "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                            "frac": project_bytes / (project_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}}
    dmap.close()
    mg.close()
    del u, f
    print(f"# density {rows['density']}", file=sys.stderr, flush=True)

    # LOBPCG modes on the 0.75 mm half-domain size
    shape = (91, 170, 48)
    prob = problems.frame(shape, seed=0, rhs=6)
    rho = prob["density"].ravel()
    mg = GeometricMultigrid(shape, (1, 1, 1), ke, {"precision": "float32"}, element.mass((1, 1, 1)))
    modal = MultigridModal(mg, [{"key": b"modes", "fixed": prob["fixed"]}], 1.0, 1.0,
                           settings={"modes": 6, "tracked": 3, "f1_min_hz": 1.0})
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    modal.measure(rho)
    cold = time.perf_counter() - t0
    first = modal.reports[0]
    t0 = time.perf_counter()
    modal.measure(np.clip(rho * 1.01, 0, 1))
    warm = time.perf_counter() - t0
    rows["modal"] = {"workload": f"MultigridModal.measure, 6 modes, frame {shape[0]}x{shape[1]}x{shape[2]} "
                                 f"({mg.ndof} DOF), cold then warm-started on a perturbed density",
                     "cold_s": cold, "cold_iterations": first["iterations"], "warm_s": warm,
                     "warm_iterations": modal.reports[0]["iterations"],
                     "converged": bool(first["converged"] and modal.reports[0]["converged"])}
    mg.close()
    print(f"# modal {rows['modal']}", file=sys.stderr, flush=True)

    # floating 3-2-1 support against the clamped pads on the 96 x 96 x 16 frame
    shape = CONFIGS["frame_small"]["shape"]
    prob = problems.frame(shape, seed=0, rhs=42)
    mod, _ = problems.moduli(prob["density"])
    points = element.node_points(shape)
    support = problems.relief_support(points, np.flatnonzero(points[:, 2] == 0))
    result = {}
    for name, fixed, sup in (("clamped", prob["fixed"], None), ("floating", support, support)):
        mg = GeometricMultigrid(shape, (1, 1, 1), ke, {"precision": "float32", "levels": 4})
        mg.update(mod)
        mg.solve(name, fixed, prob["forces"], sup)
        t0 = time.perf_counter()
        _, rep = mg.solve(name, fixed, prob["forces"], sup)
        result[name] = {"s": time.perf_counter() - t0, "iterations": max(rep["iterations"]),
                        "converged": rep["converged"], "kernel": rep["kernel"]}
        mg.close()
    rows["floating"] = {"workload": "GeometricMultigrid.solve with a 3-2-1 support vs the clamped pads, frame "
                                    "96x96x16, 42 RHS (host arrays in and out)", **result}
    print(f"# floating {rows['floating']}", file=sys.stderr, flush=True)

    # the optimised 0.75 mm half-frame density the reference exports (its multigrid probe stalled at 4.2e-8)
    from deep_frame_amd import formats
    density = formats.load_density_half(ROOT / "tests" / "golden" / "density_half_075mm_run3.npz")
    shape = density.shape
    prob = problems.frame(shape, seed=0, rhs=42)
    mod = 1.0 * (1e-6 + (1 - 1e-6) * density.ravel() ** 3)
    points = element.node_points(shape)
    pins = problems.relief_support(points, np.flatnonzero(points[:, 2] == 0))
    bench_mixed = {"precision": "float32", "refine": True, "coarse_krylov": 2, "krylov_levels": 2, "growth": 1.0,
                   "inner_tolerance": 3e-3}
    schemes = (("clamped", prob["fixed"], None, {"precision": "float32", "refine": True}),
               ("clamped_fp64", prob["fixed"], None, {"precision": "float64"}),
               ("clamped_flexible", prob["fixed"], None, {"precision": "float32"}),
               ("clamped_bench_mixed", prob["fixed"], None, bench_mixed),
               ("clamped_bench_mixed_block", prob["fixed"], None, {**bench_mixed, "block_krylov": True}),
               ("floating", pins, pins, {"precision": "float32", "refine": True}))
    result = {}
    for name, fixed, sup, settings in schemes:
        mg = GeometricMultigrid(shape, (1, 1, 1), ke, settings)
        mg.update(mod)
        mg.solve(name, fixed, prob["forces"], sup)
        t0 = time.perf_counter()
        _, rep = mg.solve(name, fixed, prob["forces"], sup)
        result[name] = {"s": time.perf_counter() - t0, "iterations": max(rep["iterations"]),
                        "max_rel_res": max(rep["relative_residual"]), "converged": rep["converged"],
                        "retried_columns": rep["retried_columns"], "settings": settings}
        mg.close()
        print(f"# exported_075mm {name}: {result[name]['s']:.3f} s, {result[name]['iterations']} iterations",
              file=sys.stderr, flush=True)
    rows["exported_075mm"] = {"workload": "42 unit-wrench solves on the exported optimised 0.75 mm half-frame density "
                                          "(91x170x48, 2,312,604 DOF, stiffness contrast 1e6): clamped pads with the "
                                          "mixed V-cycle, the reference's FP64 scheme, its default scheme (FP64 PCG, "
                                          "FP32 V-cycle), the benchmark's mixed "
                                          "configuration and the same as one block PCG; a floating 3-2-1 support "
                                          "(host arrays in and out)", **result}
    print(f"# exported_075mm {rows['exported_075mm']}", file=sys.stderr, flush=True)
    print(json.dumps({"metric": "callers around the solve", "n_gpus": 1, "data": "synthetic", "rows": rows}))


def run_variants(args, cfg):
    """Batched uneven variants: total DOF x RHS solved per second, per-variant error vs FP64."""
    import numpy as np
    from deep_frame_amd import variants
    vs = variants.generate(count=args.variants or cfg["count"], seed=0)
    print(f"# variants {len(vs)}, total RHS {sum(v['rhs'] for v in vs)}", file=sys.stderr, flush=True)
    variants.run(vs[:1])  # warm-up
    out = variants.run(vs, check=not args.no_fp64, streams=args.streams)
    recs = out["variants"]
    summary = {"max_iterations": max(r["iterations"] for r in recs), "max_rel_res": max(r["max_rel_res"] for r in recs),
               "all_converged": all(r["converged"] for r in recs), "seconds": out["seconds"]}
    if not args.no_fp64:
        summary["max_compliance_rel_err_vs_fp64"] = max(r["compliance_rel_err"] for r in recs)
        summary["max_sensitivity_rel_err_vs_fp64"] = max(r["sensitivity_rel_err"] for r in recs)
    print(json.dumps({
        "metric": "batched uneven variants: DOF x RHS solved to rel-res 1e-8 per second (with sensitivities)",
        "value": out["dof_rhs_per_s"], "unit": "DOF*RHS/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": out["seconds"] * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 inner PCG / fp64 refinement", "data": "synthetic",
        "config": {"workload": f"{len(vs)} frame variants, nx, ny in [48, 160], nz in [12, 32], 6-42 RHS",
                   "streams": args.streams,
                   **summary, "per_variant": recs}}))


if __name__ == "__main__":
    main()
