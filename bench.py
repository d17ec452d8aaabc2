"""Benchmark: particle-steps/s of the two-phase per-step loop at 1M particles.

Workload: the partially submerged 30-degree slope (sources_amd.scenarios.slope
at dx = 0.001, ~1.0M real + wall particles), FP64, full leapfrog steps
(time_integration) with the state resident in HBM.  Prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale S] [--no-cpu]

The timed region is K steps bracketed by device synchronisation, timed with
HIP events recorded on the library's stream.  A second, shorter pass with
per-launch HIP events gives each kernel's average duration for the roofline.
The CPU baseline times the oracle restatement of the reference algorithm
(pair lists by link_list, O(N_w N_s) direct_find_2, scatter passes; searches
and passes split over up to 8 OpenMP threads, the reference's cap) on a
bounded sample of the same scenario.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 1M two-phase SPH; learned-vs-plain speedup at ≤1% ρ err"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per particle and step of the reference loop with two
# neighbour stages (state read once per stage, final state written once,
# stage-A -> stage-B intermediates written and read once; neighbour re-reads
# on chip): water 368, soil 456, cell list 28 (both phases).  The roofline of
# the dominant kernel (the soil stage pass, soil_ab) takes the soil figure per
# soil particle of its launch, as if that pass moved all of the soil's bytes;
# "frac_split_model" beside it takes the share the library's stage split
# assigns to that pass: 124 state + 48 intermediates + 112 outputs = 284.
# Mixed precision (integrated fields x, v, rho, vof, stress, p in FP64, the
# rest FP32): water 260, soil 344 (split share 216).
ALGO_BYTES = {"fp64": {"water": 368, "soil": 456, "cell": 28},
              "mixed": {"water": 260, "soil": 344, "cell": 28}}
SPLIT_BYTES = {"fp64": {"soil_ab": 284}, "mixed": {"soil_ab": 216}}


class Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")   # the runtime libsphg.so links
        self.h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0


def cpu_baseline(args):
    """The oracle's threaded timing build on a bounded sample, in a child
    process (its OpenMP thread count is fixed before the runtime starts)."""
    import subprocess
    threads = max(1, min(8, len(os.sched_getaffinity(0))))   # the reference caps OpenMP at 8 threads
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", "--cpu-steps", str(args.cpu_steps),
           "--cpu-dx", str(args.cpu_dx)]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline child failed: {r.stderr[-2000:]}")
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out["cores"] = threads
    return out


def cpu_child(args):
    """Oracle restatement of the reference loop (liboracle_omp.so: the pair
    searches and pair passes split over OMP_NUM_THREADS threads, as the
    reference's OpenMP build splits them) with nnps = 2 (link_list) on the
    slope at a bounded size."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import subprocess
    import oracle
    from sources_amd import scenarios
    lib = os.path.join(ROOT, "oracle", "liboracle_omp.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle_omp.so"], check=True)
    sc = scenarios.slope(dx=args.cpu_dx, scale=1.0)
    sc.params.nnps = 2
    n = sc.n_particles
    steps = args.cpu_steps
    with oracle.using(lib):
        it = oracle.time_integration(sc.params, sc.water, sc.soil, 0, 1)  # warm-up step (step 1 path)
        t0 = time.perf_counter()
        oracle.time_integration(sc.params, sc.water, sc.soil, it, steps)
        dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", str(len(os.sched_getaffinity(0)))))
    print(json.dumps({
        "value": n * steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
        "sample": f"oracle restatement of the reference loop, FP64, nnps=2 link_list + O(Nw*Ns) direct_find_2, "
                  f"pair searches and pair passes split over {threads} OpenMP threads (oracle/liboracle_omp.so), "
                  f"on the same slope scenario at dx={args.cpu_dx} ({n} particles), {steps} steps after 1 "
                  f"warm-up step, {dt:.1f} s wall",
    }), flush=True)


def learned_section(args, hip):
    """The learned K-step surrogate (coarse mixed-precision step of K dt + the
    network's correction, the FP64 plain step as judge every J advances), its
    ablation (the same rollout with the network off: the coarse step alone)
    and FP64 plain steps, all from the same state of the 250k slope on one
    context."""
    import numpy as np
    from sources_amd import abi, scenarios
    from sources_amd.sph import Solver
    from sources_amd.surrogate import Surrogate

    K, S, M, J = args.learned_k, args.learned_samples, args.learned_advances, args.learned_judge
    # the slope at half the linear size: ~250k particles
    sc = scenarios.slope(dx=args.dx, scale=args.scale * args.learned_scale, max_interaction=0)
    sc.params.precision = abi.SPHG_MIXED   # the coarse steps; judge, targets and reference run FP64
    ntot = sc.water.n + sc.soil.n
    b = Solver(sc.params, sc.water, sc.soil)
    b.time_integration(args.warmup, sync=False)
    hip.sync()
    e0, e1 = hip.event(), hip.event()

    g = Surrogate(b, ksteps=K, max_samples=S + 1, seed=1)
    t0 = time.perf_counter()
    g.record(S)
    l0 = g.train(S, lr=0.0)
    l1 = g.train(args.learned_iters, lr=2e-3)
    t_train = time.perf_counter() - t0
    gains = [float(x) for x in g.gains()[:3]]
    print(f"# surrogate: {S} samples, {args.learned_iters} Adam steps, loss {l0:.3f} -> {l1:.3f} "
          f"in {t_train:.1f}s, gains {gains}", file=sys.stderr, flush=True)

    b.sync_all()
    base = sc.copy()
    it0 = b.itimestep

    def reset():
        for ph, bp in ((sc.water, base.water), (sc.soil, base.soil)):
            for k, v in vars(bp).items():
                if isinstance(v, np.ndarray):
                    np.copyto(getattr(ph, k), v)
        b.upload_all()
        b.itimestep = it0
        hip.sync()

    def timed(fn):
        hip.record(e0, b.stream)
        r = fn()
        hip.record(e1, b.stream)
        return hip.elapsed_ms(e0, e1), r

    def precision(p):
        sc.params.precision = p
        b.set_params(sc.params)

    # the plain FP64 reference: the same context and state, M K steps
    reset()
    precision(abi.SPHG_FP64)
    ms_plain0, _ = timed(lambda: b.time_integration(M * K, sync=False))
    b.sync_all()
    ref = sc.copy()
    print(f"# plain FP64: {M * K} steps in {ms_plain0:.1f} ms", file=sys.stderr, flush=True)

    def rho_err():
        e, sq, n = 0.0, 0.0, 0
        for pa, pb in ((ref.water, sc.water), (ref.soil, sc.soil)):
            real = pa.itype[:pa.n] > 0
            d = np.abs(pb.rho[:pb.n][real] - pa.rho[:pa.n][real]) / np.abs(pa.rho[:pa.n][real])
            e, sq, n = max(e, float(np.max(d))), sq + float(np.sum(d * d)), n + d.size
        return e, (sq / max(n, 1)) ** 0.5

    legs = {}
    for use_net in (True, False):
        # each leg beside a plain run of its own from the same state (the
        # GPU's speed drifts between runs: the plain time is taken again
        # right before the leg)
        reset()
        precision(abi.SPHG_FP64)
        ms_plain, _ = timed(lambda: b.time_integration(M * K, sync=False))
        reset()
        precision(abi.SPHG_MIXED)
        ms, (nfb, jmax, jrms) = timed(lambda: g.rollout(M, judge_every=J, tol=0.01, use_net=use_net))
        b.sync_all()
        emax, erms = rho_err()
        legs[use_net] = {"ms": ms, "ms_plain_fp64": ms_plain, "speedup_vs_plain_fp64": ms_plain / ms,
                         "value": ntot * M * K / (ms * 1e-3),
                         "rho_rel_err_max_vs_fp64": emax, "rho_rel_err_rms_vs_fp64": erms,
                         "meets_1pct_gate": emax <= 0.01, "fallbacks": nfb, "judge_max_rho_err": jmax,
                         "judge_rms_rho_err": jrms}
        print(f"# learned ({'network' if use_net else 'network off'}): {M} advances of {K} steps in {ms:.1f} ms "
              f"(plain {ms_plain:.1f} ms), {nfb} fallbacks, final rho err max {emax:.2e} rms {erms:.2e}",
              file=sys.stderr, flush=True)
    g.close()
    b.close()
    net, off = legs[True], legs[False]
    return {
        "learned": {
            "particles": ntot, "k_steps": K, "advances": M, "judge_every": J, "tol": 0.01,
            "steps_covered": M * K, "ms_plain_fp64": net["ms_plain_fp64"], "ms_learned": net["ms"],
            "value": net["value"], "unit": "particle-steps/s",
            "speedup_vs_plain_fp64": net["speedup_vs_plain_fp64"],
            "rho_rel_err_max_vs_fp64": net["rho_rel_err_max_vs_fp64"],
            "rho_rel_err_rms_vs_fp64": net["rho_rel_err_rms_vs_fp64"],
            "meets_1pct_gate": net["meets_1pct_gate"], "fallbacks": net["fallbacks"],
            "judge_max_rho_err": net["judge_max_rho_err"], "judge_rms_rho_err": net["judge_rms_rho_err"],
            "network_off": off,
            "network_beats_ablation": (net["rho_rel_err_max_vs_fp64"] < off["rho_rel_err_max_vs_fp64"] and
                                       net["rho_rel_err_rms_vs_fp64"] < off["rho_rel_err_rms_vs_fp64"]),
            "train": {"samples": S, "iters": args.learned_iters, "loss_before": l0, "loss_after": l1,
                      "seconds": t_train, "gains": gains},
            "rates_precision": "mixed",
        },
    }


def mixed_error_section(args, hip):
    """Mixed precision against FP64 on the 1M slope from the same state:
    norm-wise errors (max|mixed - fp64| / max|fp64|) of density, velocity and
    stress, and the largest relative density error, after 1 and N steps; the
    N - 1 steps are timed for both precisions."""
    import numpy as np
    from sources_amd import abi, scenarios
    from sources_amd.sph import Solver

    sf = scenarios.slope(dx=args.dx, scale=args.scale, max_interaction=0)
    sm = sf.copy()
    sm.params.precision = abi.SPHG_MIXED
    a = Solver(sf.params, sf.water, sf.soil)
    b = Solver(sm.params, sm.water, sm.soil)
    out = {}
    done = 0
    e0, e1 = hip.event(), hip.event()
    timing = {}
    for n in (1, args.mixed_error_steps):
        for tag, s in (("fp64", a), ("mixed", b)):
            hip.record(e0, s.stream)
            s.time_integration(n - done, sync=False)
            hip.record(e1, s.stream)
            timing[tag] = hip.elapsed_ms(e0, e1) / max(n - done, 1)
            s.sync_all()
        done = n
        e = {}
        for f, phases in (("rho", ("water", "soil")), ("vx", ("water", "soil")), ("sxx", ("soil",)),
                          ("sxy", ("soil",)), ("syy", ("soil",)), ("p", ("water", "soil"))):
            for ph in phases:
                x, y = getattr(getattr(sm, ph), f), getattr(getattr(sf, ph), f)
                nn = getattr(sf, ph).n
                e[f"{ph}_{f}"] = float(np.max(np.abs(x[:nn] - y[:nn])) / max(np.max(np.abs(y[:nn])), 1e-300))
        rr = 0.0
        for ph in ("water", "soil"):
            pf, pm = getattr(sf, ph), getattr(sm, ph)
            real = pf.itype[:pf.n] > 0
            rr = max(rr, float(np.max(np.abs(pm.rho[:pm.n][real] - pf.rho[:pf.n][real]) / np.abs(pf.rho[:pf.n][real]))))
        e["rho_max_rel"] = rr
        out[f"after_{n}_steps"] = e
        print(f"# mixed vs fp64 after {n} steps: rho max rel {rr:.2e}", file=sys.stderr, flush=True)
    a.close()
    b.close()
    ntot = sf.water.n + sf.soil.n
    return {"mixed_error_vs_fp64": out,
            "mixed": {"ms_per_step": timing["mixed"], "value": ntot / (timing["mixed"] * 1e-3),
                      "unit": "particle-steps/s", "particles": ntot, "steps_timed": args.mixed_error_steps - 1,
                      "fp64_ms_per_step_same_run": timing["fp64"],
                      "speedup_vs_fp64": timing["fp64"] / timing["mixed"]}}


def pcie_section(args, hip):
    """Drop-in (host-buffer) rate: one step per call with the whole state
    uploaded before and downloaded after, as at the reference call site."""
    from sources_amd import scenarios
    from sources_amd.sph import Solver

    sc = scenarios.slope(dx=args.dx, scale=args.scale, max_interaction=0)
    s = Solver(sc.params, sc.water, sc.soil)
    s.time_integration(2, sync=True)
    n = 3
    t0 = time.perf_counter()
    for _ in range(n):
        s.upload_all()
        s.time_integration(1, sync=True)
    ms = (time.perf_counter() - t0) * 1e3 / n
    ntot = sc.water.n + sc.soil.n
    s.close()
    return {"pcie_inclusive": {"ms_per_step": ms, "value": ntot / (ms * 1e-3), "unit": "particle-steps/s",
                               "what": "per call: upload of the whole state, one step, download of the whole state"}}


def sweep_section(args, hip):
    """64 independent submerged-column cases of uneven size (one seed):
    eager per-context launches against the batched hipGraph replays (staged:
    one launch per kernel for all cases after each period's rebuild step)."""
    from sources_amd import scenarios
    from sources_amd.batch import Batch
    from sources_amd.sph import Solver

    from sources_amd import abi
    cases = scenarios.sweep(args.sweep_cases, seed=11)
    for c in cases:
        c.params.precision = abi.SPHG_MIXED if args.sweep_precision == "mixed" else abi.SPHG_FP64
    ntot = sum(c.water.n + c.soil.n for c in cases)
    ss = [Solver(c.params, c.water, c.soil) for c in cases]
    # four captured periods (a period: the neighbour-list rebuild period, 64
    # steps)
    period = 64
    steps = 4 * period
    for s in ss:
        s.time_integration(2, sync=False)
    hip.sync()
    t0 = time.perf_counter()
    for s in ss:
        s.time_integration(steps, sync=False)
    hip.sync()
    ms_eager = (time.perf_counter() - t0) * 1e3 / steps
    b = Batch(ss)
    ns, _ = b.tune(trial_periods=1)
    b.run(period)
    hip.sync()
    t0 = time.perf_counter()
    b.run(steps)
    hip.sync()
    ms_graph = (time.perf_counter() - t0) * 1e3 / steps
    staged = b.staged
    print(f"# sweep: {len(cases)} cases, {ntot} particles: eager {ms_eager:.2f} ms/step, "
          f"batched {ms_graph:.2f} ms/step (tuned: {ns or 'eager'} streams, staged {staged})", file=sys.stderr,
          flush=True)
    b.close()
    for s in ss:
        s.close()
    return {"sweep": {
        "cases": len(cases), "particles": ntot, "steps": steps, "streams": ns,
        "ms_per_step_eager": ms_eager, "ms_per_step_batched": ms_graph,
        "batched_mode": "eager" if ns == 0 else (
            "hipGraph, every step one launch per kernel for all cases" if staged else f"hipGraph, {ns} streams"),
        "value_batched": ntot / (ms_graph * 1e-3), "unit": "particle-steps/s",
        "speedup_batched_vs_eager": ms_eager / ms_graph,
        "workload": "submerged column collapses, N ~ logU[5k,120k], aspect ~ U[0.5,2], "
                    f"phi ~ U[25,40] deg, k_s ~ logU[1e-4,1e-2], seed 11, {args.sweep_precision}",
    }}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--dx", type=float, default=0.001)
    ap.add_argument("--prof-steps", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-dx", type=float, default=0.0025, help="slope spacing of the CPU baseline sample")
    ap.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the mixed and learned-surrogate section")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched scenario sweep section")
    ap.add_argument("--sweep-cases", type=int, default=64)
    ap.add_argument("--sweep-precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--mixed-error-steps", type=int, default=1000,
                    help="mixed vs FP64 error is reported after 1 step and after this many (0: skip)")
    ap.add_argument("--no-pcie", action="store_true")
    # 3 dt is the largest coarse step that stays stable on the slope (4 dt diverges)
    ap.add_argument("--learned-k", type=int, default=4)
    ap.add_argument("--learned-samples", type=int, default=8)
    ap.add_argument("--learned-iters", type=int, default=400)
    ap.add_argument("--learned-advances", type=int, default=30)
    ap.add_argument("--learned-judge", type=int, default=5,
                    help="plain FP64 check every this many advances (5 x 4 steps: every 20 steps)")
    ap.add_argument("--learned-scale", type=float, default=0.5, help="slope size for the learned section")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="fp64: reference arithmetic; mixed: FP32 Kahan pair sums, FP64 integration")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (read MIB, write MIB) five times: "
                         "the known-size kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.cpu_child:
        cpu_child(args)
        return
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py: the library runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the slab decomposition over several GPUs is designed, not built)", file=sys.stderr)
        sys.exit(2)

    from sources_amd import scenarios
    from sources_amd.sph import Solver

    t_gen = time.perf_counter()
    sc = scenarios.slope(dx=args.dx, scale=args.scale, max_interaction=0)
    from sources_amd import abi
    sc.params.precision = abi.SPHG_MIXED if args.precision == "mixed" else abi.SPHG_FP64
    nw, ns = sc.water.n, sc.soil.n
    ntot = nw + ns
    print(f"# generated slope: water {sc.water.ntotal}+{sc.water.nvirt}, soil {sc.soil.ntotal}+{sc.soil.nvirt}, "
          f"total {ntot} in {time.perf_counter() - t_gen:.1f}s", file=sys.stderr, flush=True)
    s = Solver(sc.params, sc.water, sc.soil)
    hip = Hip()
    stream = s.stream

    s.time_integration(args.warmup, sync=False)
    hip.sync()
    e0, e1 = hip.event(), hip.event()
    w0 = time.perf_counter()
    hip.record(e0, stream)
    s.time_integration(args.steps, sync=False)
    hip.record(e1, stream)
    ms = hip.elapsed_ms(e0, e1)
    hip.sync()
    wall = time.perf_counter() - w0
    s.sync_all()  # checks error flags and pulls the state
    print(f"# timed {args.steps} steps: {ms:.3f} ms (events), {wall * 1e3:.3f} ms (wall)", file=sys.stderr, flush=True)
    import numpy as np
    finite = all(bool(np.isfinite(getattr(p, f)[:p.n]).all()) for p in (sc.water, sc.soil)
                 for f in ("x", "vx", "rho", "p"))
    ms_step = ms / args.steps
    value = ntot * args.steps / (ms * 1e-3)

    # per-kernel durations (separate pass: per-launch events add overhead)
    s.profile(True)
    s.profile_reset()
    s.time_integration(args.prof_steps, sync=False)
    hip.sync()
    kern = {}
    for name in ("cell_key", "cell_scatter", "cell_sort", "scan", "fill", "first_half", "pack", "permute", "iota",
                 "nlist", "water_x", "water_a", "soil_x", "soil_a", "soil_ab", "water_b", "soil_b", "final_water",
                 "final_soil"):
        n, t = s.profile_get(name)
        if n:
            kern[name] = {"launches": n, "avg_us": 1e3 * t / n, "total_ms_per_step": t / args.prof_steps}
    s.profile(False)
    pairs_w, pairs_s, pairs_x = s.pair_counts()
    nl_rebuilds, nl_off, nl_skin = s.nlist_stats()
    sb = ALGO_BYTES[args.precision]
    dom = "soil_ab"
    algo = sb["soil"] * ns
    algo_split = SPLIT_BYTES[args.precision][dom] * ns
    avg_s = kern[dom]["avg_us"] * 1e-6
    achieved = algo / avg_s / 1e9
    achieved_split = algo_split / avg_s / 1e9
    traffic = None
    # HBM bytes per launch of the dominant kernel at this workload, from the
    # committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated by a copy)
    tpath = os.path.join(ROOT, "profiles", f"traffic_slope_1m_{args.precision}.md")
    if os.path.exists(tpath) and args.dx == 0.001 and args.scale == 1.0:
        with open(tpath) as f:
            txt = f.read()
        blk = txt.split("```json", 1)[-1].split("```", 1)[0]
        traffic = json.loads(blk).get(dom)
    step_bytes = (sb["water"] + sb["cell"]) * nw + (sb["soil"] + sb["cell"]) * ns
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded slope scenario)",
        "config": {
            "workload": "two-phase 30-degree slope, partially submerged, dx=0.001, FP64 leapfrog time_integration",
            "particles": ntot, "water": nw, "soil": ns,
            "water_real": sc.water.ntotal, "soil_real": sc.soil.ntotal,
            "water_pairs": pairs_w, "soil_pairs": pairs_s, "cross_pairs": pairs_x,
            "skf": sc.params.skf, "precision": args.precision,
            "neighbour_lists": {"rebuilds": nl_rebuilds, "off_spells": nl_off, "skin": nl_skin},
        },
        # distinct interacting pairs (water-water, soil-soil, water-soil) per
        # second; the gather visits each same-phase pair from both ends
        "pair_interactions_per_s": (pairs_w + pairs_s + pairs_x) / (ms_step * 1e-3),
        "roofline": {
            "bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "algo_bytes_per_launch": algo, "algo_bytes_per_unit": sb["soil"], "units_per_launch": ns,
            "avg_launch_us": kern[dom]["avg_us"],
            "achieved_split_model": achieved_split, "frac_split_model": achieved_split / HBM_PEAK_GBS,
            "algo_bytes_per_launch_split_model": algo_split,
        },
        "step_roofline": {
            "algo_bytes_per_step": step_bytes, "achieved_GBs": step_bytes / (ms_step * 1e-3) / 1e9,
            "frac": step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "kernels": kern,
        "finite_state": finite,
        "wall_ms_timed": wall * 1e3,
    }
    s.close()
    if not args.no_learned:
        out.update(learned_section(args, hip))
    if not args.no_sweep:
        out.update(sweep_section(args, hip))
    if args.mixed_error_steps > 1:
        out.update(mixed_error_section(args, hip))
    if not args.no_pcie:
        out.update(pcie_section(args, hip))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
