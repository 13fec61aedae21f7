"""Benchmark: FMM evaluations/s on one MI355X for the 1M-particle vortex ring.

Workload (the configuration the headline metric is quoted on): N = 1,000,000
point vortices on a ring of radius 1 with a Gaussian core (sigma 0.1),
tangential strengths with 1e-3 relative noise; velocity + velocity gradient;
error-controlled expansion order PowerRelativeGradient with p <= 20;
leaf size 64, MAC 0.55, shrink_recenter. The reference's tolerance is a
per-interaction heuristic, so the default tolerance (1e-9) and multipole
acceptance (0.55; not fixed by the workload) are the fastest pair measured
whose velocity and velocity-gradient errors against the FP64 direct sum both
stay <= 1e-6 on every sampled target for the FP64 form; the line reports each
form's measured max / p99 errors. One step = one complete fmm!
evaluation (both trees, interaction lists, estimate_influence pre-pass, P2P,
upward, M2L, downward) on inputs already resident in HBM.

Prints one JSON line. Both forms (mixed: FP32 nearfield; fp64) are timed in
the same run, each with its error against the FP64 O(N^2) direct sum on
10,000 sampled targets. `value` is the evaluations/s of the fastest form whose
max velocity and gradient errors meet the 1e-6 bar (the mixed form if none
does, flagged).
"""
import argparse
import json
import math
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3     # MI355X dense FP32 vector/matrix peak
PEAK_FP64_TFLOPS = 78.6
FLOPS_PER_PAIR_VORTEX = 63   # v + grad v per body pair (sqrt and divide counted once)


def m2l_flops(p, lh=True):
    """Algorithmic flops of one rotation-based M2L at order p (Wigner-matrix build excluded)."""
    H = (p + 1) * (p + 2) // 2
    f = 14 * (H - 1) + 8 * sum((n + 1) * (2 * n + 1) for n in range(p + 1)) + \
        4 * sum(p - m + 1 for n in range(p + 1) for m in range(n + 1))
    return 2 * f + 8 * H if lh else f


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pmc_traffic(kernel_key):
    """HBM bytes per launch of a kernel from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (profiles/ring1m_traffic_pmc.md, calibrated against a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "ring1m_traffic_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.rstrip("\n").split("\t")
        if len(parts) == 2 and parts[0] == kernel_key:
            return float(parts[1])
    return None


def serving_sample(F, W, _lib, args):
    """Serving stream: Poisson arrivals mixing fmm! evaluations (N log-uniform in [1e3, 2e6], vortex
    rings and uniform-cube sources alternating, the headline's accuracy settings) with
    FastGaussSeidel solves of quad source panels tiling a unit sphere (N in [20k, 200k], random
    normal-velocity right-hand side, residual tolerance 1e-8), through the request server (HBM-budget
    admission, requests overlapping on the device). Every request's inputs are generated before the
    clock starts; latency is submit -> completion as the server measures it."""
    rng = np.random.default_rng(7)
    em = F.PowerRelativeGradient(args.tol)
    pool = []
    for i in range(args.serve_requests):
        if i % 5 == 4:
            m = int(round(math.sqrt(math.exp(rng.uniform(math.log(2e4), math.log(2e5))) / 6)))
            rows, kind = W.sphere_panels(m, quads=True, seed=100 + i)
            ext = rng.standard_normal(len(rows))
            pool.append(("solve", _RowSystem(kind, rows), ext))
        else:
            n = int(math.exp(rng.uniform(math.log(1e3), math.log(2e6))))
            if i % 2:
                pool.append(("fmm", _RowSystem(_lib.POINT_VORTEX, W.vortex_ring(n, seed=100 + i)), None))
            else:
                pool.append(("fmm", _RowSystem(_lib.POINT_SOURCE, W.uniform_sources(n, seed=100 + i)), None))

    def submit(server, item):
        kind, sysm, ext = item
        if kind == "fmm":
            return server.submit_fmm(sysm, leaf_size=args.leaf, expansion_order=args.pmax,
                                     multipole_acceptance=args.mac, error_tolerance=em, gradient=True,
                                     hessian=sysm.fmm_kernel == _lib.POINT_VORTEX)
        return server.submit_solve(sysm, ext, expansion_order=6, leaf_size=30, multipole_acceptance=0.5,
                                   influence=_lib.INFLUENCE_NORMAL_GRADIENT, max_iterations=30, tolerance=1e-8)

    def run(server, rate):
        arrivals = np.cumsum(rng.exponential(1.0 / rate, len(pool))) if math.isfinite(rate) else np.zeros(len(pool))
        t0 = time.perf_counter()
        tickets = []
        for (item, ta) in zip(pool, arrivals):
            delay = t0 + ta - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append((item[0], len(item[1].rows), submit(server, item)))
        lat = {"fmm": [], "solve": []}
        queue, first, last, bodies = [], math.inf, 0.0, 0
        for kind, n, t in tickets:
            _, _, st = server.wait(t)
            lat[kind].append(st["latency_ms"])
            queue.append(st["queue_ms"])
            first = min(first, st["start_ms"] - st["queue_ms"])
            last = max(last, st["end_ms"])
            bodies += n
        wall = (last - first) * 1e-3
        allv = lat["fmm"] + lat["solve"]
        pct = lambda v, q: round(float(np.percentile(v, q)), 3)
        return {"requests": len(pool), "requests_per_s": round(len(pool) / wall, 3),
                "bodies_per_s": round(bodies / wall, 1), "span_s": round(wall, 3),
                "p50_ms": pct(allv, 50), "p99_ms": pct(allv, 99), "mean_queue_ms": round(float(np.mean(queue)), 3),
                "fmm_p50_ms": pct(lat["fmm"], 50), "fmm_p99_ms": pct(lat["fmm"], 99),
                "solve_p50_ms": pct(lat["solve"], 50), "solve_p99_ms": pct(lat["solve"], 99)}

    out = {}
    with F.Server(n_slots=args.serve_slots, hbm_budget_bytes=args.serve_budget_gib * 2**30) as server:
        warm = run(server, math.inf)          # grows every worker's device buffers; not reported
        log(f"serving warm-up burst: {warm}")
        for rate in [float(r) for r in args.serve_rates.split(",")]:
            key = "burst" if not math.isfinite(rate) else f"{rate:g}/s"
            out[key] = run(server, rate)
            log(f"serving at {key}: {out[key]}")
    return {"workload": f"{len(pool)} requests: fmm! N log-uniform [1e3, 2e6] (rings / uniform cubes) and "
                        "FastGaussSeidel quad-panel sphere solves N in [2e4, 2e5] (1 in 5), Poisson arrivals",
            "slots": args.serve_slots, "hbm_budget_gib": args.serve_budget_gib, "rates": out}


def batch_sample(F, W, _lib, args):
    """Batched uneven instances: 512 independent systems, N = floor(exp(U[ln 1e3, ln 2e5])) (seed 4),
    even instances uniform-cube sources (potential + gradient, p 8, MAC 0.5), odd instances vortex rings
    (velocity + gradient, PowerRelativeGradient at the headline's tolerance and MAC, p <= 20); FP32
    nearfield with compensated sums (the mixed form). Each kind's leaf size is chosen by autotune (measured
    device time) on one instance of the body-weighted median size. Each kind runs as one batch evaluation (fmm_batch); for comparison all systems
    are also submitted at once to the request server. Timed regions are the second of two identical rounds
    (the first grows the buffers)."""
    rng = np.random.default_rng(4)
    sizes = np.floor(np.exp(rng.uniform(math.log(1e3), math.log(2e5), args.batch_systems))).astype(int)
    em = F.PowerRelativeGradient(args.tol)
    # tuned on one instance of the body-weighted median size (where half the bodies sit in smaller systems)
    srt = np.sort(sizes)
    med = int(srt[np.searchsorted(np.cumsum(srt), srt.sum() / 2)])
    t0 = time.perf_counter()
    tuned = {}
    for kind in ("cube", "ring"):
        if kind == "cube":
            s = _RowSystem(_lib.POINT_SOURCE, W.uniform_sources(med, seed=5))
            params, ev, _ = F.autotune(s, expansion_order=8, multipole_acceptance=0.5, free_slots=(1,), rounds=1,
                                       scalar_potential=True, gradient=True, precision="mixed")
        else:
            s = _RowSystem(_lib.POINT_VORTEX, W.vortex_ring(med, seed=6))
            params, ev, _ = F.autotune(s, error_tolerance=em, expansion_order=args.pmax, multipole_acceptances=(args.mac,),
                                       free_slots=(1,), rounds=1, gradient=True, hessian=True, precision="mixed")
        ev.close()
        tuned[kind] = params
    t_tune = time.perf_counter() - t0
    systems = []
    for i, n in enumerate(sizes):
        if i % 2 == 0:
            systems.append(("cube", _RowSystem(_lib.POINT_SOURCE, W.uniform_sources(int(n), seed=1000 + i))))
        else:
            systems.append(("ring", _RowSystem(_lib.POINT_VORTEX, W.vortex_ring(int(n), seed=1000 + i))))

    def opts(kind):
        t = tuned[kind]
        if kind == "cube":
            return dict(leaf_size=t["leaf_size"], expansion_order=8, multipole_acceptance=0.5,
                        scalar_potential=True, gradient=True, precision="mixed")
        return dict(leaf_size=t["leaf_size"], expansion_order=args.pmax, multipole_acceptance=args.mac,
                    error_tolerance=em, gradient=True, hessian=True, precision="mixed")

    def summary(dt, lat, status):
        return {"systems": len(systems), "bodies": int(sizes.sum()), "seconds": round(dt, 3),
                "systems_per_s": round(len(systems) / dt, 1), "bodies_per_s": round(float(sizes.sum()) / dt, 1),
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "status": status}

    # one batch evaluation per kind (fmm_batch: a forest of the kind's systems, every pass in shared launches);
    # a system's latency is the end of its kind's batch
    # each kind on a context (stream) of its own, the two batches run concurrently from two host threads
    evs = {k: F.Evaluation(F.Context()) for k in ("cube", "ring")}
    staged = {}   # each kind's systems as one host buffer of concatenated columns (staged before the clock)
    for kind in ("cube", "ring"):
        members = [sysm for k, sysm in systems if k == kind]
        staged[kind] = (members[0].fmm_kernel, np.ascontiguousarray(np.concatenate([m.rows for m in members])),
                        np.array([len(m.rows) for m in members], dtype=np.int64))

    def run_kind(kind, t0, res):
        kernel, buf, counts = staged[kind]
        ev, kw = evs[kind], opts(kind)
        ctx = ev.ctx
        ctx.check(ctx.lib.fmmhip_eval_set_batch(ev._h, kernel, len(counts), counts.ctypes.data, buf.ctypes.data, 0))
        ev.nt = ev.ns = int(counts.sum())
        ev.counts = counts
        ls = kw.pop("leaf_size")
        t1 = time.perf_counter()
        st = ev.fmm(F.make_options(leaf_size_source=ls, leaf_size_target=ls, **kw))
        t2 = time.perf_counter()
        rows = ev.targets()    # results back to the host, as the server's wait delivers them
        t3 = time.perf_counter()
        _, sys_ok = ev.batch_stats()
        outs[kind] = rows
        res[kind] = (len(counts), int(sys_ok.sum()), (t3 - t0) * 1e3,
                     {"copy_in_ms": round((t1 - t0) * 1e3, 1), "fmm_ms": round((t2 - t1) * 1e3, 1),
                      "copy_out_ms": round((t3 - t2) * 1e3, 1),
                      **{k: round(getattr(st, k), 1) for k in ("ms_tree", "ms_lists", "ms_estimate", "ms_p2p",
                                                              "ms_upward", "ms_m2l", "ms_downward")}})

    fused = None
    outs = {}   # each kind's results of the last round, in its systems' order
    for rnd in range(2):
        res = {}
        t0 = time.perf_counter()
        threads = [threading.Thread(target=run_kind, args=(k, t0, res)) for k in ("cube", "ring")]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        dt = time.perf_counter() - t0
        if len(res) != 2:
            raise RuntimeError("a batch evaluation failed")
        lat, status = [], {"ok": 0, "tolerance_missed": 0}
        for kind, (n, n_ok, t_end, _) in res.items():
            lat += [t_end] * n
            status["ok"] += n_ok
            status["tolerance_missed"] += n - n_ok
        fused = summary(dt, lat, status)
        fused["phases"] = {k: v[3] for k, v in res.items()}
        log(f"batch (fused) round {rnd}: {fused}")
    # measured accuracy of the last round: each system's gradient slots (ring: velocity) on 32 sampled
    # targets against the FP64 direct sum over that system's own sources (the status above is the error
    # controller's own verdict, which reports a miss once an interaction's estimate reaches the order cap)
    accuracy = {}
    for kind in ("cube", "ring"):
        kernel, buf, counts = staged[kind]
        ctx = evs[kind].ctx
        offs = np.concatenate([[0], np.cumsum(counts)])
        errs = []
        for k in range(len(counts)):
            src = np.ascontiguousarray(buf[offs[k]:offs[k + 1]])
            pick = np.sort(np.random.default_rng(k).choice(len(src), size=min(32, len(src)), replace=False))
            txyz = np.ascontiguousarray(src[pick, :3])
            ref = np.zeros((len(pick), 16))
            ctx.check(ctx.lib.fmmhip_direct(ctx._h, kernel, len(pick), txyz.ctypes.data, len(src), src.ctypes.data, 0, 1, 0,
                                            _lib.FP64, ref.ctypes.data, 0))
            errs.append(W.relative_error(outs[kind][offs[k] + pick, 4:7], ref[:, 4:7])[0])
        accuracy[kind] = {"slots": "velocity" if kind == "ring" else "gradient",
                          "max_rel_err": float(np.max(errs)), "median_rel_err": float(np.median(errs)),
                          "systems_within_1e-6": int(np.sum(np.array(errs) <= 1e-6)), "systems": len(errs)}
    log(f"batch accuracy: {accuracy}")
    for ev in evs.values():
        ctx = ev.ctx
        ev.close()
        ctx.close()

    server = None
    with F.Server(n_slots=args.batch_slots, hbm_budget_bytes=args.serve_budget_gib * 2**30) as srv:
        for rnd in range(2):
            t0 = time.perf_counter()
            tickets = [srv.submit_fmm(sysm, **opts(kind)) for kind, sysm in systems]
            sts = [srv.wait(t)[2] for t in tickets]
            dt = time.perf_counter() - t0
            status = {"ok": sum(x["status"] == _lib.FMMHIP_OK for x in sts),
                      "tolerance_missed": sum(x["status"] == _lib.FMMHIP_ETOL for x in sts)}
            server = summary(dt, [x["latency_ms"] for x in sts], status)
            log(f"batch (server) round {rnd}: {server}")
    out = dict(fused)
    out.update({"workload": f"{len(systems)} systems, N log-uniform [1e3, 2e5] (seed 4), even: uniform-cube sources "
                            "(potential + gradient, p 8, MAC 0.5), odd: vortex rings (velocity + gradient, "
                            f"PowerRelativeGradient({args.tol}), p <= {args.pmax}, MAC {args.mac}); mixed form",
                "mode": "two batch evaluations (one per kind: a forest of its systems, every pass in shared "
                        "launches), concurrently on two streams; a system's latency is the end of its kind's batch",
                "server": dict(server, slots=args.batch_slots,
                               mode="512 requests to the request server at once, overlapping on the device"),
                "accuracy": dict(accuracy, sample="32 targets per system, FP64 direct sum over the system's sources"),
                "tuned_leaf_size": {k: v["leaf_size"] for k, v in tuned.items()},
                "tuned_on": f"autotune on one instance of each kind at N = {med} (body-weighted median)",
                "tuned_error_success": {k: bool(v["error_success"]) for k, v in tuned.items()},
                "tune_seconds": round(t_tune, 2)})
    return out


def precision_sweep(F, W, _lib, args):
    """Precision sweep: N = 4,194,304 point sources from a Plummer sphere (scale 1, truncated at r = 20),
    strengths N(0,1); potential + gradient at fixed expansion orders 4 / 8 / 16 (leaf args.sweep_leaf,
    MAC args.sweep_mac) in each precision form; error of the gradient (and potential) against the FP64
    O(N^2) direct sum on 10,000 seeded sampled targets; phase times per evaluation (inputs resident in
    HBM, warm-up evaluation first)."""
    n = args.sweep_n
    sbuf = W.plummer_sources(n, seed=3)
    ctx = F.default_context()
    ev = F.Evaluation(ctx)
    ev.set_systems(_lib.POINT_SOURCE, None, sbuf)
    rng = np.random.default_rng(3)
    idx = np.sort(rng.choice(n, size=min(args.samples, n), replace=False))
    txyz = np.ascontiguousarray(sbuf[idx, :3])
    ref = np.zeros((len(idx), 16))
    ref[:, :3] = txyz
    ctx.check(ctx.lib.fmmhip_direct(ctx._h, _lib.POINT_SOURCE, len(idx), txyz.ctypes.data, n,
                                    np.ascontiguousarray(sbuf).ctypes.data, 1, 1, 0, _lib.FP64, ref.ctypes.data, 0))
    out = {}
    for form in args.sweep_forms.split(","):
        for p in (4, 8, 16):
            opt = F.make_options(expansion_order=p, leaf_size_source=args.sweep_leaf, multipole_acceptance=args.sweep_mac,
                                 scalar_potential=True, gradient=True, precision=form)
            ev.fmm(opt)
            t0 = time.perf_counter()
            st = [ev.fmm(opt).as_dict() for _ in range(args.sweep_steps)]
            dt = (time.perf_counter() - t0) / args.sweep_steps
            got = ev.targets()[idx]
            e_g = W.relative_error(got[:, 4:7], ref[:, 4:7])
            e_p = W.relative_error(got[:, 3:4], ref[:, 3:4])
            ph = {k: round(float(np.mean([x[k] for x in st])), 3) for k in
                  ("ms_tree", "ms_lists", "ms_p2p", "ms_upward", "ms_m2l", "ms_downward", "ms_total")}
            s = st[-1]
            rec = {"ms_per_eval": round(dt * 1e3, 3), "evals_per_s": round(n / dt, 1),
                   "rel_err_gradient": {"max": e_g[0], "p99": e_g[1], "median": e_g[2]},
                   "rel_err_potential": {"max": e_p[0], "p99": e_p[1]},
                   "phases_ms": ph, "n_m2l": s["n_m2l"], "p2p_pairs": s["n_p2p_pairs"],
                   "m2l_per_s": round(s["n_m2l"] / (ph["ms_m2l"] * 1e-3), 1) if ph["ms_m2l"] > 0 else None,
                   "m2l_tflops": round(s["n_m2l"] * m2l_flops(p, lh=False) / (ph["ms_m2l"] * 1e-3) / 1e12, 3)
                   if ph["ms_m2l"] > 0 else None}
            out[f"{form}/p{p}"] = rec
            log(f"sweep {form} p={p}: {rec}")
    ev.close()
    return {"workload": f"Plummer sphere N={n} (scale 1, r <= 20), N(0,1) strengths, potential + gradient, "
                        f"leaf {args.sweep_leaf}, MAC {args.sweep_mac}, fixed p; error vs FP64 direct on "
                        f"{len(idx)} sampled targets", "results": out}


def cpu_baseline(W, args):
    """The oracle (C++ FP64 restatement of the reference's fmm! path, OpenMP over targets as the
    reference's multithreaded drivers split the work) on the host cores, on bounded samples of the
    headline workload (same distribution and settings): all cores on an N = args.cpu_sample ring and one
    core on an N = args.cpu_sample_1 ring. Protocol as the reference's benchmark script
    (scripts/benchmark_singlethread_20250402.jl:20-30): one untimed warm-up run, then the median (and
    min) of args.cpu_runs timed runs."""
    from oracle import oracle as O
    cores = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))

    def leg(m, threads, runs):
        cb = W.vortex_ring(m, seed=2)
        O.threads(threads)
        ts = []
        for k in range(runs + 1):
            t0 = time.perf_counter()
            O.Run(O.KIND_VORTEX, cb[:, :3], cb, p=args.pmax, leaf_size_source=args.leaf,
                  multipole_acceptance=args.mac, error_kind=O.ERR_POWER_GRADIENT, relative=True, eps=args.tol,
                  gradient=True, hessian=True).fmm()
            if k:
                ts.append(time.perf_counter() - t0)
        med = float(np.median(ts))
        log(f"cpu baseline: {m} bodies on {threads} thread(s): median {med:.2f} s, min {min(ts):.2f} s over {runs}")
        return {"value": round(m / med, 1), "unit": "evals/s", "cores": threads, "n": m,
                "median_s": round(med, 3), "min_s": round(min(ts), 3), "runs": runs}

    full = leg(args.cpu_sample, cores, args.cpu_runs)
    one = leg(args.cpu_sample_1, 1, max(3, args.cpu_runs - 2)) if args.cpu_sample_1 > 0 else None
    O.threads(cores)
    out = dict(full)
    out.update({"kind": "port",
                "sample": f"oracle (C++ FP64 restatement, OpenMP) full fmm! on a {args.cpu_sample}-particle ring of the "
                          f"headline distribution and settings on {cores} host threads; warm-up, then the median of "
                          f"{args.cpu_runs} runs",
                "single_core": one})
    return out


class _RowSystem:
    """A system handed to the request server as source rows plus a built-in kernel."""

    def __init__(self, kernel, rows):
        self.fmm_kernel = kernel
        self.rows = np.ascontiguousarray(rows, dtype=np.float64)

    def source_buffer(self):
        return self.rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--tol", type=float, default=1e-9)
    ap.add_argument("--leaf", type=int, default=64)
    ap.add_argument("--mac", type=float, default=0.55)
    ap.add_argument("--pmax", type=int, default=20)
    ap.add_argument("--p2p-free-slots", type=int, default=None,
                    help="workgroup slots per CU the nearfield leaves to the far-field stream (library default 1)")
    ap.add_argument("--samples", type=int, default=10000)
    ap.add_argument("--cpu-sample", type=int, default=100000, help="ring size of the all-cores CPU baseline")
    ap.add_argument("--cpu-sample-1", type=int, default=30000, help="ring size of the one-core CPU baseline (0: skip)")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed CPU-baseline runs after one warm-up")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="mixed,fp64")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 256 MiB device to device three times (the FETCH_SIZE / WRITE_SIZE calibration)")
    ap.add_argument("--serve-requests", type=int, default=30)
    ap.add_argument("--serve-rates", default="6,inf")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-budget-gib", type=float, default=96.0)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-sweep", action="store_true", help="skip the 4M Plummer precision sweep")
    ap.add_argument("--sweep-n", type=int, default=4194304)
    ap.add_argument("--sweep-leaf", type=int, default=50)
    ap.add_argument("--sweep-mac", type=float, default=0.5)
    ap.add_argument("--sweep-steps", type=int, default=2)
    ap.add_argument("--sweep-forms", default="fp64,mixed")
    ap.add_argument("--batch-systems", type=int, default=512)
    ap.add_argument("--batch-slots", type=int, default=8)
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes": 3 * a.numel() * 4}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only (see DESIGN.md)"}))
        sys.exit(2)

    import fastmultipole_amd as F
    from fastmultipole_amd import _lib, workloads as W

    n = args.n
    sbuf = W.vortex_ring(n, seed=2)
    ctx = F.default_context()
    if args.p2p_free_slots is not None:
        ctx.set_p2p_free_slots(args.p2p_free_slots)
    ev = F.Evaluation(ctx)
    ev.set_systems(_lib.POINT_VORTEX, None, sbuf)          # host -> HBM once, outside the timed region
    em = F.PowerRelativeGradient(args.tol)

    # reference velocities on sampled targets: FP64 O(N^2) direct on the GPU
    rng = np.random.default_rng(3)
    idx = np.sort(rng.choice(n, size=min(args.samples, n), replace=False))
    txyz = np.ascontiguousarray(sbuf[idx, :3])
    ref = np.zeros((len(idx), 16))
    ref[:, :3] = txyz
    t0 = time.perf_counter()
    ctx.check(ctx.lib.fmmhip_direct(ctx._h, _lib.POINT_VORTEX, len(idx), txyz.ctypes.data, n,
                                    np.ascontiguousarray(sbuf).ctypes.data, 0, 1, 1, _lib.FP64, ref.ctypes.data, 0))
    log(f"direct reference on {len(idx)} targets: {time.perf_counter() - t0:.2f} s")

    forms = {}
    for form in args.forms.split(","):
        opt = F.make_options(expansion_order=args.pmax, leaf_size_source=args.leaf, multipole_acceptance=args.mac,
                             error_tolerance=em, gradient=True, hessian=True, precision=form)
        for _ in range(args.warmup):
            ev.fmm(opt)
        ctx.check(ctx.lib.fmmhip_sync(ctx._h))
        t0 = time.perf_counter()
        stats, lat = [], []
        for _ in range(args.steps):
            t1 = time.perf_counter()
            stats.append(ev.fmm(opt).as_dict())     # returns after the evaluation's stream synchronisation
            lat.append(time.perf_counter() - t1)
        ctx.check(ctx.lib.fmmhip_sync(ctx._h))
        dt = (time.perf_counter() - t0) / args.steps
        out = ev.targets()
        orders = ev.m2l_orders()
        e_v = W.relative_error(out[idx, 4:7], ref[:, 4:7])
        e_g = W.relative_error(out[idx, 7:16], ref[:, 7:16])
        phases = {k: float(np.mean([s[k] for s in stats])) for k in
                  ("ms_tree", "ms_lists", "ms_estimate", "ms_p2p", "ms_upward", "ms_m2l", "ms_downward",
                   "ms_gather", "ms_total")}
        forms[form] = dict(ms=dt * 1e3, evals=n / dt, stats=stats[-1], phases=phases, err_v=e_v, err_grad=e_g,
                           orders=orders, lat=lat)
        log(f"{form}: {dt * 1e3:.2f} ms/step, {n / dt:.3e} evals/s; phases {phases}; "
            f"vel err max {e_v[0]:.2e} p99 {e_v[1]:.2e}; grad err max {e_g[0]:.2e} p99 {e_g[1]:.2e}; "
            f"pmax {stats[-1]['pmax']} m2l {stats[-1]['n_m2l']} pairs {stats[-1]['n_p2p_pairs']}")

    def meets_bar(v):   # velocity and its gradient, the workload's two outputs
        return v["err_v"][0] <= 1e-6 and v["err_grad"][0] <= 1e-6

    ok_forms = [k for k, v in forms.items() if meets_bar(v)]
    if ok_forms:
        main_form = max(ok_forms, key=lambda k: forms[k]["evals"])
    else:
        main_form = "mixed" if "mixed" in forms else next(iter(forms))
    M = forms[main_form]
    st = M["stats"]
    ph = M["phases"]
    # roofline of the form's dominant kernel, from the library's HIP-event phase timings on its stream
    p2p_tflops = st["n_p2p_pairs"] * FLOPS_PER_PAIR_VORTEX / (ph["ms_p2p"] * 1e-3) / 1e12
    hist = np.bincount(M["orders"], minlength=args.pmax + 1)
    m2l_total = float(sum(c * m2l_flops(p) for p, c in enumerate(hist)))
    m2l_tflops = m2l_total / (ph["ms_m2l"] * 1e-3) / 1e12
    key = "k_p2p_f32<1, false, true, true>" if main_form == "mixed" else "k_p2p_f64<1, false, true, true>"
    roof_p2p = {"kernel": key + " (nearfield P2P, vortex v + grad v, 63 flop/pair)",
                "bound": "vector-alu", "achieved": round(p2p_tflops, 3),
                "peak": PEAK_FP32_TFLOPS if main_form == "mixed" else PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "traffic": pmc_traffic(key)}
    # the error-controlled M2L: one persistent launch of the matrix-core engine at the order cap (the
    # smallest instantiated cap >= pmax, specialised for order == cap); achieved = F(p) flops summed over
    # the interactions at the orders they used / the M2L phase time (the launch and the per-box ordering
    # sort before it)
    cap = next(c for c in (8, 12, 16, 20) if c >= args.pmax)
    m2l_key = f"k_m2l_mx<{cap}, true, 4, {'true' if cap == args.pmax else 'false'}>"
    roof_m2l = {"kernel": m2l_key + " (rotation-based M2L with Lamb-Helmholtz terms on v_mfma_f64_16x16x4f64, FP64; "
                                    "F(p) flop/interaction)",
                "bound": "mfma", "achieved": round(m2l_tflops, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "traffic": pmc_traffic(m2l_key)}
    roofs = [roof_p2p, roof_m2l] if ph["ms_p2p"] >= ph["ms_m2l"] else [roof_m2l, roof_p2p]
    for r in roofs:
        r["traffic_unit"] = "bytes/launch (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE, profiles/ring1m_traffic_pmc.md)"
        r["frac"] = round(r["achieved"] / r["peak"], 4)
    roof, roof_second = roofs

    # the same evaluation through host buffers, as a caller handing over host arrays pays it: the source buffer
    # copied in (pageable host memory), the evaluation, the 16 x N results copied out (never `value`)
    hev = F.Evaluation(ctx)
    hopt = F.make_options(expansion_order=args.pmax, leaf_size_source=args.leaf, multipole_acceptance=args.mac,
                          error_tolerance=em, gradient=True, hessian=True, precision=main_form)
    hev.set_systems(_lib.POINT_VORTEX, None, sbuf)
    hev.fmm(hopt)
    hev.targets()
    h_in, h_run, h_out = [], [], []
    for _ in range(3):
        t1 = time.perf_counter()
        hev.set_systems(_lib.POINT_VORTEX, None, sbuf)
        t2 = time.perf_counter()
        hev.fmm(hopt)
        t3 = time.perf_counter()
        hev.targets()
        t4 = time.perf_counter()
        h_in.append(t2 - t1); h_run.append(t3 - t2); h_out.append(t4 - t3)
    hev.close()
    h_tot = float(np.median(np.array(h_in) + np.array(h_run) + np.array(h_out)))
    host_buffers = {"ms_per_eval": round(h_tot * 1e3, 3), "evals_per_s": round(n / h_tot, 1),
                    "copy_in_ms": round(float(np.median(h_in)) * 1e3, 3),
                    "copy_out_ms": round(float(np.median(h_out)) * 1e3, 3),
                    "note": f"{main_form} form; source buffer ({sbuf.nbytes / 1e6:.0f} MB) in and 16 x N results "
                            f"({16 * 8 * n / 1e6:.0f} MB) out through pageable host memory; median of 3"}
    log(f"host buffers: {host_buffers}")

    sweep = None if args.no_sweep else precision_sweep(F, W, _lib, args)
    serving = None if args.no_serve else serving_sample(F, W, _lib, args)
    batch = None if args.no_batch else batch_sample(F, W, _lib, args)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(W, args)

    line = {
        "metric": "FMM evals/s at rel.err<=1e-6 (mixed vs FP64 GPU) + p50/p99 request latency",
        "value": round(M["evals"], 1),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(M["ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 P2P + fp64 expansions" if main_form == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"vortex ring N={n} (R=1, Gaussian core sigma=0.1), velocity+gradient, "
                               f"PowerRelativeGradient({args.tol}), p<={args.pmax}, leaf {args.leaf}, MAC {args.mac}, "
                               "shrink_recenter, SelfTuningTreeStop"},
        "roofline": roof,
        "roofline_second": roof_second,
        "cpu_baseline": cpu,
        "forms": {k: {"evals_per_s": round(v["evals"], 1), "ms_per_step": round(v["ms"], 3),
                      "rel_err_velocity": {"max": v["err_v"][0], "p99": v["err_v"][1], "median": v["err_v"][2]},
                      "rel_err_gradient": {"max": v["err_grad"][0], "p99": v["err_grad"][1]},
                      "meets_error_bar": bool(meets_bar(v)),
                      "phases_ms": {k2: round(x, 3) for k2, x in v["phases"].items()},
                      "pmax": v["stats"]["pmax"], "n_m2l": v["stats"]["n_m2l"],
                      "p2p_pairs_per_s": v["stats"]["n_p2p_pairs"] / (v["phases"]["ms_p2p"] * 1e-3),
                      "m2l_per_s": v["stats"]["n_m2l"] / (v["phases"]["ms_m2l"] * 1e-3)}
                  for k, v in forms.items()},
        "mixed_over_fp64": (round(forms["mixed"]["evals"] / forms["fp64"]["evals"], 3)
                            if "mixed" in forms and "fp64" in forms else None),
        "form": main_form,
        "meets_error_bar": bool(meets_bar(M)),
        "host_buffers": host_buffers,
        "latency_ms": {"p50": round(float(np.percentile(M["lat"], 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(M["lat"], 99)) * 1e3, 3),
                       "note": "the headline workload's evaluations issued back to back, one at a time"},
        "precision_sweep": sweep,
        "serving": serving,
        "batch": batch,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
