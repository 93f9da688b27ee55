"""Benchmark of the direct-sum softened gravity step on one MI355X.

A step is one full gravity evaluation of the giant-star workload with all
inputs already resident in HBM: validation, Morton sort, gather, tile boxes,
the pair kernel (far and near chunks) and the FP64 reduction, i.e. every pair
(i, j) of N = 1,048,576 particles (N^2 ordered pair-interactions, including
i == j).  The FP32+compensated kernel is the product; the plain FP64 kernel
is timed the same way beside it.

Beside the headline workload the one JSON line carries:
  * the CPU baselines, timed first on every core the process may use: the
    oracle's C restatement of the reference's CPU gravity loop on a row
    sample of the same workload, and the reference's own Fortran
    (cpu_grav.f, built from its source into oracle/_ref) as parallel
    processes on a row sample of the two-polytrope set (the giant star
    exceeds its nmax);
  * the end-to-end time through the reference's Fortran entry points on host
    arrays (pack, host-to-device copy, kernels, device-to-host copy: the span
    of the reference's own timer);
  * the Plummer (16,384) and two-polytrope (262,144) sets, the 64-star batch
    in one graph-captured launch against 64 separate launches, and the
    adjoint on the 131,072-particle Plummer sphere;
  * per workload, the share of pairs whose arithmetic involves FP64.

Prints ONE JSON line (rank 0, N = 1 GPU only).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import multiprocessing as mp
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("G pair-interactions/s at N=1M: FP32+comp vs FP64 GPU, max rel acc err ≤1e-5")
FLOPS_PER_PAIR = 19           # 3 dr + 5 r^2 + 1 rsqrt + 3 (r^-2, m r^-1, m r^-3) + 6 acc + 1 pot
PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 vector (AMD spec)
PEAK_FP64_TFLOPS = 78.6
TC = 64                       # j per chunk (the unit of the near / wide decisions)

# The FP64 steps of the precision specification (FP32 pair arithmetic with
# hi/lo positions and compensated sums, FP64 only for these), and the FP64
# this library adds beyond them.
PRECISION_POLICY = {
    "specified_fp64": ["origin subtraction", "cross-tile final reduction",
                       "total-energy diagnostic"],
    "added_fp64": [
        "wide chunks: pair differences formed in FP64 and rounded once (chunks close to a wave "
        "relative to its spread about the workgroup origin)",
        "near chunks: pair differences formed in FP64 and rounded once, FP32 body",
        "softened and self pairs (r < max(H_i, H_j)): whole pair in FP64",
        "FP32 partial sums folded into FP64 accumulators every 32 j (in place of per-term "
        "compensated FP32 sums)",
    ],
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_workload(name):
    from starsmasher_amd import ics
    if name == "giant1m":
        d = ics.giant_with_core(1 << 20)
        desc = ("giant star N=1048576: n=3 polytrope core (60%) + r^-2.5 envelope to r=300, "
                "masses over 4 decades, hp over 3 decades, one point mass of 30% total mass "
                "(hp=0.1, range2=-0.04); cubic spline (nkernel=0)")
    elif name == "twopoly262k":
        d = ics.two_polytropes(1 << 18)
        desc = "two n=1.5 polytropes N=262144 (mass ratio 0.6, radii 1.0/0.8, 3.0 apart); nkernel=0"
    elif name == "plummer16k":
        d = ics.plummer(16384)
        desc = "Plummer sphere N=16384 (a=1, r<10a, m=1/N); nkernel=0"
    elif name == "plummer131k":
        d = ics.adjoint_set(131072)
        desc = "Plummer sphere N=131072 (a=1, r<10a, seed 5, the adjoint's set); nkernel=0"
    else:
        raise SystemExit(f"unknown workload {name}")
    return d, desc


# ------------------------------------------------------------- host facts
def host_info():
    """Logical CPUs, the ones this process may run on, and the CPU model."""
    info = {"nproc": os.cpu_count(), "allowed_cpus": len(os.sched_getaffinity(0))}
    try:
        models, keys = set(), {}
        for ln in open("/proc/cpuinfo"):
            if ":" not in ln:
                continue
            k, v = (s.strip() for s in ln.split(":", 1))
            if k == "model name":
                models.add(v)
            elif k in ("cpu family", "model", "stepping", "cpu MHz", "cache size") and k not in keys:
                keys[k] = v
        # the numeric "model" field is reported as cpu_model_number, so that
        # it does not overwrite the model name
        info.update({f"cpu_{k.replace(' ', '_').lower()}" + ("_number" if k == "model" else ""): v
                     for k, v in keys.items()})
        info["cpu_model"] = "; ".join(sorted(models))
    except OSError:
        pass
    return info


def cpu_cores():
    return max(1, len(os.sched_getaffinity(0)))


# --------------------------------------------------------- CPU baselines
def cpu_baseline(d, seconds, nthreads):
    """The oracle's C restatement of the reference's symmetric CPU gravity
    loop (cpu_grav.f:1-381) on a bounded row sample of the same workload,
    OpenMP threads over interaction-balanced row chunks (main.f:297-349)."""
    import oracle
    args = (d["x"], d["y"], d["z"], d["m"], d["hp"], d["u"])
    n = d["x"].size

    def run(rows):
        t = time.perf_counter()
        if nthreads == 1:
            ev, *_ = oracle.cpu_twin_rows(*args, 0, rows - 1, 0)
        else:
            ev, *_ = oracle.cpu_twin_rows_parallel(*args, 0, rows - 1, 0, nthreads)
        return ev, time.perf_counter() - t

    # two-point calibration: the parallel form has an O(threads N) set-up
    r1, r2 = 2 * nthreads, 8 * nthreads
    _, t1 = run(r1)
    _, t2 = run(r2)
    per_row = max((t2 - t1) / (r2 - r1), 1e-6)
    rows = int(max(r2, min(n - 1, r1 + (seconds - t1) / per_row)))
    ev, dt = run(rows)
    # each symmetric evaluation covers two ordered pair-interactions (i<-j, j<-i)
    return {"value": 2.0 * ev / dt / 1e9, "unit": "Gpairs/s", "cores": nthreads, "kind": "port",
            "sample": f"rows 1..{rows} of the symmetric j>=i loop over all {n} particles "
                      f"({ev} evaluations, {dt:.1f} s, {nthreads} OpenMP threads)"}


def _ref_worker(path, lo, hi, barrier, queue):
    """One process of the reference-CPU baseline: the reference's own
    get_gravity_using_cpus over rows lo..hi (1-based) of the set in `path`,
    as one MPI rank of the reference would run it."""
    import oracle
    a = np.load(path)
    args = tuple(np.ascontiguousarray(a[k]) for k in range(6))
    n = args[0].size
    oracle.ref_cpu_gravity(*args, ngrav_lower=n, ngrav_upper=n)   # load, touch the arrays
    barrier.wait()
    t = time.perf_counter()
    oracle.ref_cpu_gravity(*args, ngrav_lower=lo, ngrav_upper=hi)
    queue.put(time.perf_counter() - t)


def cpu_baseline_reference(seconds, nproc):
    """The reference's own Fortran CPU gravity (cpu_grav.f, compiled from its
    source into oracle/_ref) on the two-polytrope set (N = 262,144 fits its
    nmax of 900,000): `nproc` processes, each one rank's interaction-balanced
    share (main.f:297-349) of the first rows of the symmetric j >= i loop,
    started together; rate = ordered pairs / slowest process."""
    import oracle
    if not oracle.ref_available():
        return None
    from starsmasher_amd import ics
    d = ics.two_polytropes(1 << 18)
    n = d["x"].size
    ctx = mp.get_context("spawn")
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "set.npy")
        np.save(path, np.stack([d[k] for k in ("x", "y", "z", "m", "hp", "u")]))

        def run(bounds):
            barrier = ctx.Barrier(len(bounds))
            q = ctx.Queue()
            ps = [ctx.Process(target=_ref_worker, args=(path, lo, hi, barrier, q))
                  for lo, hi in bounds]
            for p in ps:
                p.start()
            times = [q.get(timeout=600) for _ in ps]
            for p in ps:
                p.join()
            return max(times)

        def evals(lo, hi):   # sum over rows i = lo..hi of (n - i + 1)
            return (hi - lo + 1) * (n + 1) - (lo + hi) * (hi - lo + 1) // 2

        # calibrate the one-process rate on a short sample
        r0 = 128
        t0 = run([(1, r0)])
        rate1 = evals(1, r0) / t0
        # rows 1..R with about `seconds` of work per process, split evenly in evaluations
        total = rate1 * seconds * nproc
        cum = np.cumsum(n - np.arange(1, n + 1) + 1, dtype=np.float64)
        R = int(min(n, np.searchsorted(cum, total) + 1))
        edges = [0] + [int(np.searchsorted(cum, cum[R - 1] * p / nproc)) + 1
                       for p in range(1, nproc)] + [R]
        bounds = [(edges[p] + 1, edges[p + 1]) for p in range(nproc) if edges[p + 1] > edges[p]]
        t = run(bounds)
        ev = evals(1, R)
    return {"value": 2.0 * ev / t / 1e9, "unit": "Gpairs/s", "cores": len(bounds),
            "kind": "reference",
            "sample": f"reference cpu_grav.f (get_gravity_using_cpus) as {len(bounds)} processes, "
                      f"rows 1..{R} of the symmetric j>=i loop over the two-polytrope set "
                      f"N={n} ({ev} evaluations, slowest process {t:.1f} s)",
            "one_core": {"value": 2.0 * rate1 / 1e9, "unit": "Gpairs/s",
                         "sample": f"rows 1..{r0}, {t0:.2f} s"}}


# ------------------------------------------------------------- GPU legs
def time_precision(g, dev, n, prec, steps, warmup):
    from starsmasher_amd.hipmem import DeviceArray, runtime
    out = [DeviceArray(n) for _ in range(4)]
    for _ in range(warmup):
        g.compute_device(n, *dev, *out, nkernel=0, precision=prec)
        g.synchronize()
    runtime().hipDeviceSynchronize()
    far_ms, near_ms, prep_ms = [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        g.compute_device(n, *dev, *out, nkernel=0, precision=prec)
        g.synchronize()
        p, f, nr = g.last_timings()
        prep_ms.append(p)
        far_ms.append(f)
        near_ms.append(nr)
    runtime().hipDeviceSynchronize()
    t1 = time.perf_counter()
    stats = g.last_precision_stats()
    res = [o.download() for o in out]
    for o in out:
        o.free()
    return {
        "ms_per_step": (t1 - t0) * 1e3 / steps,
        "far_ms": float(np.mean(far_ms)), "near_ms": float(np.mean(near_ms)),
        "prep_ms": float(np.mean(prep_ms)), "stats": stats, "result": res,
    }


def precision_shares(stats, n, wi):
    """Shares of the ordered pairs by the arithmetic that forms them, from
    the kernel's counters (sg_last_precision_stats): a chunk visit covers 64
    j for each of the wave's wi i."""
    visits, nearv, widev, fp64 = stats
    pairs = float(visits) * TC * wi
    return {"near_chunk_fraction": round(nearv / max(1, visits), 6),
            "wide_pair_share": round(widev * TC * wi / max(1.0, pairs), 6),
            "near_chunk_pair_share": round(nearv * TC * wi / max(1.0, pairs), 6),
            "fp64_difference_share": round((widev + nearv) * TC * wi / max(1.0, pairs), 6),
            "fp64_pair_share": round(fp64 / float(n) ** 2, 8)}


def rel_err(a, b):
    ea = np.linalg.norm(np.stack(a[:3], -1) - np.stack(b[:3], -1), axis=1) / \
        np.linalg.norm(np.stack(b[:3], -1), axis=1)
    ep = np.abs(a[3] - b[3]) / np.abs(b[3])
    return ea, ep


def err_summary(a, b):
    ea, ep = rel_err(a, b)
    return {"acc_max": float(ea.max()), "acc_median": float(np.median(ea)),
            "pot_max": float(ep.max()), "pot_median": float(np.median(ep))}


def time_config(g, d, desc, steps):
    """One workload in both precisions: FP32C at the library's default
    (automatic) tile shape, the FP64 comparison kernel at its fastest
    measured shape; FP32C and FP64 steps (inputs in HBM), FP32C against the
    FP64 kernel on every particle."""
    from starsmasher_amd import SG_PREC_FP32C, SG_PREC_FP64
    from starsmasher_amd.autotune import tune
    from starsmasher_amd.hipmem import DeviceArray
    n = d["x"].size
    saved = {p: g.get_config(p) for p in (SG_PREC_FP32C, SG_PREC_FP64)}
    for p in saved:
        g.set_config(p, 0, 1, 0)
    dev = [DeviceArray.from_host(d[k]) for k in ("x", "y", "z", "m", "range2")]
    r32 = time_precision(g, dev, n, SG_PREC_FP32C, steps, 2)
    wg, ipl, js = g.last_config()
    outs = [DeviceArray(n) for _ in range(4)]
    tune(g, n, dev, outs, SG_PREC_FP64, reps=1, margin=0.0)
    for o in outs:
        o.free()
    r64 = time_precision(g, dev, n, SG_PREC_FP64, max(2, steps // 2), 1)
    tile64 = dict(zip(("wg", "ipl", "jsplit"), g.last_config()))
    for a in dev:
        a.free()
    for p, c in saved.items():
        g.set_config(p, *c)
    out = {"workload": desc, "n": n, "tile": {"wg": wg, "ipl": ipl, "jsplit": js},
           "fp32c": {"ms_per_step": round(r32["ms_per_step"], 4),
                     "kernel_ms": round(r32["far_ms"], 4),
                     "gpairs_per_s": round(float(n) ** 2 / (r32["ms_per_step"] * 1e-3) / 1e9, 3)},
           "fp64": {"ms_per_step": round(r64["ms_per_step"], 4), "tile": tile64,
                    "gpairs_per_s": round(float(n) ** 2 / (r64["ms_per_step"] * 1e-3) / 1e9, 3)},
           "fp32c_over_fp64": round(r64["ms_per_step"] / r32["ms_per_step"], 3),
           "fp32c_vs_fp64_kernel_all": err_summary(r32["result"], r64["result"]),
           "precision": precision_shares(r32["stats"], n, 64 * ipl)}
    log(f"bench: {desc[:40]}: FP32C {r32['ms_per_step']:.3f} ms, FP64 {r64['ms_per_step']:.3f} ms")
    return out


def time_end_to_end(d, steps):
    """The reference's boundary on host arrays: firsthalf_grav_forces_ then
    lasthalf_grav_forces_ (by reference, 1-based i-range 1..n), timed on the
    host around the pair of calls.  Covers the host argument checks, the
    host-to-device copies of x, y, z, m and range2 (pageable memory, as the
    reference's PINNED=false), preparation, kernels and the device-to-host
    copies of the four outputs: the span of the reference's own timer
    (grav_force_direct.cu:546-548, 571-576, 638)."""
    from starsmasher_amd import _lib
    L = _lib.load()
    n = d["x"].size
    arrs = [np.ascontiguousarray(d[k]) for k in ("x", "y", "z", "m", "range2")]
    outs = [np.empty(n) for _ in range(4)]
    ip = lambda v: C.byref(C.c_int(v))
    dp = C.POINTER(C.c_double)
    os.environ["SPHGRAV_QUIET"] = "1"
    os.environ.pop("SPHGRAV_PRECISION", None)

    def one():
        L.firsthalf_grav_forces_(ip(n), ip(1), ip(n), *[a.ctypes.data_as(dp) for a in arrs], ip(1),
                                 ip(0))
        L.lasthalf_grav_forces_(ip(n), *[o.ctypes.data_as(dp) for o in outs], ip(0))
        if L.sg_last_status() != 0:
            raise RuntimeError("legacy entry points failed")

    one()                                    # device buffers of the legacy layer
    ts = []
    for _ in range(steps):
        t = time.perf_counter()
        one()
        ts.append(time.perf_counter() - t)
    ms = float(np.mean(ts)) * 1e3
    return {"ms_per_call": round(ms, 3), "calls": steps,
            "gpairs_per_s": round(float(n) ** 2 / (ms * 1e-3) / 1e9, 3),
            "bytes_in": 40 * n, "bytes_out": 32 * n,
            "span": "firsthalf_grav_forces_ + lasthalf_grav_forces_ on host arrays: checks, "
                    "H2D (pageable), prep, kernels, D2H"}, outs


def bench_batch_leg(g, steps, tune, log_fn):
    """64 independent stars of log-uniform N in [2048, 131072] (even entries
    Plummer spheres, odd entries n = 1.5 polytropes) in one persistent,
    graph-captured launch, against 64 separate single-set launches, with the
    worst per-star error against the FP64 batch."""
    from starsmasher_amd import SG_PREC_FP32C, SG_PREC_FP64, ics
    from starsmasher_amd.autotune import tune_batch
    from starsmasher_amd.hipmem import DeviceArray, runtime

    stars = ics.batch_set(64)
    sizes = np.array([s["x"].size for s in stars], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    ntot = int(off[-1])
    pairs = float((sizes.astype(np.float64) ** 2).sum())
    dev = [DeviceArray.from_host(np.concatenate([s[k] for s in stars]))
           for k in ("x", "y", "z", "m", "range2")]
    out = [DeviceArray(ntot) for _ in range(4)]
    tuning = tuning64 = None
    if tune:
        tr = tune_batch(g, off, dev, out, SG_PREC_FP32C, reps=1, log=log_fn)
        tuning = {"best": list(tr.best), "table_ms": [[r[0], r[1], round(r[3], 2)] for r in tr.table]}
        # the FP64 comparison batch at its fastest measured shape (no margin)
        tr = tune_batch(g, off, dev, out, SG_PREC_FP64, reps=1, margin=0.0, log=log_fn)
        tuning64 = {"best": list(tr.best),
                    "table_ms": [[r[0], r[1], round(r[3], 2)] for r in tr.table]}

    def run_batch(prec, k, graph=True):
        g.compute_batch_device(off, *dev, *out, precision=prec, use_graph=graph)
        g.synchronize()
        runtime().hipDeviceSynchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            g.compute_batch_device(off, *dev, *out, precision=prec, use_graph=graph)
            g.synchronize()
        runtime().hipDeviceSynchronize()
        ms = (time.perf_counter() - t0) * 1e3 / k
        return ms, g.last_precision_stats(), [o.download() for o in out]

    def run_separate(prec, k):
        for rep in range(k + 1):
            if rep == 1:
                runtime().hipDeviceSynchronize()
                t0 = time.perf_counter()
            for s in range(64):
                a, b = int(off[s]), int(off[s + 1])
                g.compute_device(b - a, *[x.ptr + 8 * a for x in dev],
                                 *[o.ptr + 8 * a for o in out], precision=prec)
            g.synchronize()
        runtime().hipDeviceSynchronize()
        return (time.perf_counter() - t0) * 1e3 / k, [o.download() for o in out]

    ms32, st32, r32 = run_batch(SG_PREC_FP32C, steps)
    log_fn(f"bench: batch FP32C {ms32:.2f} ms/step")
    ms32_sep, s32 = run_separate(SG_PREC_FP32C, max(1, steps // 2))
    log_fn(f"bench: 64 separate launches FP32C {ms32_sep:.2f} ms/step")
    ms64, _, r64 = run_batch(SG_PREC_FP64, 2)
    worst = {"acc_max": 0.0, "acc_median": 0.0, "pot_max": 0.0, "separate_acc_max": 0.0}
    for s in range(64):
        a, b = int(off[s]), int(off[s + 1])
        e = err_summary([x[a:b] for x in r32], [x[a:b] for x in r64])
        es, _ = rel_err([x[a:b] for x in s32], [x[a:b] for x in r64])
        worst["acc_max"] = max(worst["acc_max"], e["acc_max"])
        worst["acc_median"] = max(worst["acc_median"], e["acc_median"])
        worst["pot_max"] = max(worst["pot_max"], e["pot_max"])
        worst["separate_acc_max"] = max(worst["separate_acc_max"], float(es.max()))
    wg, ipl = g.get_batch_config(SG_PREC_FP32C)
    wg64, ipl64 = g.get_batch_config(SG_PREC_FP64)
    for a in dev + out:
        a.free()
    return {"workload": "64 stars, N log-uniform in [2048, 131072] (seed 4), Plummer / n=1.5 "
                        "polytrope alternating, nkernel=0",
            "n_total": ntot, "sum_n2": pairs, "tile": {"wg": wg, "ipl": ipl},
            "graph_ms_per_step": round(ms32, 3),
            "gpairs_per_s": round(pairs / (ms32 * 1e-3) / 1e9, 3),
            "separate_launches_ms": round(ms32_sep, 3),
            "batch_speedup_vs_separate": round(ms32_sep / ms32, 3),
            "fp64_batch_ms": round(ms64, 3), "fp64_tile": {"wg": wg64, "ipl": ipl64},
            "fp32c_over_fp64": round(ms64 / ms32, 3),
            "per_star_worst_vs_fp64_batch": worst,
            "precision": precision_shares(st32, 1, 64 * ipl) | {
                "fp64_pair_share": round(st32[3] / pairs, 8)},
            "autotune": tuning, "autotune_fp64": tuning64}


def time_adjoint(g, n_adj, steps, log_fn):
    """The adjoint on the differentiation workload: a seeded Plummer sphere of
    n_adj particles with FP32 unit-Gaussian cotangents (w, u).  Times the
    FP32C and FP64 vector-Jacobian products (inputs resident in HBM) and
    compares them on every particle."""
    from starsmasher_amd import SG_PREC_FP32C, SG_PREC_FP64, ics
    from starsmasher_amd.hipmem import DeviceArray, runtime
    d = ics.adjoint_set(n_adj, seed=5)
    ins = [DeviceArray.from_host(d[k]) for k in ("x", "y", "z", "m", "range2")]
    cot = [DeviceArray.from_host(np.ascontiguousarray(d["w"][:, c])) for c in range(3)]
    cot.append(DeviceArray.from_host(d["cot_u"]))
    out = [DeviceArray(n_adj) for _ in range(5)]
    res = {"n": n_adj, "workload": f"Plummer sphere N={n_adj} (a=1, r<10a, 64 neighbours "
                                   "within 2h), FP32 unit-Gaussian cotangents w, u; loss "
                                   "sum w.a + u.phi; gradients w.r.t. x, m, H"}
    grads = {}
    for prec, tag in ((SG_PREC_FP32C, "fp32c"), (SG_PREC_FP64, "fp64")):
        g.vjp_device(n_adj, *ins, *cot, *out, precision=prec)
        g.synchronize()
        runtime().hipDeviceSynchronize()
        far, near = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            g.vjp_device(n_adj, *ins, *cot, *out, precision=prec)
            g.synchronize()
            _, f, nr = g.last_timings()
            far.append(f)
            near.append(nr)
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1e3 / steps
        res[tag] = {"ms_per_step": round(ms, 3),
                    "gpairs_per_s": round(float(n_adj) ** 2 / (ms * 1e-3) / 1e9, 3),
                    "far_ms": round(float(np.mean(far)), 3),
                    "near_and_reduce_ms": round(float(np.mean(near)), 3)}
        grads[tag] = [o.download() for o in out]
        log_fn(f"bench: adjoint {tag} {ms:.2f} ms/step")
    res["fp32c_over_fp64"] = round(res["fp64"]["ms_per_step"] / res["fp32c"]["ms_per_step"], 3)
    # FP32C against the FP64 adjoint on every particle (the FP64 adjoint is
    # checked against the oracle's FP64 gradient in the GPU tests)
    errs = {}
    a, b = grads["fp32c"], grads["fp64"]
    for key, got, want in (("x", np.stack(a[:3], 1), np.stack(b[:3], 1)), ("m", a[3], b[3]),
                           ("H", a[4], b[4])):
        dd = np.abs(got - want) if got.ndim == 1 else np.linalg.norm(got - want, axis=1)
        mag = np.abs(want) if want.ndim == 1 else np.linalg.norm(want, axis=1)
        errs[key] = {"rel_median": float(np.median(dd / np.maximum(mag, 1e-300))),
                     "rms_normalised_max": float(dd.max() / np.sqrt(np.mean(mag ** 2)))}
    res["fp32c_vs_fp64_all"] = errs
    for x in ins + cot + out:
        x.free()
    return res


def profiled_traffic(wg, ipl):
    """HBM bytes per launch of the far kernel for this tile shape, from the
    committed rocprofv3 FETCH_SIZE + WRITE_SIZE summary (None if absent)."""
    path = os.path.join(ROOT, "profiles", "k_far_fp32c_giant1m.md")
    try:
        for ln in open(path):
            if ln.startswith("traffic_bytes_per_launch"):
                # traffic_bytes_per_launch = <bytes> (wg=<wg>, ipl=<ipl>)
                val = float(ln.split("=")[1].split()[0])
                if f"wg={wg}," in ln and f"ipl={ipl})" in ln:
                    return val
    except OSError:
        pass
    return None


def bench_batch(args):
    """--workload batch64: the batch leg alone, as its own JSON line."""
    from starsmasher_amd import Gravity
    g = Gravity(0)
    leg = bench_batch_leg(g, args.steps, args.tune, log)
    g.close()
    line = {
        "metric": "G pair-interactions/s, 64 independent stars in one graph-captured persistent "
                  "launch (FP32+comp)",
        "value": leg["gpairs_per_s"], "unit": "Gpairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": leg["graph_ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": leg["workload"], "n_total": leg["n_total"],
                   "sum_n2": leg["sum_n2"], "tile": leg["tile"]},
        "batch": leg,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="giant1m")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the headline workload (no end-to-end, C1/C2, batch legs)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--adjoint-n", type=int, default=131072,
                    help="particles of the adjoint workload (0 skips it)")
    ap.add_argument("--tune", action=argparse.BooleanOptionalAction, default=True,
                    help="autotune the tile shape of both kernels before timing")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="first run one device-to-device copy of BYTES (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against a known transfer)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(several GPUs would split the i-range with the j-set replicated on each; "
              "only the one-GPU share is built and measured)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch64":
        return bench_batch(args)
    log(f"bench: generating workload {args.workload}")
    d, desc = make_workload(args.workload)
    n = d["x"].size

    # CPU baselines first, on every core this process may use, before the
    # GPU runtime starts its threads in this process
    host = host_info()
    cpu = cpu1 = cpu_ref = None
    if not args.no_cpu:
        nth = cpu_cores()
        log(f"bench: CPU baseline (port) on {nth} threads")
        cpu = cpu_baseline(d, args.cpu_seconds, nth)
        cpu.update({"nproc": host["nproc"], "cpu_model": host.get("cpu_model")})
        log("bench: CPU baseline (port) on 1 thread")
        cpu1 = cpu_baseline(d, args.cpu_seconds / 4, 1)
        log(f"bench: CPU baseline (reference Fortran) on {nth} processes")
        cpu_ref = cpu_baseline_reference(args.cpu_seconds, nth)

    from starsmasher_amd import SG_PREC_FP32C, SG_PREC_FP64, Gravity
    from starsmasher_amd.hipmem import DeviceArray

    g = Gravity(0)
    # inputs resident in HBM before the timed region
    dev = [DeviceArray.from_host(d[k]) for k in ("x", "y", "z", "m", "range2")]
    if args.copy_probe > 0:
        src = DeviceArray(args.copy_probe // 8)
        dst = DeviceArray(args.copy_probe // 8)
        dst.copy_from(src)
        log(f"bench: copy probe of {args.copy_probe} bytes done")
        del src, dst
    tuning = {}
    if args.tune:
        from starsmasher_amd.autotune import tune
        outs = [DeviceArray(n) for _ in range(4)]
        # FP32C keeps its default shape unless another is 2% faster (the
        # profiled shape); the FP64 comparison kernel takes its fastest
        # measured shape, so the ratio is taken against its best configuration
        for prec in (SG_PREC_FP32C,) + (() if args.no_fp64 else (SG_PREC_FP64,)):
            tr = tune(g, n, dev, outs, prec, reps=1, log=log,
                      margin=0.02 if prec == SG_PREC_FP32C else 0.0)
            tuning["fp32c" if prec == SG_PREC_FP32C else "fp64"] = {
                "best": list(tr.best), "table_ms": [[r[0], r[1], r[2], round(r[3], 2)]
                                                     for r in tr.table]}
        for o in outs:
            o.free()
    log(f"bench: FP32C n={n} steps={args.steps} warmup={args.warmup}")
    r32 = time_precision(g, dev, n, SG_PREC_FP32C, args.steps, args.warmup)
    wg, ipl, js = g.last_config()   # the tile shape the timed steps ran with
    log(f"bench: FP32C {r32['ms_per_step']:.2f} ms/step (far {r32['far_ms']:.2f} ms)")
    r64 = None
    if not args.no_fp64:
        log("bench: FP64")
        r64 = time_precision(g, dev, n, SG_PREC_FP64, max(1, args.steps // 2), 1)
        log(f"bench: FP64 {r64['ms_per_step']:.2f} ms/step (far {r64['far_ms']:.2f} ms)")
    for a in dev:
        a.free()

    pairs = float(n) * float(n)
    value = pairs / (r32["ms_per_step"] * 1e-3) / 1e9
    visits, nearv = r32["stats"][:2]
    near_frac = nearv / max(1, visits)
    padded = -(-n // 256) * 256
    # the FP32C pass is one kernel over every (i, padded j) pair: far chunks
    # and near chunks alike, each counted at the same 19 flops
    kernel_pairs = float(n) * padded
    achieved = FLOPS_PER_PAIR * kernel_pairs / (r32["far_ms"] * 1e-3) / 1e12

    acc = {}
    if r64 is not None:
        acc["fp32c_vs_fp64_kernel_all"] = err_summary(r32["result"], r64["result"])

    extra = {}
    if not args.no_extra:
        log("bench: end-to-end through the reference's entry points (host arrays)")
        e2e, outs = time_end_to_end(d, 3)
        e2e["kernel_share"] = round(r32["far_ms"] / e2e["ms_per_call"], 4)
        e2e["vs_device_resident_step"] = round(e2e["ms_per_call"] / r32["ms_per_step"], 4)
        e2e["vs_device_resident_fp32c"] = err_summary(outs, r32["result"])
        extra["end_to_end"] = e2e
        log(f"bench: end-to-end {e2e['ms_per_call']:.2f} ms per call")
        small = {}
        for name, steps in (("plummer16k", 40), ("twopoly262k", 8)):
            dd, dsc = make_workload(name)
            small[name] = time_config(g, dd, dsc, steps)
        extra["configs"] = small
        log("bench: 64-star batch")
        extra["batch64"] = bench_batch_leg(g, max(3, args.steps // 2), args.tune, log)

    adj = None
    if args.adjoint_n > 0:
        log(f"bench: adjoint n={args.adjoint_n}")
        adj = time_adjoint(g, args.adjoint_n, max(2, args.steps), log)

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gpairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(r32["ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": desc, "n": n, "nkernel": 0, "tile": {"wg": wg, "ipl": ipl,
                   "jsplit": js}},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": (profiled_traffic(wg, ipl) if args.workload == "giant1m" else None),
                     "kernel": "k_far_fp32c",
                     "flops_per_pair": FLOPS_PER_PAIR,
                     "pairs_per_launch": kernel_pairs, "kernel_ms": round(r32["far_ms"], 3)},
        "phases_ms": {"prep": round(r32["prep_ms"], 3), "pair_kernel": round(r32["far_ms"], 3),
                      "reduce": round(r32["near_ms"], 3)},
        "near_chunk_fraction": round(near_frac, 5),
        "precision": precision_shares(r32["stats"], n, 64 * ipl),
        "precision_policy": PRECISION_POLICY,
        "accuracy": acc,
        "cpu_baseline": cpu,
        "cpu_baseline_1core": cpu1,
        "cpu_baseline_reference": cpu_ref,
        "host": host,
        "autotune": tuning or None,
        "adjoint": adj,
    }
    line.update(extra)
    if r64 is not None:
        v64 = pairs / (r64["ms_per_step"] * 1e-3) / 1e9
        # the FP64 far kernel skips the near chunks (a separate kernel takes them)
        a64 = FLOPS_PER_PAIR * kernel_pairs * (1.0 - near_frac) / (r64["far_ms"] * 1e-3) / 1e12
        line["fp64_kernel"] = {"value": round(v64, 3), "unit": "Gpairs/s",
                               "ms_per_step": round(r64["ms_per_step"], 3),
                               "far_kernel_ms": round(r64["far_ms"], 3),
                               "far_tflops": round(a64, 3),
                               "frac_of_fp64_peak": round(a64 / PEAK_FP64_TFLOPS, 4)}
        line["fp32c_over_fp64"] = round(value / v64, 3)
    g.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
