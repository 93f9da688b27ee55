"""Benchmark of the direct-sum softened gravity step on one MI355X.

A step is one full gravity evaluation of the giant-star workload with all
inputs already resident in HBM: validation, Morton sort, gather, tile boxes,
far-field pair kernel, near-field kernel and the FP64 reduction, i.e. every
pair (i, j) of N = 1,048,576 particles (N^2 ordered pair-interactions,
including i == j).  The FP32+compensated kernel is the product; the plain FP64
kernel is timed the same way beside it.

Prints ONE JSON line (rank 0, N = 1 GPU only).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("G pair-interactions/s at N=1M: FP32+comp vs FP64 GPU, max rel acc err ≤1e-5")
FLOPS_PER_PAIR = 19           # 3 dr + 5 r^2 + 1 rsqrt + 3 (r^-2, m r^-1, m r^-3) + 6 acc + 1 pot
PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 vector (AMD spec)
PEAK_FP64_TFLOPS = 78.6


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
    visits, nearv = g.last_tile_stats()
    res = [o.download() for o in out]
    return {
        "ms_per_step": (t1 - t0) * 1e3 / steps,
        "far_ms": float(np.mean(far_ms)), "near_ms": float(np.mean(near_ms)),
        "prep_ms": float(np.mean(prep_ms)), "visits": visits, "near_visits": nearv,
        "result": res,
    }


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
    for a in ins + cot + out:
        a.free()
    return res


def rel_err(a, b):
    ea = np.linalg.norm(np.stack(a[:3], -1) - np.stack(b[:3], -1), axis=1) / \
        np.linalg.norm(np.stack(b[:3], -1), axis=1)
    ep = np.abs(a[3] - b[3]) / np.abs(b[3])
    return ea, ep


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


def cpu_baseline(d, seconds, nthreads):
    """The oracle's C restatement of the reference's symmetric CPU gravity
    loop (cpu_grav.f:1-381) on a bounded row sample of the same workload."""
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
                      f"({ev} evaluations, {dt:.1f} s)"}


def bench_batch(args):
    """64 independent stars of log-uniform N in [2048, 131072] (even entries
    Plummer spheres, odd entries n = 1.5 polytropes) in one persistent,
    graph-captured launch, against 64 separate single-set launches."""
    from starsmasher_amd import SG_PREC_FP32C, SG_PREC_FP64, Gravity, ics
    from starsmasher_amd.autotune import tune_batch
    from starsmasher_amd.hipmem import DeviceArray, runtime

    log("bench: generating 64-star batch")
    stars = ics.batch_set(64)
    sizes = np.array([s["x"].size for s in stars], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    ntot = int(off[-1])
    pairs = float((sizes.astype(np.float64) ** 2).sum())
    g = Gravity(0)
    dev = [DeviceArray.from_host(np.concatenate([s[k] for s in stars]))
           for k in ("x", "y", "z", "m", "range2")]
    out = [DeviceArray(ntot) for _ in range(4)]
    tuning = {}
    if args.tune:
        for prec in (SG_PREC_FP32C, SG_PREC_FP64):
            tr = tune_batch(g, off, dev, out, prec, reps=1, log=log)
            tuning["fp32c" if prec == SG_PREC_FP32C else "fp64"] = {
                "best": list(tr.best), "table_ms": [[r[0], r[1], round(r[3], 2)] for r in tr.table]}

    def run_batch(prec, steps, graph=True):
        g.compute_batch_device(off, *dev, *out, precision=prec, use_graph=graph)
        g.synchronize()
        runtime().hipDeviceSynchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.compute_batch_device(off, *dev, *out, precision=prec, use_graph=graph)
            g.synchronize()
        runtime().hipDeviceSynchronize()
        return (time.perf_counter() - t0) * 1e3 / steps, [o.download() for o in out]

    def run_separate(prec, steps):
        for _ in range(steps + 1):
            if _ == 1:
                runtime().hipDeviceSynchronize()
                t0 = time.perf_counter()
            for s in range(64):
                a, b = int(off[s]), int(off[s + 1])
                g.compute_device(b - a, *[x.ptr + 8 * a for x in dev], *[o.ptr + 8 * a for o in out],
                                 precision=prec)
            g.synchronize()
        runtime().hipDeviceSynchronize()
        return (time.perf_counter() - t0) * 1e3 / steps, [o.download() for o in out]

    log("bench: batch FP32C")
    ms32, r32 = run_batch(SG_PREC_FP32C, args.steps)
    log(f"bench: batch FP32C {ms32:.2f} ms/step")
    ms32_eager, _ = run_batch(SG_PREC_FP32C, args.steps, graph=False)
    ms32_sep, s32 = run_separate(SG_PREC_FP32C, args.steps)
    log(f"bench: 64 separate launches FP32C {ms32_sep:.2f} ms/step")
    ms64, r64 = run_batch(SG_PREC_FP64, max(1, args.steps // 2))
    ms64_sep, _ = run_separate(SG_PREC_FP64, max(1, args.steps // 2))

    # per star: FP32C batch and FP32C separate launches against the FP64
    # batch (the FP64 kernels are checked against the oracle in the GPU tests)
    per_star = []
    for s in range(64):
        a, b = int(off[s]), int(off[s + 1])
        ea, ep = rel_err([x[a:b] for x in r32], [x[a:b] for x in r64])
        es, _ = rel_err([x[a:b] for x in s32], [x[a:b] for x in r64])
        per_star.append({"n": b - a, "acc_max_vs_fp64": float(ea.max()),
                         "acc_median_vs_fp64": float(np.median(ea)),
                         "pot_max_vs_fp64": float(ep.max()),
                         "separate_acc_max_vs_fp64": float(es.max())})
    line = {
        "metric": "G pair-interactions/s, 64 independent stars in one graph-captured persistent "
                  "launch (FP32+comp)",
        "value": round(pairs / (ms32 * 1e-3) / 1e9, 3), "unit": "Gpairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": round(ms32, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "64 stars, N log-uniform in [2048, 131072] (seed 4), Plummer / "
                               "n=1.5 polytrope alternating, nkernel=0",
                   "n_total": ntot, "sum_n2": pairs,
                   "tile": dict(zip(("wg", "ipl"), g.get_batch_config(SG_PREC_FP32C)))},
        "batch_eager_ms": round(ms32_eager, 3),
        "separate_launches_ms": round(ms32_sep, 3),
        "batch_speedup_vs_separate": round(ms32_sep / ms32, 3),
        "fp64_batch_ms": round(ms64, 3), "fp64_separate_ms": round(ms64_sep, 3),
        "fp32c_over_fp64": round(ms64 / ms32, 3),
        "per_star_worst": {k: max(p[k] for p in per_star) for k in per_star[0] if k != "n"},
        "per_star": per_star,
        "autotune": tuning or None,
    }
    g.close()
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="giant1m")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
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

    from starsmasher_amd import SG_PREC_FP32C, SG_PREC_FP64, Gravity
    from starsmasher_amd.hipmem import DeviceArray

    if args.workload == "batch64":
        return bench_batch(args)
    log(f"bench: generating workload {args.workload}")
    d, desc = make_workload(args.workload)
    n = d["x"].size
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
        for prec in (SG_PREC_FP32C,) + (() if args.no_fp64 else (SG_PREC_FP64,)):
            tr = tune(g, n, dev, outs, prec, reps=1, log=log)
            tuning["fp32c" if prec == SG_PREC_FP32C else "fp64"] = {
                "best": list(tr.best), "table_ms": [[r[0], r[1], r[2], round(r[3], 2)]
                                                     for r in tr.table]}
        del outs
    wg, ipl, js = g.get_config(SG_PREC_FP32C)

    log(f"bench: FP32C n={n} steps={args.steps} warmup={args.warmup}")
    r32 = time_precision(g, dev, n, SG_PREC_FP32C, args.steps, args.warmup)
    log(f"bench: FP32C {r32['ms_per_step']:.2f} ms/step (far {r32['far_ms']:.2f} ms)")
    r64 = None
    if not args.no_fp64:
        log("bench: FP64")
        r64 = time_precision(g, dev, n, SG_PREC_FP64, max(1, args.steps // 2), 1)
        log(f"bench: FP64 {r64['ms_per_step']:.2f} ms/step (far {r64['far_ms']:.2f} ms)")

    pairs = float(n) * float(n)
    value = pairs / (r32["ms_per_step"] * 1e-3) / 1e9
    near_frac = r32["near_visits"] / max(1, r32["visits"])
    padded = -(-n // 256) * 256
    # the FP32C pass is one kernel over every (i, padded j) pair: far chunks
    # and near chunks alike, each counted at the same 19 flops
    kernel_pairs = float(n) * padded
    achieved = FLOPS_PER_PAIR * kernel_pairs / (r32["far_ms"] * 1e-3) / 1e12

    acc = {}
    if r64 is not None:
        ea, ep = rel_err(r32["result"], r64["result"])
        acc["fp32c_vs_fp64_kernel_all"] = {
            "acc_max": float(ea.max()), "acc_median": float(np.median(ea)),
            "pot_max": float(ep.max()), "pot_median": float(np.median(ep))}
    cpu = cpu1 = None
    if not args.no_cpu:
        nth = min(16, os.cpu_count() or 1)
        log(f"bench: CPU baseline on {nth} threads")
        cpu = cpu_baseline(d, args.cpu_seconds, nth)
        log("bench: CPU baseline on 1 thread")
        cpu1 = cpu_baseline(d, args.cpu_seconds / 3, 1)
        try:
            cpu["cpu_model"] = next(ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo")
                                    if ln.startswith("model name"))
        except (OSError, StopIteration):
            pass

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
        "accuracy": acc,
        "cpu_baseline": cpu,
        "cpu_baseline_1core": cpu1,
        "autotune": tuning or None,
        "adjoint": adj,
    }
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
