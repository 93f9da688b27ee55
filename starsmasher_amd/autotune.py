"""Tile-shape autotuner for the pair kernels.

For a precision and a particle set already resident on the device, times
every supported (workgroup size, i-particles per lane, j-partitions) shape of
the far kernel and keeps the fastest whole evaluation (prep + far + near +
reduce, HIP events).  The shape installed when tuning starts (for the
automatic workgroup size, the shape it resolves to on this set) stays unless
another is faster by more than ``margin`` (default 2%), so that run-to-run
noise does not flip between shapes of equal speed.  The chosen shape is
installed on the context with ``sg_set_config`` (an automatic configuration
that resolves to the pick is left automatic) and returned with the full
timing table.
"""
from __future__ import annotations

from dataclasses import dataclass, field

from ._lib import SG_PREC_FP32C, SG_PREC_FP64

CANDIDATES = {
    SG_PREC_FP32C: [(64, 1), (64, 2), (128, 1), (128, 2), (256, 1), (256, 2)],
    SG_PREC_FP64: [(64, 1), (64, 2), (128, 1), (128, 2), (256, 1), (256, 2)],
}


@dataclass
class TuneResult:
    precision: int
    best: tuple
    best_ms: float
    table: list = field(default_factory=list)   # (wg, ipl, jsplit, total_ms, far_ms, near_ms)


def _pick(table, incumbent, margin):
    """Fastest row, unless the incumbent shape is within ``margin`` of it."""
    best = min(table, key=lambda r: r[3])
    for r in table:
        if tuple(r[:len(incumbent)]) == tuple(incumbent) and r[3] <= best[3] * (1.0 + margin):
            return r
    return best


def tune(g, n, inputs, outputs, precision=SG_PREC_FP32C, nkernel=0, candidates=None,
         jsplits=(0,), reps=2, margin=0.02, log=None) -> TuneResult:
    """Time each candidate shape on the given device-resident set.

    g: a :class:`starsmasher_amd.Gravity`; inputs: 5 device arrays
    (x, y, z, m, range2); outputs: 4 device arrays of length n.
    """
    cands = candidates or CANDIDATES[precision]
    configured = tuple(g.get_config(precision))
    incumbent = configured
    if configured[0] == 0:
        # the automatic workgroup size: the shape it resolves to at this n
        g.compute_device(n, *inputs, *outputs, nkernel=nkernel, precision=precision)
        g.synchronize()
        incumbent = (g.last_config()[0],) + configured[1:]
    table = []
    for wg, ipl in cands:
        for js in jsplits:
            g.set_config(precision, wg, ipl, js)
            g.compute_device(n, *inputs, *outputs, nkernel=nkernel, precision=precision)
            g.synchronize()            # warm-up (buffers, code objects)
            tot = far = near = 0.0
            for _ in range(reps):
                g.compute_device(n, *inputs, *outputs, nkernel=nkernel, precision=precision)
                g.synchronize()
                p, f, nr = g.last_timings()
                tot += p + f + nr
                far += f
                near += nr
            row = (wg, ipl, js, tot / reps, far / reps, near / reps)
            table.append(row)
            if log:
                log(f"autotune prec={precision} wg={wg} ipl={ipl} jsplit={js}: "
                    f"{row[3]:.2f} ms (far {row[4]:.2f}, near {row[5]:.2f})")
    pick = _pick(table, incumbent, margin)
    best = tuple(pick[:3])
    # an automatic configuration that resolves to the pick stays automatic
    # (it keeps choosing by size on other sets)
    g.set_config(precision, *(configured if best == incumbent else best))
    return TuneResult(precision, best, pick[3], table)


BATCH_CANDIDATES = [(64, 1), (64, 2), (128, 1), (128, 2), (256, 1), (256, 2)]


def tune_batch(g, offsets, inputs, outputs, precision=SG_PREC_FP32C, nkernel=0,
               candidates=None, reps=2, margin=0.02, log=None) -> TuneResult:
    """Time each (workgroup size, i per lane) shape of the persistent batch
    kernel on the given device-resident batch (graph replays) and install the
    fastest with ``sg_set_batch_config``."""
    cands = candidates or BATCH_CANDIDATES
    incumbent = tuple(g.get_batch_config(precision))
    table = []
    for wg, ipl in cands:
        g.set_batch_config(precision, wg, ipl)
        g.compute_batch_device(offsets, *inputs, *outputs, nkernel=nkernel, precision=precision,
                               use_graph=True)
        g.synchronize()
        tot = 0.0
        for _ in range(reps):
            g.compute_batch_device(offsets, *inputs, *outputs, nkernel=nkernel,
                                   precision=precision, use_graph=True)
            g.synchronize()
            p, f, nr = g.last_timings()
            tot += p + f + nr
        row = (wg, ipl, 0, tot / reps, 0.0, 0.0)
        table.append(row)
        if log:
            log(f"autotune batch prec={precision} wg={wg} ipl={ipl}: {row[3]:.2f} ms")
    pick = _pick(table, incumbent, margin)
    best = tuple(pick[:2])
    g.set_batch_config(precision, *best)
    return TuneResult(precision, best, pick[3], table)
