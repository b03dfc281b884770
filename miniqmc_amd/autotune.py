"""Delay-rank and spline-tile autotuner for the batched FP32 sweep.

For each candidate (delay rank k, spline stencil slices S) a batch is built on
the real walker configuration, equilibrated for one sweep and timed over a
few sweeps with HIP events (FP64 refresh every sweep, as the FP32 form runs);
the fastest pair wins.  The delay rank trades the per-move cost of applying
pending rows (grows with k) against the per-flush cost of rewriting the
inverse (shrinks as 1/k); ranks whose stale rows do not fit the LDS run in the
streamed form (stale rows and capacitance inverse in HBM, a flush launch per
chunk).  S sets how many lanes of the proposals kernel share
one point's 4x4x4 stencil (S = 1, the library's default, keeps the partial sums in
registers; S > 1 reduces them through LDS).
"""
from __future__ import annotations

from . import qmcx

DELAYS = (8, 16, 32, 64, 128)
TILES = (1, 2, 4)


def tune(dev: qmcx.System, nwalkers: int, R0, seed: int = 11, tau: float = 0.3, sweeps: int = 2,
         delays=DELAYS, tiles=TILES, mode: int = qmcx.ACCEPT_REFERENCE):
    results = {}
    best = None
    for k in delays:
        for tile in tiles:
            try:
                b = qmcx.WalkerBatch(dev, nwalkers, delay_rank=k, precision=qmcx.FP32, tile=tile)
            except qmcx.QmcxError:
                continue   # does not fit (LDS) or tile invalid for this system
            b.set_positions(R0)
            b.evaluate_log()
            b.run(1, seed + 7919, tau, mode, recompute_every=1)
            b.run(sweeps, seed + 7919, tau, mode, recompute_every=1)
            ms = b.last_sweep_ms() / sweeps
            streamed = b.delay_form() == 1
            b.close()
            results[f"k{k}_s{tile}" + ("_streamed" if streamed else "")] = round(ms, 3)
            if best is None or ms < best[0]:
                best = (ms, k, tile)
    if best is None:
        raise qmcx.QmcxError("no delay rank / tile fits this system")
    return best[1], best[2], results
