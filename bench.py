#!/usr/bin/env python3
"""Walker-moves/s of the VMC particle-move sweep at NiO-64 on one MI355X.

One step = one sweep of every electron of every walker (W walkers x 768
single-electron moves), followed by evaluateGL, exactly the scope of the
reference's "Diffusion" timer (src/Drivers/miniqmc_sync_move.cpp:437-513).
Inputs (spline table, walker state) are resident in HBM before the timed
region.  The headline is the batched FP32 form (FP32 spline/distance/Jastrow/
Woodbury, FP64 log-det and acceptance product); the FP64 delay-1 GPU path and
the oracle's CPU path are timed beside it in the same run.

Prints one JSON line.
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

from miniqmc_amd import nlpp, qmcx, systems  # noqa: E402
from miniqmc_amd import autotune  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Metropolis legs: the FP32 form decides moves with |u - R^2| <= 2e-3 R^2 on the FP64 ratio (wider
# bands measured 1e-3 .. 3e-1: the walkers that still part are the ill-conditioned ones, at any band)
ACCEPT_BAND = 2e-3
METRIC = "walker-moves/s at NiO-64 on 1 MI355X, batched FP32 vs FP64 delay-1; log|Psi| err"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def path_bytes_per_move(sysd, k, acc, s=4):
    """Algorithmic HBM bytes per walker-move of the whole path (every kernel of a sweep), the
    compulsory-traffic model of the path: spline stencil 64*Np*s, positions 3(N+M)s, Jastrow
    state a*2*5*N*s, and the determinant s*n*(1 + 2c) + a*2*n*s + a*2*n^2*s/k with c = k/2 the
    mean delay count a move sees (0 for k = 1): the inverse row and the U/V panels, the U/V
    row writes of an accept, and one read and write of the inverse per flush (per accept with
    Sherman-Morrison).  NiO-64 FP32, k = 32, a = 0.5: 194,304 B.
    """
    n, N, M, Np = sysd.nup, sysd.nel, sysd.nions, sysd.npad
    cbar = k / 2.0 if k > 1 else 0.0
    b_spline = 64 * Np * s
    b_pos = 3 * (N + M) * s
    b_j = acc * 2 * 5 * N * s
    b_det = s * n * (1 + 2 * cbar) + acc * 2 * n * s + acc * 2 * n * n * s / k
    return b_spline + b_pos + b_j + b_det, dict(spline=b_spline, pos=b_pos, jastrow=b_j, det=b_det, cbar=cbar)


def move_kernel_bytes_per_move(sysd, k, acc, moves_per_walker, flushes, s=4):
    """Algorithmic HBM bytes per move of the fused move kernel (the spline stencil is read by
    the proposals kernel): the proposal's SPO row 5*n*s and record (J1 terms, displacement,
    uniform: 9 doubles), the moving electron's dpsiM row 3*n*s and J1 state 4*s, with k > 1 its
    stale inverse row n*s (read once per chunk); on accept (fraction a) the new A, dpsiM, d2psiM
    rows and the stale inverse row 6*n*s (written for a flush launch, or the U row read back by
    the kernel's own flush) and the J1 state 5*s.  Per launch and walker (moves_per_walker
    moves): the positions (3*N doubles) and J2 state (5*N*s) in and out, and per delayed-update
    flush done inside the kernel the inverse read and written, 2*n^2*s."""
    n, N = sysd.nup, sysd.nel
    per = 5 * n * s + 9 * 8 + 3 * n * s + 4 * s + (n * s if k > 1 else 0) + acc * (6 * n * s + 5 * s)
    per_launch = 2 * (3 * N * 8 + 5 * N * s) + flushes * 2 * n * n * s
    return per + per_launch / moves_per_walker


def proposal_bytes(sysd, s=4):
    """Algorithmic HBM bytes per proposal of the proposals kernel: the 4x4x4 stencil of the
    table, 64*Np*s, the SPO row written, 5*n*s, and the J1 terms, 5 doubles."""
    return 64 * sysd.npad * s + 5 * sysd.nup * s + 5 * 8


def measured_traffic(sysd, k, moves_per_launch):
    """HBM bytes per move-kernel launch from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by a copy of known size), when it was taken on
    this workload and launch shape; None otherwise."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{sysd.name}_fp32_k{k}.json")
    try:
        with open(path) as fh:
            rec = json.load(fh)
    except OSError:
        return None, None
    if rec.get("moves_per_launch") != moves_per_launch:
        return None, None
    return round(rec["bytes_per_launch"]), os.path.relpath(path, ROOT)


def cpu_baseline(sysd, R0, seed, tau, nthreads, mode, target_s=15.0):
    """The all-FP64 CPU oracle on the host's cores over a bounded sample of the same workload:
    two walkers per thread, a warm-up sweep and two timed calibration sweeps, then as many sweeps
    (same chains) as fill about target_s seconds."""
    from oracle import pyoracle as po
    orc = po.OracleSystem(sysd, spline_prec=8)
    nwalk = min(2 * nthreads, len(R0))
    first = orc.run_vmc(R0[:nwalk], 1, seed, tau, mode=mode, delay=1, nthreads=nthreads)
    cal = orc.run_vmc(first["R"], 2, seed + 1, tau, mode=mode, delay=1, nthreads=nthreads)
    nsw = int(max(1, min(2000, round(target_s / max(cal["seconds"] / 2, 1e-3)))))
    res = orc.run_vmc(cal["R"], nsw, seed + 2, tau, mode=mode, delay=1, nthreads=nthreads)
    moves = nwalk * sysd.nel * nsw
    return dict(value=moves / res["seconds"], unit="walker-moves/s", cores=nthreads, kind="port",
                sample=f"{nwalk} walkers x {nsw} sweeps of {sysd.name} ({moves} moves, {res['seconds']:.1f} s), "
                       f"all-FP64 oracle, delay 1, OpenMP over walkers (2 per thread), "
                       f"{'u < 0.5' if mode == qmcx.ACCEPT_REFERENCE else 'Metropolis'} acceptance")


def measure_leg(dev, sysd, R0, k, tile, mode, steps, warmup, fp64_steps, err_steps, recompute, seed, tau, one_walker):
    """One acceptance rule: the FP32-vs-FP64 error leg, the timed FP32 sweeps, the FP64 delay-1 sweeps."""
    W, N = len(R0), sysd.nel
    # ---- error leg: FP32 batched vs FP64 delay-1 from the same start, same random streams.
    # A walker's sequences agree until its first accept flip (a move whose u falls between
    # the two forms' ratio^2); log|Psi| is compared on the walkers without a flip.
    b32 = qmcx.WalkerBatch(dev, W, delay_rank=k, precision=qmcx.FP32, tile=tile)
    b64 = qmcx.WalkerBatch(dev, W, delay_rank=1, precision=qmcx.FP64)
    band = ACCEPT_BAND if mode == qmcx.ACCEPT_METROPOLIS else 0.0
    b32.set_accept_band(band)
    for b in (b32, b64):
        b.set_positions(R0)
        b.evaluate_log()
        b.set_trace(True)
    # the same start without the FP64 refresh: log|Psi| accumulated from the FP32 ratios alone
    bnr = qmcx.WalkerBatch(dev, W, delay_rank=k, precision=qmcx.FP32, tile=tile)
    bnr.set_accept_band(band)
    bnr.set_positions(R0)
    bnr.evaluate_log()
    hnr = bnr.run(1, seed, tau, mode, recompute_every=0)
    bnr.close()
    flipped = np.zeros(W, dtype=bool)
    flips = 0
    ratio_err, grad_err, chunk_acc = [], [], []
    lerr_sweeps = []
    n = sysd.nup
    for sw in range(err_steps):
        h32 = b32.run(1, seed, tau, mode, recompute_every=recompute)
        h64 = b64.run(1, seed, tau, mode)
        t32, t64 = b32.trace(), b64.trace()
        f = t32[:, :, 2] != t64[:, :, 2]
        flips += int((f.any(axis=1) & ~flipped).sum())   # first divergence of a walker's chain
        clean_moves = ~(flipped[:, None] | (np.cumsum(f, axis=1) > 0))
        rel = np.abs(t32[:, :, 0] - t64[:, :, 0]) / np.abs(t64[:, :, 0])
        g64 = t64[:, :, 3:9]
        gerr = (np.abs(t32[:, :, 3:9] - g64) / (1.0 + np.abs(g64))).max(axis=2)
        # accepted moves in each move's delay-rank chunk (k electrons of one determinant)
        cid = (np.arange(N) // n) * ((n + k - 1) // k) + (np.arange(N) % n) // k
        acc = np.stack([np.bincount(cid, weights=t64[w_, :, 2], minlength=cid.max() + 1) for w_ in range(W)])
        cacc = acc[:, cid]
        ratio_err.append(rel[clean_moves])
        grad_err.append(gerr[clean_moves])
        chunk_acc.append(cacc[clean_moves])
        flipped |= f.any(axis=1)
        lerr_sweeps.append(np.abs(h32[0] - h64[0]) / np.abs(h64[0]))
        if sw == 0:
            same = ~f.any(axis=1)
            lerr_norefresh = (np.abs(hnr[0] - h64[0]) / np.abs(h64[0]))[same]
    clean = ~flipped
    lerr = np.array(lerr_sweeps)[:, clean] if clean.any() else np.full((1, 1), np.nan)
    ratio_err, grad_err, chunk_acc = np.concatenate(ratio_err), np.concatenate(grad_err), np.concatenate(chunk_acc)
    for b in (b32, b64):
        b.set_trace(False)

    def pct(x):
        return {"p50": float(np.percentile(x, 50)), "p99": float(np.percentile(x, 99)),
                "p99.9": float(np.percentile(x, 99.9)), "max": float(x.max()), "moves": int(x.size)}

    bins = {}
    for name, lo, hi in (("accepts<=16", 0, 16), ("accepts17-24", 17, 24), ("accepts>24", 25, k)):
        sel = (chunk_acc >= lo) & (chunk_acc <= hi)
        if sel.any():
            bins[name] = pct(ratio_err[sel])
    errors = {"ratio_rel": pct(ratio_err), "grad_rel": pct(grad_err), "ratio_rel_by_chunk_accepts": bins,
              "log_psi_rel_before_refresh": {"max": float(lerr_norefresh.max()), "mean": float(lerr_norefresh.mean()),
                                             "walkers": int(lerr_norefresh.size), "sweeps": 1},
              "note": "FP32 delay-rank form vs the FP64 delay-1 GPU path move by move on walkers whose accept "
                      "sequence is still identical; ratio error |r32 - r64| / |r64|, gradient error "
                      "|g32 - g64| / (1 + |g64|) (largest component); the tail is set by moves whose Slater "
                      "matrix is ill-conditioned (the GPU parity tests bound each move by the oracle's "
                      "conditioning)"}
    log(f"error leg ({err_steps} sweeps): walkers with identical accept sequences {int(clean.sum())}/{W}, "
        f"flips {flips}; max rel |dlog| on those {np.nanmax(lerr):.2e} (before the refresh "
        f"{lerr_norefresh.max():.2e}); ratio err p50 {errors['ratio_rel']['p50']:.1e} p99 "
        f"{errors['ratio_rel']['p99']:.1e} max {ratio_err.max():.1e}; by chunk accepts " +
        ", ".join(f"{b_}: max {v['max']:.1e} ({v['moves']} moves)" for b_, v in bins.items()))

    # ---- headline: batched FP32 (FP64 LU recompute every `recompute` sweeps inside the timed region)
    b32.run(warmup, seed, tau, mode, recompute_every=recompute)
    b32.set_kernel_timing(True)
    a0 = b32.accepts().sum()
    t0 = time.perf_counter()
    b32.run(steps, seed, tau, mode, recompute_every=recompute)
    wall = time.perf_counter() - t0
    a1 = b32.accepts().sum()
    kt = b32.kernel_timing()
    b32.set_kernel_timing(False)
    mv_ms, mv_n = kt["move"]
    moves = W * N * steps
    value = moves / wall
    acc_timed = float(a1 - a0) / moves
    stages = {st: {"ms_per_step": ms / steps, "launches": n, "ms_per_launch": ms / max(n, 1)}
              for st, (ms, n) in kt.items()}
    log(f"FP32 k={k} ({'u < 0.5' if mode == qmcx.ACCEPT_REFERENCE else 'Metropolis'}): {value:.4e} moves/s, {1e3 * wall / steps:.2f} ms/sweep; stages " +
        ", ".join(f"{st} {v['ms_per_step']:.2f} ms/sweep" for st, v in stages.items()))
    # the same sweeps replayed from a captured HIP graph (one capture, then a timed replay)
    b32.set_graph(True)
    b32.run(steps, seed, tau, mode, recompute_every=recompute)
    t0 = time.perf_counter()
    b32.run(steps, seed, tau, mode, recompute_every=recompute)
    wall_g = time.perf_counter() - t0
    b32.set_graph(False)
    graph = {"value": moves / wall_g, "ms_per_step": 1e3 * wall_g / steps, "steps": steps,
             "note": "the run captured once into a HIP graph (qmcx_batch_set_graph) and replayed"}
    log(f"  graph replay: {graph['value']:.4e} moves/s, {graph['ms_per_step']:.2f} ms/sweep")

    # ---- FP64 delay-1 GPU path: batched (same walkers, continuing its own chain) and one walker alone
    b64.run(1, seed, tau, mode)
    t0 = time.perf_counter()
    b64.run(fp64_steps, seed, tau, mode)
    wall64 = time.perf_counter() - t0
    value64 = W * N * fp64_steps / wall64
    value64_1 = None
    if one_walker:
        b1 = qmcx.WalkerBatch(dev, 1, delay_rank=1, precision=qmcx.FP64)
        b1.set_positions(R0[:1])
        b1.evaluate_log()
        b1.run(1, seed, tau, mode)
        t0 = time.perf_counter()
        b1.run(1, seed, tau, mode)
        value64_1 = N / (time.perf_counter() - t0)
        b1.close()
    log(f"FP64 k=1: batched {value64:.4e} moves/s ({1e3 * wall64 / fp64_steps:.2f} ms/sweep)" +
        (f", one walker {value64_1:.4e} moves/s" if one_walker else ""))
    narb = int(b32.arbitrations().sum())
    b32.close()
    b64.close()
    return dict(graph=graph, band=band, arbitrations=narb, flips=flips, clean=clean, lerr=lerr, ratio_err=ratio_err, errors=errors, wall=wall, kt=kt, moves=moves,
                value=value,
                acc_timed=acc_timed, stages=stages, wall64=wall64, value64=value64, value64_1=value64_1)



def measure_nlpp(dev, sysd, R0, prec, steps, seed, tau, rmax=1.7):
    """NonLocalPP::evaluate over the batch (qmcx_nlpp_evaluate) after evaluateLog and one sweep:
    virtual-move ratios per second, outputs left in HBM (only the pair counts come back)."""
    W = R0.shape[0]
    wb = qmcx.WalkerBatch(dev, W, delay_rank=1, precision=prec)
    wb.set_positions(R0)
    wb.evaluate_log()
    wb.run(1, seed, tau, qmcx.ACCEPT_REFERENCE)
    pp = nlpp.NonLocalPP(sysd, rmax=rmax, seed=seed)
    spheres = [pp.randomize_walkers(W) for _ in range(steps + 1)]   # one rotation per walker and evaluation
    _, cnt = wb.nlpp_evaluate(rmax, spheres[0], max_pairs=8)
    ratios = int(cnt.sum()) * nlpp.NKNOTS
    wb.nlpp_evaluate(rmax, spheres[0], max_pairs=8, copy_out=False)
    t0 = time.perf_counter()
    for st in range(steps):
        wb.nlpp_evaluate(rmax, spheres[st + 1], max_pairs=8, copy_out=False)
    dt = (time.perf_counter() - t0) / steps
    s = 4 if prec == qmcx.FP32 else 8
    # per ratio: the value-only 4x4x4 stencil of every orbital, the inverse row, the other
    # electrons' positions and the J1/J2 state read once
    bpr = 64 * sysd.npad * s + sysd.nup * s + 3 * 8 * sysd.nel
    gbs = bpr * ratios / dt / 1e9
    traffic = None
    tp = os.path.join(ROOT, "profiles", "hbm_traffic_k_nlpp_ratios_nio64.json")
    if sysd.name == "nio64" and R0.shape[0] == 128 and os.path.exists(tp):
        traffic = json.load(open(tp))["f32" if prec == qmcx.FP32 else "f64"]["bytes_per_launch"]
    return {"value": ratios / dt, "unit": "virtual-move ratios/s", "ms_per_evaluate": 1e3 * dt, "steps": steps,
            "dtype": "f32" if prec == qmcx.FP32 else "f64", "rmax": rmax, "electron_ion_pairs": int(cnt.sum()),
            "ratios_per_evaluate": ratios,
            "roofline": {"bound": "hbm", "kernel": "k_nlpp_ratios", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                         "achieved": round(gbs, 2), "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_ratio": bpr,
                         "traffic": traffic, "traffic_source": "profiles/hbm_traffic_k_nlpp_ratios_nio64.json",
                         "note": "algorithmic bytes over the wall clock of an evaluation; knot k of the pairs around "
                                 "one ion at similar distances shares stencil rows in L2, so HBM traffic is below "
                                 "the algorithmic bytes"}}


MIXED = (("nio16", 64), ("nio32", 64), ("nio64", 32))


def measure_mixed(dev64, sys64, mode, steps, err_steps, seed, tau):
    """The uneven mixed batch: 64 NiO-16 + 64 NiO-32 + 32 NiO-64 walkers, each system with its own
    seeded spline table and Jastrows, advanced together by one grouped move launch and one grouped
    proposals launch per stage (qmcx_group_*), FP64 refresh every sweep of FP64 Slater rows
    (qmcx_batch_set_fp64_rows: the refresh carries no FP32 rounding of ill-conditioned matrices,
    which the reference driver's u < 0.5 rule visits at NiO-32).  Per size: the autotuned
    delay rank, moves/s (that size's moves over the group's wall time), log|Psi| error vs the FP64
    delay-1 GPU path from the same start and streams.  Beside it the same batches run one after
    the other, and the group replayed from a captured HIP graph."""
    parts, off = [], 0
    for name, nw in MIXED:
        sysd = sys64 if name == sys64.name else systems.make_system(name, seed=seed)
        dev = dev64 if name == sys64.name else qmcx.System(sysd, precisions=(qmcx.FP32, qmcx.FP64))
        R0 = systems.initial_positions(sysd, nw, seed=seed, walker0=off)
        k, tile, tune = autotune.tune(dev, nw, R0, seed=seed, tau=tau, mode=mode, delays=(16, 32), tiles=(1,))
        parts.append(dict(name=name, sysd=sysd, dev=dev, nw=nw, R0=R0, k=k, off=off, tune=tune))
        off += nw

    def batches(prec=qmcx.FP32):
        out = []
        for p in parts:
            b = qmcx.WalkerBatch(p["dev"], p["nw"], delay_rank=p["k"] if prec == qmcx.FP32 else 1, precision=prec,
                                 walker0=p["off"])
            b.set_fp64_rows(True)
            b.set_positions(p["R0"])
            b.evaluate_log()
            out.append(b)
        return out

    # error leg: the group vs FP64 delay-1 batches, same start, same streams
    grp_b, ref_b = batches(), batches(qmcx.FP64)
    g = qmcx.Group(grp_b)
    hg = g.run(err_steps, seed, tau, mode, recompute_every=1)
    errs = []
    for i, p in enumerate(parts):
        hr = ref_b[i].run(err_steps, seed, tau, mode)
        same = grp_b[i].accepts() == ref_b[i].accepts()
        e = np.abs(hg[i] - hr) / np.abs(hr)
        errs.append(dict(max_rel=float(e[:, same].max()) if same.any() else None,
                         walkers_same_accepts=int(same.sum())))
        ref_b[i].close()
    # timed: the group launched, then replayed from a graph, then the batches one after the other
    g.run(1, seed, tau, mode, recompute_every=1)
    t0 = time.perf_counter()
    g.run(steps, seed, tau, mode, recompute_every=1)
    wall = time.perf_counter() - t0
    g.set_graph(True)
    g.run(steps, seed, tau, mode, recompute_every=1)
    t0 = time.perf_counter()
    g.run(steps, seed, tau, mode, recompute_every=1)
    wall_g = time.perf_counter() - t0
    g.close()
    t0 = time.perf_counter()
    for b in grp_b:
        b.run(steps, seed, tau, mode, recompute_every=1)
    wall_sep = time.perf_counter() - t0
    for b in grp_b:
        b.close()
    total = sum(p["nw"] * p["sysd"].nel for p in parts) * steps
    sizes = {}
    for i, p in enumerate(parts):
        mv = p["nw"] * p["sysd"].nel * steps
        sizes[p["name"]] = {"walkers": p["nw"], "electrons": p["sysd"].nel, "delay_rank": p["k"],
                            "moves_per_s": mv / wall, "log_psi_err_vs_fp64": errs[i], "autotune": p["tune"]}
        log(f"mixed {p['name']}: {p['nw']} walkers, delay {p['k']}, {mv / wall:.4e} moves/s in the group, "
            f"log err vs FP64 {errs[i]['max_rel']:.2e} ({errs[i]['walkers_same_accepts']}/{p['nw']} same accepts)")
    log(f"mixed batch: {total / wall:.4e} moves/s grouped ({1e3 * wall / steps:.2f} ms/sweep), graph replay "
        f"{total / wall_g:.4e}, the batches one after the other {total / wall_sep:.4e}")
    return {"value": total / wall, "unit": "walker-moves/s", "ms_per_step": 1e3 * wall / steps, "steps": steps,
            "graph_replay_value": total / wall_g, "separate_launches_value": total / wall_sep, "sizes": sizes,
            "note": "one grouped move launch and one grouped proposals launch per determinant advance every "
                    "walker of the three systems; FP64 refresh every sweep per batch (FP64 Slater rows); log|Psi| after "
                    f"{err_steps} sweeps vs the FP64 delay-1 GPU path on walkers with the same accept counts"}


def measure_nio32(mode, steps, err_steps, fp64_steps, seed, tau, walkers=256, delay=32, rows=16):
    """The NiO-32 batched precision form: 256 walkers, FP32 at delay rank 32, with the FP64 inverse
    recomputed either every sweep or every 16 accepted rows of a walker's determinant (no per-sweep
    refresh), both inverting FP64 Slater rows.  Per cadence: walker-moves/s launched and replayed
    from a captured HIP graph, and
    log|Psi| after err_steps sweeps vs the FP64 delay-1 GPU path from the same start and streams;
    beside them the FP64 delay-1 path's own rate."""
    sysd = systems.make_system("nio32", seed=seed)
    dev = qmcx.System(sysd, precisions=(qmcx.FP32, qmcx.FP64))
    R0 = systems.initial_positions(sysd, walkers, seed=seed)
    moves = walkers * sysd.nel * steps
    b64 = qmcx.WalkerBatch(dev, walkers, delay_rank=1, precision=qmcx.FP64)
    b64.set_positions(R0)
    b64.evaluate_log()
    h64 = b64.run(err_steps, seed, tau, mode)
    a64 = b64.accepts()
    t0 = time.perf_counter()
    b64.run(fp64_steps, seed, tau, mode)
    value64 = walkers * sysd.nel * fp64_steps / (time.perf_counter() - t0)
    b64.close()
    out = {"walkers": walkers, "electrons": sysd.nel, "delay_rank": delay, "steps": steps,
           "fp64_delay1_value": value64, "cadences": {}}
    for name, nrows, every in (("refresh_every_sweep", 0, 1), (f"recompute_every_{rows}_accepted_rows", rows, 0)):
        b = qmcx.WalkerBatch(dev, walkers, delay_rank=delay, precision=qmcx.FP32)
        b.set_recompute_rows(nrows)
        b.set_fp64_rows(True)   # the FP64 inversions take FP64 Slater rows (no FP32 rounding of A)
        b.set_positions(R0)
        b.evaluate_log()
        h = b.run(err_steps, seed, tau, mode, recompute_every=every)
        same = b.accepts() == a64
        err = (np.abs(h - h64) / np.abs(h64))[:, same]
        b.run(1, seed, tau, mode, recompute_every=every)
        t0 = time.perf_counter()
        b.run(steps, seed, tau, mode, recompute_every=every)
        wall = time.perf_counter() - t0
        b.set_graph(True)
        b.run(steps, seed, tau, mode, recompute_every=every)
        t0 = time.perf_counter()
        b.run(steps, seed, tau, mode, recompute_every=every)
        wall_g = time.perf_counter() - t0
        b.close()
        out["cadences"][name] = {"value": moves / wall, "ms_per_step": 1e3 * wall / steps,
                                 "graph_replay_value": moves / wall_g,
                                 "speedup_vs_fp64_delay1": moves / wall / value64,
                                 "log_psi_err_vs_fp64": {"max_rel": float(err.max()) if err.size else None,
                                                         "mean_rel": float(err.mean()) if err.size else None,
                                                         "walkers_same_accepts": int(same.sum()),
                                                         "sweeps": err_steps}}
        log(f"nio32 {walkers} walkers, delay {delay}, {name}: {moves / wall:.4e} moves/s ({1e3 * wall / steps:.2f} "
            f"ms/sweep), graph replay {moves / wall_g:.4e}; log err vs FP64 after {err_steps} sweeps "
            f"{out['cadences'][name]['log_psi_err_vs_fp64']['max_rel']:.2e}")
    log(f"nio32 FP64 delay 1: {value64:.4e} moves/s")
    return out


def measure_j3(mode, steps, err_steps, fp64_steps, seed, tau, walkers=256, delay=32, rmax=1.7):
    """NiO-32 with the three-body Jastrow (the reference's -j; at NiO-64 the FP64 path's walker state
    with the J3 pair slots exceeds the LDS), 256 walkers: the FP32 form at delay rank 32 with FP64
    Slater rows and the per-sweep FP64 refresh, against the FP64 delay-1 GPU path from the same start
    and streams: walker-moves/s of both, log|Psi| error and identical accept sequences, and one
    NonLocalPP evaluation (FP32) with the J3 factor at every knot."""
    sysd = systems.make_system("nio32", seed=seed, j3=True)
    dev = qmcx.System(sysd, precisions=(qmcx.FP32, qmcx.FP64))
    R0 = systems.initial_positions(sysd, walkers, seed=seed)
    b64 = qmcx.WalkerBatch(dev, walkers, delay_rank=1, precision=qmcx.FP64)
    b64.set_positions(R0)
    b64.evaluate_log()
    h64 = b64.run(err_steps, seed, tau, mode)
    a64 = b64.accepts()
    t0 = time.perf_counter()
    b64.run(fp64_steps, seed, tau, mode)
    value64 = walkers * sysd.nel * fp64_steps / (time.perf_counter() - t0)
    b64.close()
    b = qmcx.WalkerBatch(dev, walkers, delay_rank=delay, precision=qmcx.FP32)
    b.set_fp64_rows(True)
    b.set_positions(R0)
    b.evaluate_log()
    h = b.run(err_steps, seed, tau, mode, recompute_every=1)
    same = b.accepts() == a64
    err = (np.abs(h - h64) / np.abs(h64))[:, same]
    b.run(1, seed, tau, mode, recompute_every=1)
    t0 = time.perf_counter()
    b.run(steps, seed, tau, mode, recompute_every=1)
    wall = time.perf_counter() - t0
    spheres = nlpp.NonLocalPP(sysd, rmax=rmax, seed=seed).randomize_walkers(walkers)
    b.nlpp_evaluate(rmax, spheres, max_pairs=8, ratios=False)
    t0 = time.perf_counter()
    _, cnt = b.nlpp_evaluate(rmax, spheres, max_pairs=8, ratios=False)
    nl_wall = time.perf_counter() - t0
    form = b.delay_form()
    b.close()
    moves = walkers * sysd.nel * steps
    nratios = int(np.minimum(cnt, 8).sum()) * 12
    out = {"workload": f"nio32 with the three-body Jastrow (-j), {walkers} walkers, tau {tau}", "walkers": walkers,
           "delay_rank": delay, "delay_form": "streamed" if form else "lds_resident", "steps": steps,
           "value": moves / wall, "ms_per_step": 1e3 * wall / steps, "fp64_delay1_value": value64,
           "speedup_vs_fp64_delay1": moves / wall / value64,
           "log_psi_err_vs_fp64": {"max_rel": float(err.max()) if err.size else None,
                                   "walkers_same_accepts": int(same.sum()), "sweeps": err_steps},
           "nlpp_f32": {"ratios_per_s": nratios / nl_wall, "ms": 1e3 * nl_wall, "ratios": nratios}}
    log(f"nio32 -j: FP32 k={delay} {moves / wall:.4e} moves/s ({1e3 * wall / steps:.2f} ms/sweep) vs FP64 delay 1 "
        f"{value64:.4e}; log err {out['log_psi_err_vs_fp64']['max_rel']}; same accepts {int(same.sum())}/{walkers}; "
        f"NonLocalPP {nratios / nl_wall:.4e} ratios/s")
    return out


def measure_single_walker(seed, tau, mode):
    """The two one-walker configurations: the plain FP64 delay-1 drop-in (Sherman-Morrison, one
    workgroup) on the CPU-runnable NiO-8 cubic cell (96 e-, 5 sweeps, tau 0.5) and on NiO-32 (384 e-, 20
    sweeps, tau 0.3): walker-moves/s over the sweeps."""
    out = {}
    for name, nsw, t in (("nio8-cubic", 5, 0.5), ("nio32", 20, tau)):
        sysd = systems.make_system(name, seed=seed)
        dev = qmcx.System(sysd, precisions=(qmcx.FP64,))
        b = qmcx.WalkerBatch(dev, 1, delay_rank=1, precision=qmcx.FP64)
        b.set_positions(systems.initial_positions(sysd, 1, seed=seed))
        b.evaluate_log()
        b.run(1, seed, t, mode)
        t0 = time.perf_counter()
        b.run(nsw, seed, t, mode)
        wall = time.perf_counter() - t0
        b.close()
        out[name] = {"value": sysd.nel * nsw / wall, "unit": "walker-moves/s", "sweeps": nsw, "tau": t,
                     "ms_per_step": 1e3 * wall / nsw, "form": "FP64, delay rank 1, one walker"}
        log(f"{name} one walker, FP64 delay 1: {sysd.nel * nsw / wall:.4e} moves/s ({1e3 * wall / nsw:.2f} ms/sweep)")
    return out


def measure_evaluate_log(dev, sysd, R0, prec, steps):
    """WaveFunction::evaluateLog from scratch for the whole batch (SPO rows at every electron, FP64 LU,
    J1/J2 over all pairs, evaluateGL): walkers per second."""
    W = R0.shape[0]
    wb = qmcx.WalkerBatch(dev, W, delay_rank=1, precision=prec)
    wb.set_positions(R0)
    wb.evaluate_log()
    t0 = time.perf_counter()
    for _ in range(steps):
        wb.evaluate_log()
    dt = (time.perf_counter() - t0) / steps
    return {"value": W / dt, "unit": "walkers/s", "ms_per_batch": 1e3 * dt, "steps": steps, "walkers": W,
            "dtype": "f32" if prec == qmcx.FP32 else "f64"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--system", default="nio64")
    ap.add_argument("--walkers", type=int, default=128)
    ap.add_argument("--delay", type=int, default=0, help="delay rank (0: autotune delay rank and spline tile, miniqmc_amd.autotune)")
    ap.add_argument("--tau", type=float, default=0.3)
    ap.add_argument("--accept", choices=("reference", "metropolis"), default="reference",
                    help="acceptance rule: the reference driver's u < r with its default r = 0.5 "
                         "(miniqmc.cpp:202,445), or Metropolis u < |ratio|^2")
    ap.add_argument("--other-steps", type=int, default=20,
                    help="timed sweeps of the other acceptance rule, reported beside the headline (0: skip)")
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--fp64-steps", type=int, default=4)
    ap.add_argument("--recompute", type=int, default=1,
                    help="FP32 form: FP64 LU inverse recompute every this many sweeps (0: never)")
    ap.add_argument("--err-steps", type=int, default=2, help="sweeps of the FP32-vs-FP64 error leg")
    ap.add_argument("--nlpp-steps", type=int, default=5, help="timed NonLocalPP evaluations per precision (0: skip)")
    ap.add_argument("--mixed-steps", type=int, default=10,
                    help="timed sweeps of the mixed NiO-16/32/64 batch in one grouped launch (0: skip)")
    ap.add_argument("--nio32-steps", type=int, default=10,
                    help="timed sweeps of the NiO-32 batched leg (256 walkers, two recompute cadences; 0: skip)")
    ap.add_argument("--j3-steps", type=int, default=5,
                    help="timed sweeps of NiO-32 with the three-body Jastrow (-j; 0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device five times (a kernel of known HBM traffic "
                         "to scale the FETCH_SIZE/WRITE_SIZE counters by) and exit")
    args = ap.parse_args()
    mode = qmcx.ACCEPT_REFERENCE if args.accept == "reference" else qmcx.ACCEPT_METROPOLIS
    if args.calibrate_copy:
        import torch
        nbytes = args.calibrate_copy << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy": {"bytes_read": nbytes, "bytes_written": nbytes, "dispatches": 5}}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: walkers are independent, so a multi-GPU run is one "
              "replica of this process per GPU with no exchange inside a sweep; --gpus must be 1.",
              file=sys.stderr)
        sys.exit(2)

    t_setup = time.time()
    sysd = systems.make_system(args.system, seed=args.seed)
    dev = qmcx.System(sysd, precisions=(qmcx.FP32, qmcx.FP64))
    W, N = args.walkers, sysd.nel
    R0 = systems.initial_positions(sysd, W, seed=args.seed)
    log(f"setup {time.time() - t_setup:.1f}s: {sysd.name} N={N} norb={sysd.norb} ng={sysd.ng} "
        f"table={sysd.table_bytes() / 1e6:.0f} MB W={W}")

    # ---- delay rank / spline tile (project autotuner)
    if args.delay > 0:
        k, tile = args.delay, 0
        tune = None
    else:
        k, tile, tune = autotune.tune(dev, W, R0, seed=args.seed, tau=args.tau, mode=mode)
        log(f"autotune: delay {k}, tile {tile}: {tune}")

    L = measure_leg(dev, sysd, R0, k, tile, mode, args.steps, args.warmup, args.fp64_steps, args.err_steps,
                    args.recompute, args.seed, args.tau, one_walker=True)
    flips, clean, lerr, ratio_err = L["flips"], L["clean"], L["lerr"], L["ratio_err"]
    wall, kt, moves, value, acc_timed = L["wall"], L["kt"], L["moves"], L["value"], L["acc_timed"]
    stages, wall64, value64, value64_1 = L["stages"], L["wall64"], L["value64"], L["value64_1"]
    mv_ms, mv_n = kt["move"]
    # the other acceptance rule, shorter, reported beside the headline
    other_rule = "metropolis" if args.accept == "reference" else "reference"
    other = None
    if args.other_steps > 0:
        O = measure_leg(dev, sysd, R0, k, tile,
                        qmcx.ACCEPT_METROPOLIS if other_rule == "metropolis" else qmcx.ACCEPT_REFERENCE,
                        args.other_steps, 1, 2, args.err_steps, args.recompute, args.seed, args.tau, one_walker=False)
        other = {"accept_rule": other_rule, "value": O["value"], "ms_per_step": 1e3 * O["wall"] / args.other_steps,
                 "steps": args.other_steps, "acceptance": O["acc_timed"],
                 "fp64_delay1_value": O["value64"], "speedup_vs_fp64_delay1": O["value"] / O["value64"],
                 "accept_flips": O["flips"], "walkers_identical_sequence": int(O["clean"].sum()),
                 "log_psi_err_max_rel": float(np.nanmax(O["lerr"])), "errors": O["errors"],
                 "accept_band": O["band"], "fp64_arbitrations": O["arbitrations"]}

    # ---- rooflines.  The dominant kernel (most of the sweep's time) is the fused move kernel:
    # one workgroup per walker walks its electrons in order, a latency-bound chain that moves
    # few bytes per move.  The proposals kernel carries the spline's table traffic; the whole
    # sweep is also put against the path's algorithmic bytes.
    bpm_path, parts = path_bytes_per_move(sysd, k, acc_timed)
    # moves per walker and launch from the launch count: the whole determinant (n moves) when
    # the kernel applies each chunk's delayed updates itself (no flush launches) or k = 1, else
    # one chunk of k; in-kernel flushes per launch: every chunk's but the last one's of the
    # first determinant when the FP64 refresh follows each sweep
    per_walker = moves // (W * max(mv_n, 1))
    nch = -(-sysd.nup // k) if k > 1 else 0
    inflush = k > 1 and kt["flush"][1] == 0
    flushes = (nch - (0.5 if args.recompute == 1 else 0.0)) if inflush else 0.0
    bpm_impl = move_kernel_bytes_per_move(sysd, k, acc_timed, per_walker, flushes)
    # the move kernel's share of the path's algorithmic bytes: everything but the spline stencil,
    # which the proposals kernel reads
    bpm = bpm_path - parts["spline"]
    launch_ms = mv_ms / max(mv_n, 1)
    achieved = bpm * W * per_walker / (launch_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(sysd, k, W * per_walker)
    roofline = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, traffic_source=traffic_src,
                    kernel="k_moves<float>", bytes_per_move=round(bpm, 1),
                    bytes_per_move_kernel_count=round(bpm_impl, 1),
                    moves_per_launch=W * per_walker, flushes_per_launch=flushes,
                    avg_launch_ms=round(launch_ms, 4),
                    note="bytes_per_move: the path's algorithmic bytes per move less the spline stencil; "
                         "bytes_per_move_kernel_count: the kernel's own rows read and written per move")
    pr_ms, pr_n = kt["proposals"]
    pr_launch = pr_ms / max(pr_n, 1)
    pr_units = W * sysd.nup
    pr_ach = proposal_bytes(sysd) * pr_units / (pr_launch * 1e-3) / 1e9 if pr_n else None
    roofline_proposals = dict(bound="hbm", kernel="k_spo_proposals<float>", unit="GB/s", peak=HBM_PEAK_GBS,
                              achieved=round(pr_ach, 2) if pr_ach else None,
                              frac=round(pr_ach / HBM_PEAK_GBS, 4) if pr_ach else None,
                              bytes_per_proposal=proposal_bytes(sysd), proposals_per_launch=pr_units,
                              avg_launch_ms=round(pr_launch, 4),
                              note="runs on its own stream beside the move kernel")
    sweep_gbs = bpm_path * moves / wall / 1e9
    roofline_sweep = dict(bound="hbm", unit="GB/s", peak=HBM_PEAK_GBS, achieved=round(sweep_gbs, 2),
                          frac=round(sweep_gbs / HBM_PEAK_GBS, 4), bytes_per_move=round(bpm_path, 1),
                          note="whole path's algorithmic bytes per walker-move over the timed sweeps")

    nl = {}
    if args.nlpp_steps > 0:
        for prec in (qmcx.FP32, qmcx.FP64):
            r = measure_nlpp(dev, sysd, R0, prec, args.nlpp_steps, args.seed, args.tau)
            nl[r["dtype"]] = r
            log(f"NonLocalPP evaluate ({r['dtype']}): {r['value']:.4e} ratios/s, {r['ms_per_evaluate']:.2f} ms for "
                f"{r['ratios_per_evaluate']} ratios ({r['electron_ion_pairs']} electron-ion pairs)")

    elog = {}
    if args.nlpp_steps > 0:
        for prec in (qmcx.FP32, qmcx.FP64):
            r = measure_evaluate_log(dev, sysd, R0, prec, args.nlpp_steps)
            elog[r["dtype"]] = r
            log(f"evaluateLog ({r['dtype']}): {r['ms_per_batch']:.2f} ms for {W} walkers")

    mixed = None
    if args.mixed_steps > 0:
        mixed = measure_mixed(dev, sysd, mode, args.mixed_steps, args.err_steps, args.seed, args.tau)

    nio32 = None
    if args.nio32_steps > 0:
        nio32 = measure_nio32(mode, args.nio32_steps, args.err_steps, 2, args.seed, args.tau)
        nio32["one_walker_configs"] = measure_single_walker(args.seed, args.tau, mode)

    j3 = None
    if args.j3_steps > 0:
        j3 = measure_j3(mode, args.j3_steps, args.err_steps, 2, args.seed, args.tau)

    cpu = None
    if not args.no_cpu:
        nthreads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(sysd, R0, args.seed, args.tau, nthreads, mode)
        log(f"CPU oracle: {cpu['value']:.4e} moves/s on {nthreads} threads")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "walker-moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{sysd.name} VMC sweep: {N} e-, {sysd.norb} orbitals, {sysd.ng[0]}^3 FP32 spline "
                        f"({sysd.table_bytes() / 1e6:.0f} MB), {W} walkers, tau {args.tau}, "
                        f"{'u < 0.5 acceptance (the reference driver)' if args.accept == 'reference' else 'Metropolis'}; "
                        f"batched FP32, delay rank {k}",
            "walkers": W, "electrons": N, "orbitals": sysd.norb, "grid": list(sysd.ng), "delay_rank": k,
            "accept_rule": args.accept,
            "spline_tile": tile, "moves_per_step": W * N,
        },
        "roofline": roofline,
        "roofline_proposals": roofline_proposals,
        "roofline_sweep": roofline_sweep,
        "cpu_baseline": cpu,
        "fp64_delay1": {"value": value64, "unit": "walker-moves/s", "ms_per_step": 1e3 * wall64 / args.fp64_steps,
                        "steps": args.fp64_steps, "walkers": W,
                        "form": "batched, 1 workgroup per walker, in-kernel Sherman-Morrison",
                        "one_walker_value": value64_1},
        "speedup_vs_fp64_delay1": value / value64,
        "graph_replay": L["graph"],
        "other_accept_rule": other,
        "log_psi_err": {"max_rel": float(np.nanmax(lerr)), "mean_rel": float(np.nanmean(lerr)),
                        "walkers_compared": int(clean.sum()), "sweeps": args.err_steps,
                        "vs": "FP64 delay-1 GPU path (parity-pinned to the FP64 oracle), same start and streams; "
                              "walkers whose accept sequence stayed identical"},
        "ratio_err_max_rel": float(ratio_err.max()),
        "errors": L["errors"],
        "accept_flips": flips,
        "walkers_identical_sequence": int(clean.sum()),
        "acceptance": acc_timed,
        "recompute_every_sweeps": args.recompute,
        "stages": stages,
        "bytes_per_move_parts": parts,
        "autotune": tune,
        "nlpp_virtual_ratios": nl or None,
        "evaluate_log": elog or None,
        "mixed_batch": mixed,
        "nio32_batched": nio32,
        "three_body_jastrow": j3,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
