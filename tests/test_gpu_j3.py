"""Three-body (eeI) Jastrow on the device (ThreeBodyJastrow<PolynomialFunctor3D>,
src/QMCWaveFunctions/Jastrow/ThreeBodyJastrow.h:30-677; miniqmc's -j option) against the FP64
oracle, through the C ABI: evaluateLog with its G and L, whole sweeps (both acceptance rules,
Sherman-Morrison and delayed updates, FP64 and FP32), the flex_* stage loop, NiO-64, and the
virtual-move and NonLocalPP ratios.

Tolerances (stated per test): FP64 ratios 1e-8 of max(|r|, 1), log|Psi| 1e-11 relative, G/L 1e-8;
FP32 (the FP32 form keeps J3 in FP32 like J2) ratios 2e-3 of max(|r|, 1) under Metropolis, and at
NiO-64 under u < 0.5 scaled by the oracle's conditioning of each move (8 2^-24 max(s_1..s_m), at
least 2e-5); accept sequences identical.
"""
import numpy as np
import pytest

from miniqmc_amd import qmcx, systems
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def cubic_j3():
    sysd = systems.make_system("nio8-cubic", j3=True)
    return sysd, qmcx.System(sysd), po.OracleSystem(sysd)


@pytest.mark.parametrize("prec,tol", [(qmcx.FP64, 1e-11), (qmcx.FP32, 1e-6)])
def test_j3_evaluate_log_vs_oracle(cubic_j3, prec, tol):
    sysd, dev, orc = cubic_j3
    R0 = systems.initial_positions(sysd, 3, seed=21)
    _say("J3 evaluateLog, precision", prec)
    wb = qmcx.WalkerBatch(dev, 3, delay_rank=1, precision=prec)
    wb.set_positions(R0)
    log = wb.evaluate_log()
    G, L, _ = wb.gl()
    for w in range(3):
        v, Gr, Lr = orc.evaluate_log(R0[w])
        parts = orc.evaluate_log_parts(R0[w])[0]
        assert parts[4] != 0.0
        _say(f"  walker {w}: log {log[w]:.12f} oracle {v:.12f} (J3 part {parts[4]:.6f})")
        assert abs(log[w] - v) <= tol * abs(v)
        np.testing.assert_allclose(G[w], Gr, rtol=tol * 1e3, atol=tol * 1e3)
        np.testing.assert_allclose(L[w], Lr, rtol=tol * 1e3, atol=tol * 1e4)


EPS32 = 2.0 ** -24


def _run_compare(sysd, dev, orc, nw, nsweeps, delay, prec, mode, tau, seed, tol_ratio, tol_log, cond=0.0, setup=None,
                 recompute=0):
    """Sweeps of the device against the oracle's (same delay rank).  Ratios: |dr| / max(|r|, 1) <=
    tol_ratio, or, with cond > 0, |dr| / |r| <= max(cond 2^-24 max(s_1 .. s_m), tol_ratio) with s the
    oracle's conditioning of the move (its relative sensitivity to relative perturbations of the
    Slater matrix and the new row, from a delay-1 run whose chain is the same).  Gradients: 100
    tol_ratio (resp. ten times the bound) relative to 1 + |g|.  log|Psi| after each sweep (no FP64
    refresh, or after it with recompute) within tol_log relative, plus with cond the summed bounds of
    the accepted moves.  setup(batch) runs before the positions are set."""
    R0 = systems.initial_positions(sysd, nw, seed=seed)
    ref = orc.run_vmc(R0, nsweeps, seed, tau, mode=mode, delay=delay, trace=True)
    sens = orc.run_vmc(R0, nsweeps, seed, tau, mode=mode, delay=1, trace=True, sensitivity=True) if cond else None
    wb = qmcx.WalkerBatch(dev, nw, delay_rank=delay, precision=prec)
    if setup is not None:
        setup(wb)
    wb.set_positions(R0)
    np.testing.assert_allclose(wb.evaluate_log(), ref["log_init"], rtol=tol_log)
    wb.set_trace(True)
    n = sysd.nup
    for s in range(nsweeps):
        hist = wb.run(1, seed, tau, mode, recompute_every=recompute)
        tr, rt = wb.trace(), ref["trace"][:, s]
        assert np.array_equal(tr[:, :, 2], rt[:, :, 2]), ("accept sequence", s)
        gerr = (np.abs(tr[:, :, 3:9] - rt[:, :, 3:9]) / (1 + np.abs(rt[:, :, 3:9]))).max(axis=2)
        dlog = np.abs(hist[0] - ref["log"][:, s])
        allowed = tol_log * np.abs(ref["log"][:, s])
        if cond:
            assert np.array_equal(sens["trace"][:, s, :, 2], rt[:, :, 2])
            sm = sens["sensitivity"][:, s]
            smax = np.concatenate([np.maximum.accumulate(sm[:, :n], axis=1),
                                   np.maximum.accumulate(sm[:, n:], axis=1)], axis=1)
            bound = np.maximum(cond * EPS32 * smax, tol_ratio)
            rel = np.abs(tr[:, :, 0] - rt[:, :, 0]) / np.abs(rt[:, :, 0])
            gbound = 10 * bound
            allowed = allowed + (bound * (rt[:, :, 2] > 0)).sum(axis=1)
        else:
            bound = np.full(rt.shape[:2], tol_ratio)
            rel = np.abs(tr[:, :, 0] - rt[:, :, 0]) / np.maximum(np.abs(rt[:, :, 0]), 1.0)
            gbound = 100 * bound
        _say(f"  sweep {s}: ratio err {rel.max():.2e} (worst err/bound {(rel / bound).max():.3f}), "
             f"grad err {gerr.max():.2e}, |dlog| {dlog.max():.2e} (allowed {allowed.min():.2e})")
        assert np.all(rel <= bound), (rel / bound).max()
        assert np.all(gerr <= gbound), (gerr / gbound).max()
        assert np.all(dlog <= allowed), (dlog, allowed)
    G, L, _ = wb.gl()
    return ref, G, L, wb


@pytest.mark.parametrize("delay,mode", [(1, qmcx.ACCEPT_METROPOLIS), (4, qmcx.ACCEPT_METROPOLIS),
                                        (16, qmcx.ACCEPT_METROPOLIS), (1, qmcx.ACCEPT_REFERENCE)])
def test_j3_sweep_fp64_vs_oracle(cubic_j3, delay, mode):
    """FP64: ratios within 1e-8 of max(|r|, 1) (the reference driver's u < 0.5 also visits
    near-nodal configurations, where rounding is amplified; there the delayed updates are tested
    in FP32 with conditioning-scaled bounds, test_j3_nio64_fp32_vs_oracle)."""
    sysd, dev, orc = cubic_j3
    _say("J3 sweep FP64 delay", delay, "accept mode", mode)
    ref, G, L, wb = _run_compare(sysd, dev, orc, 3, 2, delay, qmcx.FP64, mode, 0.5, 11, 1e-8, 1e-11)
    assert np.array_equal(wb.accepts(), ref["naccept"])
    np.testing.assert_allclose(wb.positions(), ref["R"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(G, ref["G"], rtol=1e-8, atol=1e-8)
    np.testing.assert_allclose(L, ref["L"], rtol=1e-8, atol=1e-7)


@pytest.mark.parametrize("delay", [1, 8])
def test_j3_sweep_fp32_vs_oracle(cubic_j3, delay):
    """FP32 under Metropolis acceptance (as the FP32 parity tests without J3): the accept sequence
    of the FP64 oracle, ratios within 2e-3 of max(|r|, 1), log|Psi| within 1e-6."""
    sysd, dev, orc = cubic_j3
    _say("J3 sweep FP32 delay", delay)
    _run_compare(sysd, dev, orc, 3, 2, delay, qmcx.FP32, qmcx.ACCEPT_METROPOLIS, 0.5, 11, 2e-3, 1e-6)


@pytest.mark.parametrize("prec,delay,tol,tol_log", [(qmcx.FP64, 1, 1e-9, 1e-11), (qmcx.FP64, 4, 1e-9, 1e-11),
                                                     (qmcx.FP32, 4, 2e-3, 1e-6)])
def test_j3_flex_loop_vs_oracle(cubic_j3, prec, delay, tol, tol_log):
    """The flex_* stage calls with caller displacements retrace the oracle's sweep with J3 (delay rank 4:
    accepted rows pending across the calls in the streamed form)."""
    sysd, dev, orc = cubic_j3
    nw, seed, tau = 2, 11, 0.5
    R0 = systems.initial_positions(sysd, nw, seed=seed)
    ref = orc.run_vmc(R0, 1, seed, tau, mode=qmcx.ACCEPT_METROPOLIS, delay=1, trace=True)
    wb = qmcx.WalkerBatch(dev, nw, delay_rank=delay, precision=prec)
    wb.set_positions(R0)
    wb.evaluate_log()
    for iel in range(sysd.nel):
        displ = np.array([np.sqrt(tau) * np.asarray(po.move_randoms(seed, w, 0, iel)[0]) for w in range(nw)])
        ratio, gnow, gnew = wb.flex_ratio_grad(iel, displ)
        rt = ref["trace"][:, 0, iel]
        assert (np.abs(ratio - rt[:, 0]) / np.maximum(np.abs(rt[:, 0]), 1.0)).max() < tol, iel
        assert (np.abs(gnew - rt[:, 6:9]) / (1 + np.abs(rt[:, 6:9]))).max() < 100 * tol, iel
        assert (np.abs(gnow - rt[:, 3:6]) / (1 + np.abs(rt[:, 3:6]))).max() < 100 * tol, iel
        wb.flex_accept_restore(iel, rt[:, 2].astype(np.int32))
    wb.flex_complete_updates()
    log = wb.flex_evaluate_gl()
    np.testing.assert_allclose(log, ref["log"][:, 0], rtol=tol_log)


def test_j3_nio64_fp32_vs_oracle():
    """The headline system with the reference's eeI Jastrow: NiO-64, 2 walkers, one sweep, FP32
    delay 16 (the largest delay rank whose walker state fits the LDS with a three-body Jastrow),
    the reference driver's acceptance."""
    sysd = systems.make_system("nio64", j3=True)
    dev, orc = qmcx.System(sysd), po.OracleSystem(sysd)
    _say("J3 NiO-64 FP32 delay 16")
    _run_compare(sysd, dev, orc, 2, 1, 16, qmcx.FP32, qmcx.ACCEPT_REFERENCE, 0.3, 5, 2e-5, 1e-6, cond=8.0)
    # the FP64 form's walker state with a three-body Jastrow exceeds the LDS at NiO-64
    with pytest.raises(qmcx.QmcxError, match="LDS"):
        qmcx.WalkerBatch(dev, 2, delay_rank=1, precision=qmcx.FP64)


@pytest.mark.parametrize("prec,delay,rows", [(qmcx.FP64, 1, 4), (qmcx.FP64, 1, 16), (qmcx.FP32, 1, 5),
                                             (qmcx.FP32, 8, 4), (qmcx.FP32, 16, 16)])
def test_j3_recompute_cadence_vs_oracle(cubic_j3, prec, delay, rows):
    """The FP64 recompute after every `rows` accepted moves of a walker's determinant (the walker
    stops, its inverse is rebuilt, it resumes from its own electron; Sherman-Morrison moves, or the
    FP32 form's delayed updates applied inside the move kernel) with the three-body Jastrow:
    the oracle's sweeps, FP64 to 1e-8 in the ratios and 1e-11 in log|Psi| with identical positions,
    G and L; FP32 to 2e-3 and 1e-6 (Metropolis acceptance)."""
    sysd, dev, orc = cubic_j3
    fp64 = prec == qmcx.FP64
    _say("J3 recompute cadence", prec, "delay", delay, "rows", rows)
    ref, G, L, wb = _run_compare(sysd, dev, orc, 3, 2, delay, prec, qmcx.ACCEPT_METROPOLIS, 0.5, 11,
                                 1e-8 if fp64 else 2e-3, 1e-11 if fp64 else 1e-6,
                                 setup=lambda b: b.set_recompute_rows(rows))
    assert np.array_equal(wb.accepts(), ref["naccept"])
    if fp64:
        np.testing.assert_allclose(wb.positions(), ref["R"], rtol=0, atol=1e-12)
        np.testing.assert_allclose(G, ref["G"], rtol=1e-8, atol=1e-8)
        np.testing.assert_allclose(L, ref["L"], rtol=1e-8, atol=1e-7)


@pytest.mark.parametrize("delay", [1, 8])
def test_j3_fp32_fp64_rows_refresh_vs_oracle(cubic_j3, delay):
    """The FP32 form with FP64 Slater rows and the per-sweep FP64 refresh inverting them, with the
    three-body Jastrow: ratios within 2e-3 of max(|r|, 1) and, after each sweep's refresh, log|Psi|
    within 1e-7 of the oracle (FP64 determinant, FP32 Jastrow sums)."""
    sysd, dev, orc = cubic_j3
    _say("J3 FP32 with FP64 rows, delay", delay)
    _run_compare(sysd, dev, orc, 3, 2, delay, qmcx.FP32, qmcx.ACCEPT_METROPOLIS, 0.5, 11, 2e-3, 1e-7,
                 setup=lambda b: b.set_fp64_rows(True), recompute=1)


@pytest.mark.parametrize("delay", [8, 16])
def test_j3_near_tie_band_vs_oracle(cubic_j3, delay):
    """The FP32 form under Metropolis with the near-tie band and the three-body Jastrow: a move whose
    uniform lies within band R^2 of the FP32 R^2 is decided on its FP64 ratio (the determinant ratio
    refined against the FP64 Slater rows, the J1, J2 and J3 terms in FP64).  A wide band (0.5) sends
    a large share of the moves through the re-evaluation.  One sweep against the FP64 oracle:
    identical accept sequences, ratios (the re-evaluated ones included, whose J3 factor is up to 2e-2
    here) within 2e-3 of max(|r|, 1), log|Psi| within 1e-6."""
    sysd, dev, orc = cubic_j3
    _say("J3 near-tie band, delay", delay)
    ref, G, L, wb = _run_compare(sysd, dev, orc, 3, 1, delay, qmcx.FP32, qmcx.ACCEPT_METROPOLIS, 0.5, 11, 2e-3, 1e-6,
                                 setup=lambda b: b.set_accept_band(0.5))
    narb = int(wb.arbitrations().sum())
    _say(f"  re-evaluated moves {narb} of {3 * sysd.nel}")
    assert narb > 20
    assert np.array_equal(wb.accepts(), ref["naccept"])


def test_j3_nio64_near_tie_band_vs_oracle():
    """NiO-64 with the three-body Jastrow, the FP32 form at delay 16 under Metropolis with the near-tie
    band 1e-2 (the re-evaluation's LDS region beside the J3 pair slots: 156 KB), 2 walkers, one sweep:
    the oracle's accept sequence, ratios within the conditioning bound 8 2^-24 max(s_1..s_m) (at least
    2e-5), log|Psi| within 1e-6."""
    sysd = systems.make_system("nio64", j3=True)
    dev, orc = qmcx.System(sysd), po.OracleSystem(sysd)
    _say("J3 NiO-64 FP32 delay 16, near-tie band")
    ref, G, L, wb = _run_compare(sysd, dev, orc, 2, 1, 16, qmcx.FP32, qmcx.ACCEPT_METROPOLIS, 0.3, 5, 2e-5, 1e-6,
                                 cond=8.0, setup=lambda b: b.set_accept_band(1e-2))
    _say(f"  re-evaluated moves {int(wb.arbitrations().sum())}")
    assert wb.delay_form() == 0


@pytest.mark.parametrize("prec,delay,tol", [(qmcx.FP64, 1, 1e-10), (qmcx.FP64, 4, 1e-10), (qmcx.FP32, 4, 2e-3)])
def test_j3_virtual_ratios_vs_oracle(cubic_j3, prec, delay, tol):
    """WaveFunction::evaluateRatios with the three-body Jastrow (ThreeBodyJastrow.h:262-273) through
    qmcx_flex_evaluate_ratios at the NonLocalPP quadrature points of every electron of 2 walkers,
    after a Metropolis sweep (the ion masks and sums then come from the move kernel's accepts), against
    the oracle.  Ratios within tol of max(|r|, 1)."""
    from miniqmc_amd import nlpp
    sysd, dev, orc = cubic_j3
    _say("J3 virtual ratios, precision", prec, "delay", delay)
    nw = 2
    wb = qmcx.WalkerBatch(dev, nw, delay_rank=delay, precision=prec)
    wb.set_positions(systems.initial_positions(sysd, nw, seed=4))
    wb.evaluate_log()
    wb.run(1, 6, 0.3, qmcx.ACCEPT_METROPOLIS, recompute_every=1)
    R = wb.positions()
    pp = nlpp.NonLocalPP(sysd, rmax=1.7, seed=3)
    sphere = pp.randomize_walkers(nw)
    worst, count, moved = 0.0, 0, 0.0
    for jel in range(sysd.nel):
        vpos, npair = pp.virtual_positions(R, jel, sphere)
        if npair.max() == 0:
            continue
        got = wb.flex_evaluate_ratios(jel, vpos)
        for w in range(nw):
            ref = orc.evaluate_ratios(R[w], jel, vpos[w])
            err = np.abs(got[w] - ref) / np.maximum(np.abs(ref), 1.0)
            worst = max(worst, float(err.max()))
            count += len(ref)
            assert err.max() < tol, (jel, w, err.max(), got[w][:4], ref[:4])
    _say(f"  {count} ratios, worst {worst:.2e}")
    assert count > 100


@pytest.mark.parametrize("prec,tol", [(qmcx.FP64, 1e-10), (qmcx.FP32, 2e-3)])
def test_j3_nlpp_evaluate_vs_oracle(cubic_j3, prec, tol):
    """qmcx_nlpp_evaluate with the three-body Jastrow (the reference's -j runs NonLocalPP::evaluate
    with it, miniqmc.cpp:470-495): every work-unit shape, one rotation per walker, against the
    oracle's evaluateRatios of each pair's knots.  Ratios within tol of max(|r|, 1)."""
    from miniqmc_amd import nlpp
    sysd, dev, orc = cubic_j3
    nw, rmax = 2, 1.7
    wb = qmcx.WalkerBatch(dev, nw, delay_rank=1, precision=prec)
    wb.set_positions(systems.initial_positions(sysd, nw, seed=9))
    wb.evaluate_log()
    wb.run(1, 2, 0.3, qmcx.ACCEPT_METROPOLIS, recompute_every=1)
    R = wb.positions()
    for unit in (1, 2, 4, 6):
        _say("J3 NonLocalPP, precision", prec, "knots per unit", unit)
        wb.set_nlpp_unit(unit)
        spheres = nlpp.NonLocalPP(sysd, rmax=rmax, seed=unit).randomize_walkers(nw)
        rat, cnt = wb.nlpp_evaluate(rmax, spheres, max_pairs=4)
        checked = 0
        for w in range(nw):
            iels, vps, slots = [], [], []
            for jel in range(sysd.nel):
                disp, dist = nlpp.electron_ion_displacements(sysd.lattice, sysd.ion_pos, R[w, jel])
                if np.any(np.abs(dist - rmax) < 1e-5):
                    continue
                near = np.nonzero(dist < rmax)[0]
                assert cnt[w, jel] == len(near)
                for p, c in enumerate(near[:4]):
                    iels.append(jel)
                    vps.append(dist[c] * spheres[w] + disp[c] + R[w, jel])
                    slots.append((jel, p))
            ref = orc.evaluate_ratios_many(R[w], iels, np.array(vps))
            for (jel, p), rr in zip(slots, ref):
                err = np.abs(rat[w, jel, p] - rr) / np.maximum(np.abs(rr), 1.0)
                assert err.max() < tol, (unit, w, jel, p, err.max())
                checked += 1
        _say(f"  {checked} pairs checked")
        assert checked > 10
