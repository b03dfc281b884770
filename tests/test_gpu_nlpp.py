"""Virtual-move ratios (WaveFunction::evaluateRatios, src/QMCWaveFunctions/WaveFunction.cpp:281-297)
through qmcx_flex_evaluate_ratios, against the oracle's restatement on the same positions.

The oracle's ratios are pinned on the CPU against exp(log|Psi(R')| - log|Psi(R)|) from two full
evaluateLog calls and against ratioGrad's ratio in a sweep's trace.  Here the GPU batch is compared
after evaluateLog and after sweeps (the inverse rows and the J1/J2 sums then come from the move
kernel's incremental updates), at the NonLocalPP quadrature spheres (src/Drivers/NonLocalPP.hpp:115-140).
"""
import numpy as np
import pytest

from miniqmc_amd import nlpp, qmcx, systems
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _check_all(sysd, orc, wb, R, pp, tol, electrons):
    sphere = pp.randomize_walkers(R.shape[0])
    worst, count = 0.0, 0
    for jel in electrons:
        vpos, npair = pp.virtual_positions(R, jel, sphere)
        if npair.max() == 0:
            continue
        got = wb.flex_evaluate_ratios(jel, vpos)
        for w in range(R.shape[0]):
            ref = orc.evaluate_ratios(R[w], jel, vpos[w])
            err = np.abs(got[w] - ref) / np.maximum(np.abs(ref), 1.0)
            worst = max(worst, float(err.max()))
            count += len(ref)
            assert err.max() < tol, (jel, w, err.max(), got[w][:4], ref[:4])
    return worst, count


@pytest.mark.parametrize("prec,delay,tol", [(qmcx.FP64, 1, 1e-10), (qmcx.FP64, 4, 1e-10), (qmcx.FP32, 4, 2e-3)])
def test_virtual_ratios_nio8_vs_oracle(prec, delay, tol):
    sysd = systems.make_system("nio8-cubic")
    _say("virtual ratios nio8", prec, delay)
    nw = 3
    orc = po.OracleSystem(sysd)
    dev = qmcx.System(sysd)
    wb = qmcx.WalkerBatch(dev, nw, delay_rank=delay, precision=prec)
    wb.set_positions(systems.initial_positions(sysd, nw, seed=7))
    wb.evaluate_log()
    pp = nlpp.NonLocalPP(sysd, rmax=1.7, seed=3)
    worst, count = _check_all(sysd, orc, wb, wb.positions(), pp, tol, range(sysd.nel))
    _say(f"  after evaluateLog: {count} ratios, worst {worst:.2e}")
    assert count > 0
    # after two sweeps the state is the move kernel's
    wb.run(2, 5, 0.4, qmcx.ACCEPT_METROPOLIS)
    worst, count = _check_all(sysd, orc, wb, wb.positions(), pp, tol, range(sysd.nel))
    _say(f"  after sweeps: {count} ratios, worst {worst:.2e}")
    # the electron's own position: ratio one
    R = wb.positions()
    one = wb.flex_evaluate_ratios(2, R[:, 2][:, None, :])
    assert np.abs(one - 1.0).max() < (1e-10 if prec == qmcx.FP64 else 1e-3)


@pytest.mark.parametrize("delay", [16, 32])
def test_virtual_ratios_nio32_fp32_after_sweeps(delay):
    sysd = systems.make_system("nio32")
    _say("virtual ratios nio32 fp32", delay)
    nw = 4
    orc = po.OracleSystem(sysd)
    dev = qmcx.System(sysd)
    wb = qmcx.WalkerBatch(dev, nw, delay_rank=delay, precision=qmcx.FP32)
    wb.set_positions(systems.initial_positions(sysd, nw, seed=9))
    wb.evaluate_log()
    wb.run(1, 3, 0.3, qmcx.ACCEPT_REFERENCE, recompute_every=1)
    pp = nlpp.NonLocalPP(sysd, rmax=1.7, seed=4)
    electrons = list(range(0, sysd.nel, 37)) + [sysd.nup - 1, sysd.nup, sysd.nel - 1]
    worst, count = _check_all(sysd, orc, wb, wb.positions(), pp, 2e-3, electrons)
    _say(f"  {count} ratios, worst {worst:.2e}")
    assert count > 0


def test_nlpp_evaluate_pairs():
    """NonLocalPP::evaluate over the batch: one 12-ratio set per electron-ion pair inside rmax."""
    sysd = systems.make_system("nio8-cubic")
    nw = 2
    dev = qmcx.System(sysd)
    wb = qmcx.WalkerBatch(dev, nw, delay_rank=1, precision=qmcx.FP64)
    wb.set_positions(systems.initial_positions(sysd, nw, seed=7))
    wb.evaluate_log()
    pp = nlpp.NonLocalPP(sysd, rmax=1.7, seed=3)
    out, sphere = pp.evaluate(wb)
    orc = po.OracleSystem(sysd)
    R = wb.positions()
    for w in range(nw):
        expect = []
        for jel in range(sysd.nel):
            disp, dist = nlpp.electron_ion_displacements(sysd.lattice, sysd.ion_pos, R[w, jel])
            for c in np.nonzero(dist < 1.7)[0]:
                expect.append((jel, dist[c] * sphere[w] + disp[c] + R[w, jel]))
        assert len(out[w]) == len(expect) > 0
        for (jel, rat), (jel2, v) in zip(out[w], expect):
            assert jel == jel2
            np.testing.assert_allclose(rat, orc.evaluate_ratios(R[w], jel, v), rtol=1e-10, atol=1e-12)


def test_virtual_ratios_argument_errors():
    sysd = systems.make_system("nio8-cubic")
    dev = qmcx.System(sysd)
    wb = qmcx.WalkerBatch(dev, 2, delay_rank=1, precision=qmcx.FP64)
    wb.set_positions(systems.initial_positions(sysd, 2))
    wb.evaluate_log()
    with pytest.raises(qmcx.QmcxError):
        wb.flex_evaluate_ratios(sysd.nel, np.zeros((2, 12, 3)))
    with pytest.raises(qmcx.QmcxError):
        wb.flex_evaluate_ratios(-1, np.zeros((2, 12, 3)))


def _expected_pairs(sysd, R, jel, rmax, sphere):
    disp, dist = nlpp.electron_ion_displacements(sysd.lattice, sysd.ion_pos, R[jel])
    near = np.nonzero(dist < rmax)[0]
    border = np.any(np.abs(dist - rmax) < 1e-5)
    return [dist[c] * sphere + disp[c] + R[jel] for c in near], border


def _check_device_evaluate(sysd, wb, rmax, spheres, tol, stride, max_pairs=4, walkers=None):
    """qmcx_nlpp_evaluate against the oracle: each electron's ion list inside rmax (ion order) and,
    for every stride-th electron, its 12 ratios per pair at its walker's own knots."""
    rat, cnt = wb.nlpp_evaluate(rmax, spheres, max_pairs=max_pairs)
    R = wb.positions()
    orc = po.OracleSystem(sysd)
    checked = 0
    for w in (range(wb.nw) if walkers is None else walkers):
        iels, vps, slots = [], [], []
        for jel in range(sysd.nel):
            exp_v, border = _expected_pairs(sysd, R[w], jel, rmax, spheres[w])
            if border:
                continue
            assert cnt[w, jel] == len(exp_v), (w, jel)
            if jel % stride or not exp_v:
                continue
            for p, v in enumerate(exp_v):
                iels.append(jel)
                vps.append(v)
                slots.append((jel, p))
        if not iels:
            continue
        ref = orc.evaluate_ratios_many(R[w], iels, np.array(vps))
        for (jel, p), rr in zip(slots, ref):
            err = np.abs(rat[w, jel, p] - rr) / np.maximum(np.abs(rr), 1.0)
            assert err.max() < tol, (w, jel, p, err.max())
            checked += 1
    _say(f"  pairs {int(cnt.sum())}, checked {checked}")
    assert checked > 0
    return rat, cnt


@pytest.mark.parametrize("name,prec,delay,tol,stride", [
    ("nio8-cubic", qmcx.FP64, 4, 1e-10, 1),
    ("nio8-cubic", qmcx.FP32, 4, 2e-3, 1),
    ("nio32", qmcx.FP32, 32, 2e-3, 23),
])
def test_nlpp_device_evaluate_vs_oracle(name, prec, delay, tol, stride):
    """qmcx_nlpp_evaluate with one random rotation per walker (miniqmc.cpp:470)."""
    sysd = systems.make_system(name)
    _say("nlpp device evaluate", name, prec, delay)
    nw, rmax = 3, 1.7
    dev = qmcx.System(sysd)
    wb = qmcx.WalkerBatch(dev, nw, delay_rank=delay, precision=prec)
    wb.set_positions(systems.initial_positions(sysd, nw, seed=13))
    wb.evaluate_log()
    wb.run(1, 9, 0.3, qmcx.ACCEPT_METROPOLIS, recompute_every=1)
    pp = nlpp.NonLocalPP(sysd, rmax=rmax, seed=8)
    spheres = pp.randomize_walkers(nw)
    assert not np.allclose(spheres[0], spheres[1])
    _check_device_evaluate(sysd, wb, rmax, spheres, tol, stride)


@pytest.mark.parametrize("prec", [qmcx.FP64, qmcx.FP32])
@pytest.mark.parametrize("unit", [1, 2, 4, 6])
def test_nlpp_work_unit_shapes_vs_oracle(prec, unit):
    """Every work-unit shape of k_nlpp_ratios (P knots of a pair per unit), forced through the
    tuning knob, against the oracle (NiO-8, 2 walkers)."""
    sysd = systems.make_system("nio8-cubic")
    _say("nlpp work unit", unit, prec)
    dev = qmcx.System(sysd)
    wb = qmcx.WalkerBatch(dev, 2, delay_rank=1, precision=prec)
    wb.set_nlpp_unit(unit)
    wb.set_positions(systems.initial_positions(sysd, 2, seed=21))
    wb.evaluate_log()
    spheres = nlpp.NonLocalPP(sysd, rmax=1.7, seed=unit).randomize_walkers(2)
    _check_device_evaluate(sysd, wb, 1.7, spheres, 1e-10 if prec == qmcx.FP64 else 2e-3, 1)


def test_nlpp_nio64_fp64_vs_oracle():
    """The automatic shape at NiO-64 in FP64 (four knots per unit, 384 orbitals), one walker of two
    checked against the oracle on every 16th electron."""
    sysd = systems.make_system("nio64")
    _say("nlpp nio64 fp64")
    dev = qmcx.System(sysd)
    wb = qmcx.WalkerBatch(dev, 2, delay_rank=1, precision=qmcx.FP64)
    wb.set_positions(systems.initial_positions(sysd, 2, seed=5))
    wb.evaluate_log()
    spheres = nlpp.NonLocalPP(sysd, rmax=1.7, seed=2).randomize_walkers(2)
    _check_device_evaluate(sysd, wb, 1.7, spheres, 1e-10, 16, max_pairs=8, walkers=[1])


def test_nlpp_device_evaluate_overflow_is_an_error():
    sysd = systems.make_system("nio8-cubic")
    dev = qmcx.System(sysd)
    wb = qmcx.WalkerBatch(dev, 2, delay_rank=1, precision=qmcx.FP64)
    wb.set_positions(systems.initial_positions(sysd, 2))
    wb.evaluate_log()
    with pytest.raises(qmcx.QmcxError):
        wb.nlpp_evaluate(5.0, nlpp.quadrature_knots(), max_pairs=1)
    with pytest.raises(qmcx.QmcxError):
        wb.nlpp_evaluate(1.7, nlpp.quadrature_knots(), max_pairs=0)


@pytest.mark.parametrize("prec,tol", [(qmcx.FP64, 1e-10), (qmcx.FP32, 2e-3)])
def test_nlpp_skewed_cell_vs_oracle(prec, tol):
    """A non-orthogonal cell (the 'tiny' system): distances take the general eight-corner search."""
    sysd = systems.make_system("tiny")
    L = sysd.lattice
    vol = abs(np.linalg.det(L))
    width = min(vol / np.linalg.norm(np.cross(L[(i + 1) % 3], L[(i + 2) % 3])) for i in range(3))
    rmax = 0.4 * width
    _say("nlpp skewed", prec, f"rmax {rmax:.3f}")
    nw = 2
    dev = qmcx.System(sysd)
    wb = qmcx.WalkerBatch(dev, nw, delay_rank=1, precision=prec)
    wb.set_positions(systems.initial_positions(sysd, nw, seed=4))
    wb.evaluate_log()
    wb.run(1, 3, 0.4, qmcx.ACCEPT_METROPOLIS)
    orc = po.OracleSystem(sysd)
    pp = nlpp.NonLocalPP(sysd, rmax=rmax, seed=6)
    worst, count = _check_all(sysd, orc, wb, wb.positions(), pp, tol, range(sysd.nel))
    assert count > 0
    sphere = pp.randomize_walkers(nw)
    rat, cnt = wb.nlpp_evaluate(rmax, sphere, max_pairs=8)
    R = wb.positions()
    checked = 0
    for w in range(nw):
        for jel in range(sysd.nel):
            exp_v, border = _expected_pairs(sysd, R[w], jel, rmax, sphere[w])
            if border:
                continue
            assert cnt[w, jel] == len(exp_v)
            for p, v in enumerate(exp_v):
                ref = orc.evaluate_ratios(R[w], jel, v)
                assert (np.abs(rat[w, jel, p] - ref) / np.maximum(np.abs(ref), 1.0)).max() < tol
                checked += 1
    _say(f"  {count} flex ratios (worst {worst:.2e}), {checked} pairs")
    assert checked > 0


@pytest.mark.parametrize("extra", [["-k", "4", "--stack-timers"], ["-k", "1", "--fp64", "--metropolis"],
                                   ["-k", "4", "-j", "-N", "2", "-v"]])
def test_driver_runs(tmp_path, extra):
    """The miniqmc-style driver end to end on a small system: sweeps, NonLocalPP, report, info XML."""
    from miniqmc_amd import driver
    _say("driver", extra)
    rc = driver.main(["--system", "nio8-cubic", "-w", "2", "-n", "2", "--info-dir", str(tmp_path), *extra])
    assert rc == 0
    assert (tmp_path / "info_1_1_1.xml").exists()
