"""The batched flex_* stage loop (src/Drivers/miniqmc_sync_move.cpp:459-513) through the C ABI.

Each electron of every walker: flex_ratio_grad with caller displacements, the caller's
Metropolis decision, flex_accept_restore; then flex_completeUpdates and flex_evaluateGL.
Displacements are the oracle's Philox stream, so the loop must retrace the oracle's sweep:
same ratios and gradients, same accept sequence, same final positions and log|Psi|.
"""
import numpy as np
import pytest

from miniqmc_amd import qmcx, systems
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cubic():
    return systems.make_system("nio8-cubic")


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("prec,delay,tol", [(qmcx.FP64, 1, 1e-10), (qmcx.FP64, 4, 1e-10), (qmcx.FP64, 48, 1e-10),
                                            (qmcx.FP32, 4, 2e-3), (qmcx.FP32, 16, 2e-3)])
def test_flex_loop_vs_oracle(cubic, prec, delay, tol):
    """With delay rank k > 1 an accepted flex move stays a pending delayed row of its walker until the
    walker has k of them (DelayedUpdate.h:151-152) or flex_completeUpdates; later ratios read the
    pending rows through getInvRow (k = 48: nothing is flushed before completeUpdates in the first
    determinant's 48 moves)."""
    _say("flex loop", prec, delay)
    nw, seed, tau = 3, 11, 0.5
    R0 = systems.initial_positions(cubic, nw, seed=seed)
    orc = po.OracleSystem(cubic)
    ref = orc.run_vmc(R0, 1, seed, tau, mode=qmcx.ACCEPT_METROPOLIS, delay=1, trace=True)
    dev = qmcx.System(cubic)
    wb = qmcx.WalkerBatch(dev, nw, delay_rank=delay, precision=prec)
    wb.set_positions(R0)
    wb.evaluate_log()
    N = cubic.nel
    flips = 0
    for iel in range(N):
        displ = np.zeros((nw, 3))
        uacc = np.zeros(nw)
        for w in range(nw):
            g, u = po.move_randoms(seed, w, 0, iel)
            displ[w] = np.sqrt(tau) * g
            uacc[w] = u
        ratio, gnow, gnew = wb.flex_ratio_grad(iel, displ)
        rt = ref["trace"][:, 0, iel]
        rel = np.abs(ratio - rt[:, 0]) / np.maximum(np.abs(rt[:, 0]), 1.0)
        assert rel.max() < tol, ("ratio", iel, rel.max())
        assert (np.abs(gnow - rt[:, 3:6]) / (1 + np.abs(rt[:, 3:6]))).max() < tol * 100
        assert (np.abs(gnew - rt[:, 6:9]) / (1 + np.abs(rt[:, 6:9]))).max() < tol * 100
        accepted = (uacc < ratio * ratio).astype(np.int32)
        flips += int((accepted != rt[:, 2]).sum())
        # follow the oracle's decisions so that a near-tie cannot fork the two runs
        wb.flex_accept_restore(iel, rt[:, 2].astype(np.int32))
    wb.flex_complete_updates()
    log = wb.flex_evaluate_gl()
    _say(f"  flips {flips}, log {log} oracle {ref['log'][:, 0]}")
    assert flips == 0
    np.testing.assert_allclose(wb.positions(), ref["R"], rtol=0, atol=1e-12 if prec == qmcx.FP64 else 1e-5)
    np.testing.assert_allclose(log, ref["log"][:, 0], rtol=1e-11 if prec == qmcx.FP64 else 1e-6)
    np.testing.assert_array_equal(wb.accepts(), ref["naccept"])


def test_flex_state_errors(cubic):
    dev = qmcx.System(cubic)
    wb = qmcx.WalkerBatch(dev, 2, delay_rank=1, precision=qmcx.FP64)
    wb.set_positions(systems.initial_positions(cubic, 2))
    wb.evaluate_log()
    with pytest.raises(qmcx.QmcxError):
        wb.flex_accept_restore(0, np.ones(2, np.int32))   # no pending ratio_grad
    with pytest.raises(qmcx.QmcxError):
        wb.flex_ratio_grad(cubic.nel, np.zeros((2, 3)))     # electron out of range


def test_flex_pending_rows_applied_before_other_calls(cubic):
    """Delayed rows left pending by flex moves (delay rank 16, a dozen accepts) reach the inverse before
    any call that reads it: the inverse read back, and a following sweep, equal those of a batch whose
    pending rows were applied by flex_completeUpdates first (the same kernels, so bit for bit)."""
    nw, seed, tau = 2, 5, 0.5
    R0 = systems.initial_positions(cubic, nw, seed=seed)
    dev = qmcx.System(cubic)
    a = qmcx.WalkerBatch(dev, nw, delay_rank=16, precision=qmcx.FP64)
    b = qmcx.WalkerBatch(dev, nw, delay_rank=16, precision=qmcx.FP64)
    for wb in (a, b):
        wb.set_positions(R0)
        wb.evaluate_log()
        for iel in range(12):
            displ = np.array([np.sqrt(tau) * np.asarray(po.move_randoms(seed, w, 0, iel)[0]) for w in range(nw)])
            wb.flex_ratio_grad(iel, displ)
            wb.flex_accept_restore(iel, np.ones(nw, np.int32))
    b.flex_complete_updates()
    for w in range(nw):
        np.testing.assert_array_equal(a.inverse(w, 0), b.inverse(w, 0))
    np.testing.assert_array_equal(a.run(1, seed, tau, qmcx.ACCEPT_METROPOLIS), b.run(1, seed, tau, qmcx.ACCEPT_METROPOLIS))
    np.testing.assert_array_equal(a.positions(), b.positions())
