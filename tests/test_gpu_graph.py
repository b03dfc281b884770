"""Graph replay of the sweep (qmcx_batch_set_graph): a run of sweeps is captured into one HIP graph
(its three streams' launches and event dependencies) and replayed; the random streams' sweep numbers
and the log history slots come from device counters.  Replays must continue the chains exactly as
launching the kernels does: two batches from the same start, one run eagerly and one by graph
replay, agree bit for bit after every run (positions, log|Psi| history, G, L, accept counts), with
the host's own eager work and synchronization between the replays.  Against the FP64 oracle as well.
"""
import numpy as np
import pytest

from miniqmc_amd import qmcx, systems
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name,prec,delay,recompute,form,rows", [
    ("nio8-cubic", qmcx.FP32, 8, 1, 0, 0), ("nio8-cubic", qmcx.FP64, 1, 0, 0, 0), ("nio32", qmcx.FP32, 32, 1, 0, 0),
    # the streamed delayed-update form (move and flush launches per chunk) and the FP64 recompute
    # cadence (rounds of move launches and inversions of the stopped walkers, device-side positions)
    ("nio8-cubic", qmcx.FP32, 16, 1, 1, 0), ("nio8-cubic", qmcx.FP64, 8, 0, 1, 0),
    ("nio8-cubic", qmcx.FP32, 8, 0, 0, 6), ("nio32", qmcx.FP32, 32, 0, 0, 16)])
def test_graph_replays_equal_launches(name, prec, delay, recompute, form, rows):
    sysd = systems.make_system(name)
    dev = qmcx.System(sysd)
    nw, seed, tau, ns = 4, 11, 0.3, 3
    R0 = systems.initial_positions(sysd, nw, seed=seed)
    eager = qmcx.WalkerBatch(dev, nw, delay_rank=delay, precision=prec)
    graph = qmcx.WalkerBatch(dev, nw, delay_rank=delay, precision=prec)
    graph.set_graph(True)
    for b in (eager, graph):
        b.set_delay_form(form)
        b.set_recompute_rows(rows)
        b.set_positions(R0)
        b.evaluate_log()
    for call in range(3):
        _say(f"{name} prec {prec} delay {delay} form {form} rows {rows}: run {call}")
        he = eager.run(ns, seed, tau, qmcx.ACCEPT_METROPOLIS, recompute_every=recompute)
        hg = graph.run(ns, seed, tau, qmcx.ACCEPT_METROPOLIS, recompute_every=recompute)
        np.testing.assert_array_equal(hg, he)
        np.testing.assert_array_equal(graph.positions(), eager.positions())
        np.testing.assert_array_equal(graph.accepts(), eager.accepts())
        Ge, Le, le = eager.gl()
        Gg, Lg, lg = graph.gl()
        np.testing.assert_array_equal(Gg, Ge)
        np.testing.assert_array_equal(Lg, Le)
        # eager work on the graph batch between replays
        np.testing.assert_allclose(graph.evaluate_log(), lg, rtol=1e-5)
        eager.evaluate_log()
    if prec == qmcx.FP64 and delay == 1:
        # and the chain is the oracle's: 9 sweeps from R0
        ref = po.OracleSystem(sysd).run_vmc(R0, 3 * ns, seed, tau, mode=qmcx.ACCEPT_METROPOLIS, delay=1)
        np.testing.assert_allclose(graph.positions(), ref["R"], rtol=0, atol=1e-11)
        np.testing.assert_allclose(hg[-1], ref["log"][:, -1], rtol=1e-10)


def test_graph_recaptures_on_new_parameters():
    sysd = systems.make_system("nio8-cubic")
    dev = qmcx.System(sysd)
    R0 = systems.initial_positions(sysd, 2, seed=3)
    a = qmcx.WalkerBatch(dev, 2, delay_rank=4, precision=qmcx.FP32)
    b = qmcx.WalkerBatch(dev, 2, delay_rank=4, precision=qmcx.FP32)
    b.set_graph(True)
    for w in (a, b):
        w.set_positions(R0)
        w.evaluate_log()
    for ns, mode in ((2, qmcx.ACCEPT_REFERENCE), (1, qmcx.ACCEPT_METROPOLIS), (2, qmcx.ACCEPT_REFERENCE)):
        np.testing.assert_array_equal(b.run(ns, 7, 0.5, mode, recompute_every=1),
                                      a.run(ns, 7, 0.5, mode, recompute_every=1))
    # a setting change drops the graph; tracing then applies to the next capture
    b.set_trace(True)
    a.set_trace(True)
    np.testing.assert_array_equal(b.run(1, 7, 0.5, qmcx.ACCEPT_METROPOLIS), a.run(1, 7, 0.5, qmcx.ACCEPT_METROPOLIS))
    np.testing.assert_array_equal(b.trace(), a.trace())
