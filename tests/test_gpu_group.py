"""Mixed-size groups (qmcx_group_*): walkers of NiO-16, NiO-32 and NiO-64 batches (each its own
system: spline table, Jastrows, size, delay rank) advanced by one launch per move stage and per
proposals stage.

  * the group reproduces each batch run on its own: identical accept sequences and positions (the
    reference driver's u < 0.5 acceptance; positions move by the same FP64 displacements) and
    log|Psi| within 1e-6 relative (only the reduction order over the larger workgroup differs);
  * against the FP64 oracle per system: log|Psi| after each sweep's FP64 refresh within 1e-6
    relative, the oracle's accept counts and positions (FP64 in both forms);
  * a graph-captured group run equals the launched one bit for bit.
"""
import numpy as np
import pytest

from miniqmc_amd import qmcx, systems
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

SEED, TAU = 13, 0.3


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def mixed():
    out = []
    for name, k, nw in (("nio16", 16, 3), ("nio32", 32, 2), ("nio64", 32, 2)):
        sysd = systems.make_system(name)
        out.append((name, sysd, qmcx.System(sysd, precisions=(qmcx.FP32,)), k, nw))
    return out


def _batches(mixed, offsets=True):
    bs, R0s = [], []
    off = 0
    for name, sysd, dev, k, nw in mixed:
        b = qmcx.WalkerBatch(dev, nw, delay_rank=k, precision=qmcx.FP32, walker0=off if offsets else 0)
        R0 = systems.initial_positions(sysd, nw, seed=SEED, walker0=off)
        b.set_positions(R0)
        b.evaluate_log()
        bs.append(b)
        R0s.append(R0)
        off += nw
    return bs, R0s


def test_group_equals_separate_runs(mixed):
    grp_b, R0s = _batches(mixed)
    sep_b, _ = _batches(mixed)
    g = qmcx.Group(grp_b)
    for call in range(2):
        _say("group run", call)
        hg = g.run(2, SEED, TAU, qmcx.ACCEPT_REFERENCE, recompute_every=1)
        for i, (name, sysd, dev, k, nw) in enumerate(mixed):
            hs = sep_b[i].run(2, SEED, TAU, qmcx.ACCEPT_REFERENCE, recompute_every=1)
            np.testing.assert_array_equal(grp_b[i].accepts(), sep_b[i].accepts())
            np.testing.assert_array_equal(grp_b[i].positions(), sep_b[i].positions())
            err = np.abs(hg[i] - hs) / np.abs(hs)
            _say(f"  {name}: accepts {grp_b[i].accepts().tolist()}, log err vs separate {err.max():.2e}")
            assert err.max() < 1e-6


def test_group_vs_oracle(mixed):
    grp_b, R0s = _batches(mixed)
    g = qmcx.Group(grp_b)
    hg = g.run(2, SEED, TAU, qmcx.ACCEPT_REFERENCE, recompute_every=1)
    _say(f"group of {g.nw} walkers: {g.last_ms():.2f} ms for 2 sweeps")
    off = 0
    for i, (name, sysd, dev, k, nw) in enumerate(mixed):
        ref = po.OracleSystem(sysd).run_vmc(R0s[i], 2, SEED, TAU, mode=qmcx.ACCEPT_REFERENCE, delay=1, walker0=off)
        err = np.abs(hg[i] - ref["log"].T) / np.abs(ref["log"].T)
        _say(f"  {name}: log err vs the FP64 oracle {err.max():.2e}")
        assert err.max() < 1e-6
        np.testing.assert_allclose(grp_b[i].positions(), ref["R"], rtol=0, atol=1e-11)
        np.testing.assert_array_equal(grp_b[i].accepts(), ref["naccept"])
        off += nw


def test_group_graph_equals_launches(mixed):
    a_b, _ = _batches(mixed)
    b_b, _ = _batches(mixed)
    ga, gb = qmcx.Group(a_b), qmcx.Group(b_b)
    gb.set_graph(True)
    for call in range(3):
        ha = ga.run(2, SEED, TAU, qmcx.ACCEPT_METROPOLIS, recompute_every=1)
        hb = gb.run(2, SEED, TAU, qmcx.ACCEPT_METROPOLIS, recompute_every=1)
        for i in range(len(mixed)):
            np.testing.assert_array_equal(hb[i], ha[i])
            np.testing.assert_array_equal(b_b[i].positions(), a_b[i].positions())


def test_group_rejects_unsupported_batches(mixed):
    name, sysd, dev, k, nw = mixed[0]
    fp64_sys = qmcx.System(sysd, precisions=(qmcx.FP64,))
    with pytest.raises(qmcx.QmcxError, match="FP32"):
        qmcx.Group([qmcx.WalkerBatch(fp64_sys, 1, delay_rank=4, precision=qmcx.FP64)])
    with pytest.raises(qmcx.QmcxError, match="delay rank"):
        qmcx.Group([qmcx.WalkerBatch(dev, 1, delay_rank=1, precision=qmcx.FP32)])


def test_group_fp64_rows_vs_oracle(mixed):
    """Batches keeping FP64 Slater rows (the grouped proposals also sum the FP64 values): log|Psi|
    after each sweep's FP64 refresh within 1e-6 of the FP64 oracle (the determinant's part carries
    no FP32 rounding of the matrix; the Jastrow sums stay FP32); the oracle's chains."""
    grp_b, R0s = _batches(mixed)
    for b in grp_b:
        b.set_fp64_rows(True)
        b.evaluate_log()
    g = qmcx.Group(grp_b)
    hg = g.run(2, SEED, TAU, qmcx.ACCEPT_REFERENCE, recompute_every=1)
    off = 0
    for i, (name, sysd, dev, k, nw) in enumerate(mixed):
        ref = po.OracleSystem(sysd).run_vmc(R0s[i], 2, SEED, TAU, mode=qmcx.ACCEPT_REFERENCE, delay=1, walker0=off)
        err = np.abs(hg[i] - ref["log"].T) / np.abs(ref["log"].T)
        _say(f"  {name} with FP64 rows: log err vs the FP64 oracle {err.max():.2e}")
        assert err.max() < 1e-6
        np.testing.assert_allclose(grp_b[i].positions(), ref["R"], rtol=0, atol=1e-11)
        off += nw
    # all or none
    grp_b[0].set_fp64_rows(False)
    with pytest.raises(qmcx.QmcxError, match="FP64 Slater rows"):
        g.run(1, SEED, TAU, qmcx.ACCEPT_REFERENCE, recompute_every=1)
