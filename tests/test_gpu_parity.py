"""HIP path vs the FP64 CPU oracle, through the C ABI (libqmcx.so).

Every comparison uses the same seeded inputs on both sides.  Tolerances:
  * FP64 kernels: relative 1e-10 or tighter (only summation order differs);
  * FP32 kernels: stated per test (FP32 spline sums, FP32 Woodbury) and reported;
  * accept/reject sequences: identical.
"""
import numpy as np
import pytest

from miniqmc_amd import qmcx, systems
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def cubic():
    return systems.make_system("nio8-cubic")


@pytest.fixture(scope="module")
def cubic_dev(cubic):
    return qmcx.System(cubic)


@pytest.fixture(scope="module")
def cubic_orc(cubic):
    return po.OracleSystem(cubic)


def _rand_points(sysd, npts, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-0.3, 1.3, size=(npts, 3))   # includes points outside the cell
    return u @ sysd.lattice


@pytest.mark.parametrize("prec,tol", [(qmcx.FP64, 1e-12), (qmcx.FP32, 2e-5)])
def test_spline_vgh_vs_oracle(cubic, cubic_dev, prec, tol):
    _say("spline VGH", prec)
    r = _rand_points(cubic, 64, 3)
    gpu = cubic_dev.spline_vgh(r, prec)
    u = po.to_unit_floor(cubic.spo_lattice, r)
    ref = po.spline_vgh_unit(cubic.ng, cubic.norb, cubic.npad, cubic.coefs, u)
    for p in range(10):
        scale = np.abs(ref[:, p]).max()
        err = np.abs(gpu[:, p] - ref[:, p]).max() / scale
        assert err < tol, (p, err)


@pytest.mark.parametrize("prec,tol", [(qmcx.FP64, 1e-12), (qmcx.FP32, 2e-5)])
def test_distance_rows_vs_oracle(cubic, cubic_dev, prec, tol):
    _say("distance rows", prec)
    rng = np.random.default_rng(5)
    src = rng.uniform(-0.5, 1.5, size=(8, 3)) @ cubic.lattice
    tgt = rng.uniform(0, 1, size=(40, 3)) @ cubic.lattice
    flips = np.array([0, 3, 7, 40, 12, 0, 39, 20], dtype=np.int32)
    r, dr = cubic_dev.distance_rows(src, flips, tgt, prec)
    for p in range(len(src)):
        rr, dd = po.distance_row(cubic.lattice, src[p], tgt, int(flips[p]))
        assert np.abs(r[p] - rr).max() < tol * 10
        assert np.abs(dr[p] - dd).max() < tol * 10


def test_invert_transpose_kat():
    # test_dirac_matrix.cpp:69-107: log|det| 3.78518913425 and the known inverse
    _say("LU KAT")
    a = np.array([[2.3, 4.5, 2.6], [0.5, 8.5, 3.3], [1.8, 4.4, 4.9]])
    psiM, ld, ph = qmcx.invert_transpose(a.T.copy())
    b = np.array([[0.6159749342, -0.2408954682, -0.1646081192],
                  [0.07923894288, 0.1496231042, -0.1428117337],
                  [-0.2974298429, -0.04586322768, 0.3927890292]])
    assert abs(ld[0] - 3.78518913425) < 1e-10
    assert np.abs(psiM[0] - b).max() < 1e-9


@pytest.mark.parametrize("n", [4, 8, 24, 96, 384, 600])
@pytest.mark.parametrize("ties", [False, True])
def test_invert_transpose_batched_vs_oracle(n, ties):
    """Regression (n < 32): the LDS row that a pivot swap parks must hold a whole panel row.  384 is
    NiO-64's determinant (32-column panels), 600 takes 16-column panels; with ties the entries are
    small integers, so most pivot searches meet equal magnitudes in several rows and waves (the
    first such row must win, as idamax picks it) and the matrices are still almost surely regular."""
    _say("LU batched n =", n, "ties", ties)
    rng = np.random.default_rng(7)
    a = rng.integers(-3, 4, size=(5, n, n)).astype(np.float64) if ties else rng.uniform(-1, 1, size=(5, n, n))
    psiM, ld, ph = qmcx.invert_transpose(a)
    for m in range(5):
        ref, rld, rph = po.invert_transpose(a[m])
        assert abs(ld[m] - rld) < 1e-10 * abs(rld) + 1e-10
        assert ph[m] == rph
        assert np.abs(psiM[m] - ref).max() < 1e-9 * np.abs(ref).max()


def _compare_run(sysd, dev, orc, nw, nsweeps, delay, prec, mode, tau, tol_ratio, tol_log, seed=11, flush_kind=0,
                 tile=0, delay_form=0, band=0.0, recompute_rows=0):
    R0 = systems.initial_positions(sysd, nw, seed=seed)
    ref = orc.run_vmc(R0, nsweeps, seed, tau, mode=mode, delay=delay, trace=True)
    wb = qmcx.WalkerBatch(dev, nw, delay_rank=delay, precision=prec, tile=tile)
    wb.set_flush(flush_kind)
    wb.set_delay_form(delay_form)
    assert wb.delay_form() == delay_form
    if band > 0.0:
        wb.set_accept_band(band)
    if recompute_rows:
        wb.set_recompute_rows(recompute_rows)
    wb.set_positions(R0)
    log0 = wb.evaluate_log()
    np.testing.assert_allclose(log0, ref["log_init"], rtol=tol_log)
    wb.set_trace(True)
    flips = 0
    for s in range(nsweeps):
        hist = wb.run(1, seed, tau, mode)
        # the counter-based stream is indexed by the absolute sweep number
        tr = wb.trace()
        rt = ref["trace"][:, s]
        flips += int((tr[:, :, 2] != rt[:, :, 2]).sum())
        if flips == 0:
            # scaled by max(|ratio|, 1): near-zero ratios carry the FP32 cancellation error of the dot product
            rel = np.abs(tr[:, :, 0] - rt[:, :, 0]) / np.maximum(np.abs(rt[:, :, 0]), 1.0)
            _say(f"  sweep {s}: max ratio err {rel.max():.3e}, median {np.median(rel):.3e}")
            for idx in np.argsort(rel.ravel())[-3:]:
                w, e = np.unravel_index(idx, rel.shape)
                _say(f"    walker {w} electron {e}: gpu {tr[w, e, :3]} oracle {rt[w, e, :3]}")
            assert rel.max() < tol_ratio, ("ratio", s, rel.max())
            g_err = np.abs(tr[:, :, 3:9] - rt[:, :, 3:9]) / (1 + np.abs(rt[:, :, 3:9]))
            assert g_err.max() < tol_ratio * 100, ("grad", s, g_err.max())
        lerr = np.abs(hist[0] - ref["log"][:, s]) / np.abs(ref["log"][:, s])
        _say(f"  sweep {s}: max |dlog|/|log| {lerr.max():.3e}")
        np.testing.assert_allclose(hist[0], ref["log"][:, s], rtol=tol_log)
    G, L, lv = wb.gl()
    return dict(flips=flips, ref=ref, G=G, L=L, log=lv, acc=wb.accepts(), R=wb.positions())


@pytest.mark.parametrize("prec", [qmcx.FP64, qmcx.FP32])
@pytest.mark.parametrize("tile", [1, 2, 4, 8, 16])
def test_sweep_spline_tiles_vs_oracle(cubic, cubic_dev, cubic_orc, prec, tile):
    """Every stencil split of the proposals kernel (1 slice: partial sums kept in registers and
    stored directly; 2-16 slices: reduced through LDS): FP64 ratios within 1e-9, FP32 within 2e-3
    of max(|r|, 1), identical accept sequences."""
    _say("sweep tile", tile, "precision", prec)
    fp64 = prec == qmcx.FP64
    out = _compare_run(cubic, cubic_dev, cubic_orc, nw=2, nsweeps=1, delay=1 if fp64 else 8, prec=prec,
                       mode=qmcx.ACCEPT_METROPOLIS, tau=0.5, tol_ratio=1e-9 if fp64 else 2e-3,
                       tol_log=1e-11 if fp64 else 1e-6, tile=tile)
    assert out["flips"] == 0
    assert np.array_equal(out["acc"], out["ref"]["naccept"])


@pytest.mark.parametrize("delay", [1, 4, 16])
def test_sweep_fp64_vs_oracle(cubic, cubic_dev, cubic_orc, delay):
    """FP64: a move launch per chunk of k electrons, each followed by a flush launch"""
    _say("sweep FP64 delay", delay)
    out = _compare_run(cubic, cubic_dev, cubic_orc, nw=3, nsweeps=2, delay=delay, prec=qmcx.FP64,
                       mode=qmcx.ACCEPT_METROPOLIS, tau=0.5, tol_ratio=1e-9, tol_log=1e-11)
    assert out["flips"] == 0
    ref = out["ref"]
    assert np.array_equal(out["acc"], ref["naccept"])
    np.testing.assert_allclose(out["R"], ref["R"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(out["G"], ref["G"], rtol=1e-8, atol=1e-8)
    np.testing.assert_allclose(out["L"], ref["L"], rtol=1e-8, atol=1e-7)


@pytest.mark.parametrize("delay,kind", [(1, 0), (8, 0), (32, 0), (32, 2), (8, 1), (16, 1)])
def test_sweep_fp32_vs_fp64_oracle(cubic, cubic_dev, cubic_orc, delay, kind):
    """kind 0: one move launch per determinant, each chunk's updates applied by the move kernel
    on the matrix cores (new Slater rows staged in LDS when at most 16 moves were accepted);
    kind 2: the same with those rows always read from HBM; kind 1: a move launch per chunk and
    a vector-FMA flush launch"""
    _say("sweep FP32 delay", delay, "flush kind", kind)
    # Metropolis acceptance: the FP32 path must reproduce the FP64 accept/reject sequence;
    # log|Psi| within 1e-6 relative of the FP64 oracle
    out = _compare_run(cubic, cubic_dev, cubic_orc, nw=3, nsweeps=2, delay=delay, prec=qmcx.FP32,
                       mode=qmcx.ACCEPT_METROPOLIS, tau=0.5, tol_ratio=2e-3, tol_log=1e-6, flush_kind=kind)
    assert out["flips"] == 0


@pytest.mark.parametrize("prec,delay,kind,band", [(qmcx.FP64, 4, 0, 0.0), (qmcx.FP64, 16, 0, 0.0),
                                                  (qmcx.FP32, 8, 0, 0.0), (qmcx.FP32, 32, 0, 1e-2),
                                                  (qmcx.FP32, 48, 1, 0.0)])
def test_sweep_streamed_vs_oracle(cubic, cubic_dev, cubic_orc, prec, delay, kind, band):
    """The streamed form of the delayed updates (the form of delay ranks whose stale rows do not fit
    the LDS): stale inverse rows read from HBM, p = U . Ainv[row] formed at each move, the capacitance
    inverse in the chunk's state slot, a flush launch per chunk (matrix cores, or vector FMA with kind
    1).  The same bar as the LDS-resident form: FP64 ratios within 1e-9 and log|Psi| within 1e-11 of
    the oracle with identical accept sequences, positions, G and L; FP32 ratios within 2e-3 and
    log|Psi| within 1e-6, identical accept sequences (delay 32 with the near-tie band)."""
    _say("sweep streamed form, precision", prec, "delay", delay, "flush kind", kind, "band", band)
    fp64 = prec == qmcx.FP64
    out = _compare_run(cubic, cubic_dev, cubic_orc, nw=3, nsweeps=2, delay=delay, prec=prec,
                       mode=qmcx.ACCEPT_METROPOLIS, tau=0.5, tol_ratio=1e-9 if fp64 else 2e-3,
                       tol_log=1e-11 if fp64 else 1e-6, flush_kind=kind, delay_form=1, band=band)
    assert out["flips"] == 0
    ref = out["ref"]
    assert np.array_equal(out["acc"], ref["naccept"])
    if fp64:
        np.testing.assert_allclose(out["R"], ref["R"], rtol=0, atol=1e-12)
        np.testing.assert_allclose(out["G"], ref["G"], rtol=1e-8, atol=1e-8)
        np.testing.assert_allclose(out["L"], ref["L"], rtol=1e-8, atol=1e-7)


@pytest.mark.parametrize("prec,delay,rows", [(qmcx.FP64, 1, 4), (qmcx.FP64, 1, 16), (qmcx.FP32, 1, 8),
                                             (qmcx.FP32, 8, 4), (qmcx.FP32, 16, 16), (qmcx.FP32, 32, 5)])
def test_sweep_recompute_cadence_vs_oracle(cubic, cubic_dev, cubic_orc, prec, delay, rows):
    """FP64 recompute of a walker's inverse after every `rows` accepted moves of a determinant (each
    walker stops there, its inverse and log|det| are rebuilt in FP64, and it continues from its own
    electron in the next round), no per-sweep refresh.  FP64: the oracle's sweep to 1e-9 in the ratios
    and 1e-11 in log|Psi|, identical accepts, positions, G and L (the resumed positions, Jastrow sums
    and log|det| replacement are exact).  FP32: ratios within 2e-3, log|Psi| within 1e-6."""
    _say("sweep recompute cadence", prec, "delay", delay, "rows", rows)
    fp64 = prec == qmcx.FP64
    out = _compare_run(cubic, cubic_dev, cubic_orc, nw=3, nsweeps=2, delay=delay, prec=prec,
                       mode=qmcx.ACCEPT_METROPOLIS, tau=0.5, tol_ratio=1e-9 if fp64 else 2e-3,
                       tol_log=1e-11 if fp64 else 1e-6, recompute_rows=rows)
    assert out["flips"] == 0
    ref = out["ref"]
    assert np.array_equal(out["acc"], ref["naccept"])
    np.testing.assert_allclose(out["R"], ref["R"], rtol=0, atol=1e-12 if fp64 else 1e-5)
    if fp64:
        np.testing.assert_allclose(out["G"], ref["G"], rtol=1e-8, atol=1e-8)
        np.testing.assert_allclose(out["L"], ref["L"], rtol=1e-8, atol=1e-7)


def test_nio32_recompute_cadence_reference_rule(nio32):
    """NiO-32 batched FP32 at delay 32 with the FP64 recompute every 16 accepted rows and no per-sweep
    refresh (the reference driver's u < 0.5): log|Psi| within 1e-7 of the FP64 oracle after each of two
    sweeps, identical accepts."""
    sysd, dev, orc = nio32
    R0 = systems.initial_positions(sysd, 2, seed=5)
    ref = orc.run_vmc(R0, 2, 5, 0.3, mode=qmcx.ACCEPT_REFERENCE, delay=1, trace=True)
    wb = qmcx.WalkerBatch(dev, 2, delay_rank=32, precision=qmcx.FP32)
    wb.set_recompute_rows(16)
    wb.set_positions(R0)
    wb.evaluate_log()
    wb.set_trace(True)
    for s in range(2):
        hist = wb.run(1, 5, 0.3, qmcx.ACCEPT_REFERENCE, recompute_every=0)
        assert np.array_equal(wb.trace()[:, :, 2], ref["trace"][:, s][:, :, 2])
        lerr = np.abs(hist[0] - ref["log"][:, s]) / np.abs(ref["log"][:, s])
        _say(f"  sweep {s}: max |dlog|/|log| {lerr.max():.3e}")
        assert lerr.max() < 1e-7
    assert np.array_equal(wb.accepts(), ref["naccept"])


@pytest.mark.parametrize("delay,kind", [(32, 0), (32, 2), (16, 0)])
def test_sweep_fp32_reference_rule_vs_oracle(nio32, delay, kind):
    """The reference driver's acceptance (u < 0.5, about half the moves accepted): chunks of 32
    with more than 16 accepts stage their new Slater rows beyond the 16th in the rejected moves'
    LDS slots (kind 0) or read them from HBM (kind 2).  Accept sequences match by construction;
    log|Psi| within 1e-6 of the FP64 oracle after every sweep (NiO-32, 2 walkers, 2 sweeps)."""
    sysd, dev, orc = nio32
    _say("sweep FP32 reference rule delay", delay, "flush kind", kind)
    R0 = systems.initial_positions(sysd, 2, seed=5)
    ref = orc.run_vmc(R0, 2, 5, 0.3, mode=qmcx.ACCEPT_REFERENCE, delay=delay, trace=True)
    wb = qmcx.WalkerBatch(dev, 2, delay_rank=delay, precision=qmcx.FP32)
    wb.set_flush(kind)
    wb.set_positions(R0)
    wb.evaluate_log()
    wb.set_trace(True)
    for s in range(2):
        hist = wb.run(1, 5, 0.3, qmcx.ACCEPT_REFERENCE, recompute_every=1)
        tr = wb.trace()
        assert np.array_equal(tr[:, :, 2], ref["trace"][:, s][:, :, 2])
        lerr = np.abs(hist[0] - ref["log"][:, s]) / np.abs(ref["log"][:, s])
        _say(f"  sweep {s}: max |dlog|/|log| {lerr.max():.3e}")
        assert lerr.max() < 1e-6
    assert np.array_equal(wb.accepts(), ref["naccept"])


@pytest.mark.parametrize("prec,delay", [(qmcx.FP64, 1), (qmcx.FP32, 4)])
def test_sweep_skewed_cell_vs_oracle(prec, delay):
    """A non-orthogonal cell (the 'tiny' system): the minimum image takes the general
    eight-corner search in the J2 rows rather than the orthogonal-cell shortcut."""
    _say("sweep skewed cell", prec, "delay", delay)
    sysd = systems.make_system("tiny")
    dev, orc = qmcx.System(sysd), po.OracleSystem(sysd)
    if prec == qmcx.FP64:
        out = _compare_run(sysd, dev, orc, nw=2, nsweeps=2, delay=delay, prec=prec, mode=qmcx.ACCEPT_METROPOLIS,
                           tau=0.5, tol_ratio=1e-9, tol_log=1e-11)
        np.testing.assert_allclose(out["G"], out["ref"]["G"], rtol=1e-8, atol=1e-8)
    else:
        out = _compare_run(sysd, dev, orc, nw=2, nsweeps=2, delay=delay, prec=prec, mode=qmcx.ACCEPT_METROPOLIS,
                           tau=0.5, tol_ratio=2e-3, tol_log=1e-6)
    assert out["flips"] == 0


@pytest.fixture(scope="module")
def nio32():
    sysd = systems.make_system("nio32")
    return sysd, qmcx.System(sysd), po.OracleSystem(sysd)


@pytest.mark.parametrize("prec,tol", [(qmcx.FP64, 1e-12), (qmcx.FP32, 1e-6)])
def test_evaluate_log_nio32_vs_oracle(nio32, prec, tol):
    """evaluateLog from scratch (SPO rows, FP64 LU, J1/J2 recompute) at NiO-32, both precisions.
    Regression: the J recompute kernel's LDS must hold the minimum-image cell."""
    sysd, dev, orc = nio32
    R0 = systems.initial_positions(sysd, 3, seed=9)
    wb = qmcx.WalkerBatch(dev, 3, delay_rank=1, precision=prec)
    wb.set_positions(R0)
    got = wb.evaluate_log()
    G, L, _ = wb.gl()
    for w in range(3):
        v, Gr, Lr = orc.evaluate_log(R0[w])
        assert abs(got[w] - v) < tol * abs(v)
        assert np.abs(G[w] - Gr).max() < 1e3 * tol * (1 + np.abs(Gr).max())


def test_sweep_nio32_fp64_vs_oracle(nio32):
    """NiO-32 plain drop-in form: 1 walker, FP64, delay 1 (384 e-, 192 orbitals padded to 208)."""
    _say("NiO-32 FP64 delay 1")
    sysd, dev, orc = nio32
    out = _compare_run(sysd, dev, orc, nw=1, nsweeps=1, delay=1, prec=qmcx.FP64, mode=qmcx.ACCEPT_METROPOLIS,
                       tau=0.3, tol_ratio=1e-8, tol_log=1e-11)
    assert out["flips"] == 0


def _flip_margins(tr, rt, seed, sweep, walker0=0):
    """Relative margin |u - ratio^2| / ratio^2 of the oracle at every move where the
    accept bits differ (the acceptance uniform is regenerated from the stream)."""
    out = []
    for w, e in zip(*np.nonzero(tr[:, :, 2] != rt[:, :, 2])):
        _, u = po.move_randoms(seed, walker0 + int(w), sweep, int(e))
        p = rt[w, e, 0] ** 2
        out.append((int(w), int(e), abs(u - p) / max(p, 1e-300)))
    return out


def test_flush_mfma_vs_oracle(nio32):
    """FP32 Woodbury flush on the matrix cores (U rows from LDS or HBM) and on vector FMA against
    the FP64 oracle (NiO-32, delay 32, Metropolis, the FP32 form's FP64 inverse recompute once
    per sweep).  A walker's accept sequence may only part from the oracle's at a near tie
    (oracle margin |u - ratio^2| < 1e-3 ratio^2, the FP32 ratio error band); a walker without
    one ends the sweep with log|Psi| within 1e-6 of the oracle."""
    sysd, dev, orc = nio32
    R0 = systems.initial_positions(sysd, 2, seed=3)
    ref = orc.run_vmc(R0, 1, 5, 0.3, mode=qmcx.ACCEPT_METROPOLIS, delay=32, trace=True)
    for kind in (0, 1, 2):
        _say("flush kind", kind)
        wb = qmcx.WalkerBatch(dev, 2, delay_rank=32, precision=qmcx.FP32)
        wb.set_flush(kind)
        wb.set_positions(R0)
        wb.evaluate_log()
        wb.set_trace(True)
        hist = wb.run(1, 5, 0.3, qmcx.ACCEPT_METROPOLIS, recompute_every=1)
        tr = wb.trace()
        rt = ref["trace"][:, 0]
        flips = _flip_margins(tr, rt, 5, 0)
        first = {}
        for w, e, m in flips:
            first.setdefault(w, (e, m))
        _say(f"  first flips (walker: electron, oracle margin): {first}")
        for w, (e, m) in first.items():
            assert m < 1e-3, (w, e, m)
        for w in range(2):
            end = first.get(w, (tr.shape[1], 0))[0]
            rel = np.abs(tr[w, :end, 0] - rt[w, :end, 0]) / np.maximum(np.abs(rt[w, :end, 0]), 1)
            _say(f"  walker {w}: max ratio err before any flip {rel.max() if end else 0:.2e}")
            if w not in first:
                lerr = abs(hist[0][w] - ref["log"][w, 0]) / abs(ref["log"][w, 0])
                _say(f"  walker {w}: |dlog|/|log| {lerr:.2e}")
                assert lerr < 1e-6


def test_recompute_matches_fresh_evaluate_log(nio32):
    """The FP64 inverse recompute at the end of a sweep gives the inverse and log|det| that a
    fresh evaluateLog of the same positions gives (DiracDeterminant.cpp:237-280), in both
    precisions."""
    sysd, dev, orc = nio32
    R0 = systems.initial_positions(sysd, 3, seed=9)
    for prec, tol in ((qmcx.FP64, 1e-10), (qmcx.FP32, 1e-6)):
        _say("recompute vs fresh evaluateLog", prec)
        wb = qmcx.WalkerBatch(dev, 3, delay_rank=16 if prec == qmcx.FP32 else 1, precision=prec)
        wb.set_positions(R0)
        wb.evaluate_log()
        hist = wb.run(2, 7, 0.3, qmcx.ACCEPT_METROPOLIS, recompute_every=1)
        fresh = qmcx.WalkerBatch(dev, 3, delay_rank=1, precision=qmcx.FP64)
        fresh.set_positions(wb.positions())
        log_f = fresh.evaluate_log()
        lerr = np.abs(hist[1] - log_f) / np.abs(log_f)
        ierr = 0.0
        for w in range(3):
            for det in range(2):
                ref_inv = fresh.inverse(w, det)
                ierr = max(ierr, np.abs(wb.inverse(w, det) - ref_inv).max() / np.abs(ref_inv).max())
        _say(f"  |dlog|/|log| {lerr.max():.2e}, inverse {ierr:.2e}")
        assert lerr.max() < tol
        assert ierr < (1e-9 if prec == qmcx.FP64 else 1e-3)


def test_recompute_overlapped_history_vs_oracle(cubic, cubic_dev, cubic_orc):
    """Several sweeps in one call with the FP64 refresh every sweep: the second determinant's
    refresh lands during the next sweep and corrects the recorded log|Psi| of its own sweep.
    Every sweep's recorded value must match the oracle's (FP64, delay 4, Metropolis)."""
    _say("overlapped recompute history")
    R0 = systems.initial_positions(cubic, 3, seed=11)
    ref = cubic_orc.run_vmc(R0, 3, 11, 0.5, mode=qmcx.ACCEPT_METROPOLIS, delay=4, trace=False)
    wb = qmcx.WalkerBatch(cubic_dev, 3, delay_rank=4, precision=qmcx.FP64)
    wb.set_positions(R0)
    wb.evaluate_log()
    hist = wb.run(3, 11, 0.5, qmcx.ACCEPT_METROPOLIS, recompute_every=1)
    np.testing.assert_allclose(hist, ref["log"].T, rtol=1e-10)
    np.testing.assert_allclose(wb.positions(), ref["R"], rtol=0, atol=1e-12)
    G, L, lv = wb.gl()
    np.testing.assert_allclose(lv, ref["log"][:, -1], rtol=1e-10)
