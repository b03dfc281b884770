"""Python host layer over libqmcx.so (the C ABI declared in include/qmcx.h).

Mirrors the reference's object roles: ``System`` holds what ``build_ions``,
``buildJ1/buildJ2`` and ``build_SPOSet`` set up (src/Input/nio.hpp,
src/QMCWaveFunctions/SPOSet_builder.cpp); ``WalkerBatch`` holds the per-walker
``ParticleSet`` + ``WaveFunction`` state that the batched ``flex_*`` calls and
the VMC sweep act on (src/Drivers/miniqmc_sync_move.cpp:459-513).

There is no fallback: if libqmcx.so is missing or a call fails, an exception
is raised.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqmcx.so")

FP64 = 0
FP32 = 1
ACCEPT_REFERENCE = 0
ACCEPT_METROPOLIS = 1

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i = C.c_int

# name -> argtypes (all return int status)
_SIGS = {
    "qmcx_device_count": [C.POINTER(C.c_int)],
    "qmcx_system_create": [C.POINTER(_vp), _dp, _i, _dp, _ip, _i, _i, _i],
    "qmcx_system_destroy": [_vp],
    "qmcx_system_set_j1": [_vp, _i, _i, C.c_double, C.c_double, _dp],
    "qmcx_system_set_j2": [_vp, _i, _i, C.c_double, C.c_double, _dp],
    "qmcx_system_set_j3": [_vp, _i, _i, _i, _i, C.c_double, _dp, _i],
    "qmcx_system_set_spline": [_vp, _ip, _i, _i, _fp, _dp, _i],
    "qmcx_spline_vgh": [_vp, _i, _i, _dp, _dp],
    "qmcx_distance_rows": [_vp, _i, _i, _dp, _ip, _i, _dp, _dp, _dp],
    "qmcx_invert_transpose": [_i, _i, _dp, _dp, _dp, _dp],
    "qmcx_invert_transpose_noalloc": [_i, _i, _vp, _vp, _vp, _vp, _vp],
    "qmcx_batch_create": [C.POINTER(_vp), _vp, _i, _i, _i],
    "qmcx_batch_destroy": [_vp],
    "qmcx_batch_set_positions": [_vp, _dp],
    "qmcx_batch_get_positions": [_vp, _dp],
    "qmcx_batch_set_walker_offset": [_vp, _i],
    "qmcx_batch_set_tile": [_vp, _i],
    "qmcx_batch_set_flush": [_vp, _i],
    "qmcx_batch_set_delay_form": [_vp, _i],
    "qmcx_batch_set_recompute_rows": [_vp, _i],
    "qmcx_batch_get_delay_form": [_vp, C.POINTER(C.c_int)],
    "qmcx_batch_set_graph": [_vp, _i],
    "qmcx_evaluate_log": [_vp, _vp],
    "qmcx_run_sweeps": [_vp, _i, C.c_uint64, C.c_double, _i, _i, _vp],
    "qmcx_set_trace": [_vp, _i],
    "qmcx_get_trace": [_vp, _dp],
    "qmcx_get_gl": [_vp, _vp, _vp, _vp],
    "qmcx_get_accepts": [_vp, _lp],
    "qmcx_batch_set_accept_band": [_vp, C.c_double],
    "qmcx_batch_set_fp64_rows": [_vp, C.c_int],
    "qmcx_get_arbitrations": [_vp, _lp],
    "qmcx_get_inverse": [_vp, _i, _i, _dp],
    "qmcx_last_sweep_ms": [_vp, C.POINTER(C.c_double)],
    "qmcx_set_kernel_timing": [_vp, _i],
    "qmcx_flex_ratio_grad": [_vp, _i, _dp, _vp, _vp, _vp],
    "qmcx_flex_accept_restore": [_vp, _i, _ip],
    "qmcx_flex_complete_updates": [_vp],
    "qmcx_flex_evaluate_gl": [_vp, _vp],
    "qmcx_flex_evaluate_ratios": [_vp, _i, _i, _dp, _vp],
    "qmcx_nlpp_evaluate": [_vp, C.c_double, _dp, _i, _vp, _vp],
    "qmcx_batch_set_nlpp_unit": [_vp, _i],
    "qmcx_du_create": [C.POINTER(_vp), _i, _i, _i],
    "qmcx_du_destroy": [_vp],
    "qmcx_du_invert_transpose": [_vp, _dp, _vp, _vp, _vp],
    "qmcx_du_set_inverse": [_vp, _dp],
    "qmcx_du_get_inverse": [_vp, _dp],
    "qmcx_du_get_inv_row": [_vp, _i, _dp],
    "qmcx_du_accept_row": [_vp, _i, _dp],
    "qmcx_du_update_inv_mat": [_vp],
    "qmcx_du_delay_count": [_vp, C.POINTER(C.c_int)],
    "qmcx_get_kernel_timing": [_vp, _dp, _lp],
    "qmcx_group_create": [C.POINTER(_vp), C.POINTER(_vp), _i],
    "qmcx_group_destroy": [_vp],
    "qmcx_group_set_graph": [_vp, _i],
    "qmcx_group_run_sweeps": [_vp, _i, C.c_uint64, C.c_double, _i, _i, _vp],
    "qmcx_group_last_ms": [_vp, C.POINTER(C.c_double)],
}


class QmcxError(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Load libqmcx.so; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise QmcxError(f"{path} is not built; run `python -m miniqmc_amd.build`")
        L = C.CDLL(path)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.qmcx_last_error.restype = C.c_char_p
        L.qmcx_last_error.argtypes = []
        _lib = L
    return _lib


def exported_symbols():
    return sorted(list(_SIGS) + ["qmcx_last_error"])


def _check(rc: int):
    if rc != 0:
        raise QmcxError(f"qmcx error {rc}: {load().qmcx_last_error().decode()}")


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def device_count() -> int:
    c = C.c_int(0)
    _check(load().qmcx_device_count(C.byref(c)))
    return c.value


def invert_transpose(a: np.ndarray):
    """Batched FP64 LU inverse: a[nmat][n][n] -> psiM = (a^-1)^T, logdet, phase."""
    a = _c(a)
    if a.ndim == 2:
        a = a[None]
    nmat, n, _ = a.shape
    psiM = np.zeros_like(a)
    ld, ph = np.zeros(nmat), np.zeros(nmat)
    _check(load().qmcx_invert_transpose(n, nmat, a, psiM, ld, ph))
    return psiM, ld, ph


class DelayedUpdate:
    """Device-resident DU_TYPE engine of one determinant (DelayedUpdate.h:29-231 interface)."""

    def __init__(self, norb: int, delay: int, precision: int = FP64):
        h = _vp()
        _check(load().qmcx_du_create(C.byref(h), norb, delay, precision))
        self.h, self.n = h, norb

    def close(self):
        if getattr(self, "h", None):
            load().qmcx_du_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def invert_transpose(self, logdetT):
        """logdetT[i][j] = phi_j(r_i) -> (psiM = (logdetT^-1)^T, log|det|, phase)."""
        a = _c(logdetT).reshape(self.n, self.n)
        out = np.zeros_like(a)
        lv, ph = C.c_double(0), C.c_double(0)
        _check(load().qmcx_du_invert_transpose(self.h, a.reshape(-1), out.ctypes.data, C.byref(lv), C.byref(ph)))
        return out, lv.value, ph.value

    def set_inverse(self, Ainv):
        _check(load().qmcx_du_set_inverse(self.h, _c(Ainv).reshape(-1)))

    def inverse(self):
        out = np.zeros((self.n, self.n))
        _check(load().qmcx_du_get_inverse(self.h, out.reshape(-1)))
        return out

    def get_inv_row(self, row):
        out = np.zeros(self.n)
        _check(load().qmcx_du_get_inv_row(self.h, row, out))
        return out

    def accept_row(self, row, psiV):
        _check(load().qmcx_du_accept_row(self.h, row, _c(psiV)))

    def update_inv_mat(self):
        _check(load().qmcx_du_update_inv_mat(self.h))

    @property
    def delay_count(self):
        c = C.c_int(0)
        _check(load().qmcx_du_delay_count(self.h, C.byref(c)))
        return c.value


class System:
    """Device-resident system description (lattice, ions, J1/J2 (and J3) functors, SPO table)."""

    def __init__(self, sysd, precisions=(FP32, FP64)):
        L = load()
        self.desc = sysd
        h = _vp()
        _check(L.qmcx_system_create(C.byref(h), _c(sysd.lattice).reshape(-1), sysd.nions,
                                    _c(sysd.ion_pos).reshape(-1), _c(sysd.ion_species, np.int32), sysd.nspecies,
                                    sysd.nup, sysd.ndn))
        self.h = h
        for s, f in enumerate(sysd.j1):
            _check(L.qmcx_system_set_j1(h, s, f.nparams, f.rcut, f.cusp, _c(f.params)))
        for p, f in enumerate(sysd.j2):
            _check(L.qmcx_system_set_j2(h, p, f.nparams, f.rcut, f.cusp, _c(f.params)))
        for f in getattr(sysd, "j3", []) or []:   # ThreeBodyJastrow::addFunc
            _check(L.qmcx_system_set_j3(h, f.species, f.pair, f.neI, f.nee, f.rcut, _c(f.params), len(f.params)))
        mask = (1 if FP32 in precisions else 0) | (2 if FP64 in precisions else 0)
        _check(L.qmcx_system_set_spline(h, _c(sysd.ng, np.int32), sysd.norb, sysd.npad,
                                        np.ascontiguousarray(sysd.coefs, dtype=np.float32),
                                        _c(sysd.spo_lattice).reshape(-1), mask))

    def close(self):
        if getattr(self, "h", None):
            load().qmcx_system_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def spline_vgh(self, r, precision=FP64):
        """VGH of every orbital at Cartesian points: [npts][10][norb]."""
        r = _c(r).reshape(-1, 3)
        out = np.zeros((len(r), 10, self.desc.norb))
        _check(load().qmcx_spline_vgh(self.h, precision, len(r), r, out))
        return out

    def distance_rows(self, src, flip_ind, targets, precision=FP64):
        src = _c(src).reshape(-1, 3)
        targets = _c(targets).reshape(-1, 3)
        r = np.zeros((len(src), len(targets)))
        dr = np.zeros((len(src), len(targets), 3))
        _check(load().qmcx_distance_rows(self.h, precision, len(src), src, _c(flip_ind, np.int32), len(targets),
                                         targets, r, dr))
        return r, dr


class WalkerBatch:
    """Walker state on the device and the batched VMC sweep."""

    def __init__(self, system: System, nwalkers: int, delay_rank: int = 1, precision: int = FP64,
                 walker0: int = 0, tile: int = 0):
        L = load()
        self.system = system
        self.nw = nwalkers
        self.N = system.desc.nel
        self.n = system.desc.nup
        h = _vp()
        _check(L.qmcx_batch_create(C.byref(h), system.h, nwalkers, delay_rank, precision))
        self.h = h
        if walker0:
            _check(L.qmcx_batch_set_walker_offset(h, walker0))
        if tile:
            _check(L.qmcx_batch_set_tile(h, tile))

    def close(self):
        if getattr(self, "h", None):
            load().qmcx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_flush(self, kind):
        """0: matrix-core (MFMA) flush when possible, 1: vector-FMA flush."""
        _check(load().qmcx_batch_set_flush(self.h, kind))

    def set_delay_form(self, form):
        """0: automatic (LDS-resident delayed rows when they fit), 1: the streamed form."""
        _check(load().qmcx_batch_set_delay_form(self.h, form))

    def set_recompute_rows(self, rows):
        """FP64 recompute of a walker's inverse after every `rows` accepted rows of a determinant (0: off)."""
        _check(load().qmcx_batch_set_recompute_rows(self.h, rows))

    def delay_form(self):
        f = C.c_int(0)
        _check(load().qmcx_batch_get_delay_form(self.h, C.byref(f)))
        return f.value

    def set_positions(self, R):
        R = _c(R).reshape(self.nw, self.N, 3)
        _check(load().qmcx_batch_set_positions(self.h, R.reshape(-1)))

    def positions(self):
        R = np.zeros((self.nw, self.N, 3))
        _check(load().qmcx_batch_get_positions(self.h, R.reshape(-1)))
        return R

    def evaluate_log(self):
        out = np.zeros(self.nw)
        _check(load().qmcx_evaluate_log(self.h, out.ctypes.data))
        return out

    def set_trace(self, on=True):
        _check(load().qmcx_set_trace(self.h, 1 if on else 0))

    def trace(self):
        t = np.zeros((self.nw, self.N, 9))
        _check(load().qmcx_get_trace(self.h, t.reshape(-1)))
        return t

    def run(self, nsweeps, seed, tau, mode=ACCEPT_METROPOLIS, recompute_every=0):
        hist = np.zeros((max(nsweeps, 1), self.nw))
        _check(load().qmcx_run_sweeps(self.h, nsweeps, seed, tau, mode, recompute_every, hist.ctypes.data))
        return hist[:nsweeps]

    def gl(self):
        G, Lp, lv = np.zeros((self.nw, self.N, 3)), np.zeros((self.nw, self.N)), np.zeros(self.nw)
        _check(load().qmcx_get_gl(self.h, G.ctypes.data, Lp.ctypes.data, lv.ctypes.data))
        return G, Lp, lv

    def accepts(self):
        a = np.zeros(self.nw, dtype=np.int64)
        _check(load().qmcx_get_accepts(self.h, a))
        return a

    def set_graph(self, enable=True):
        """Capture run_sweeps into a HIP graph on first use and replay it (qmcx_batch_set_graph)."""
        _check(load().qmcx_batch_set_graph(self.h, int(bool(enable))))

    def set_accept_band(self, band):
        """FP32 Metropolis: decide moves with |u - R^2| <= band R^2 on the FP64 ratio (0: off)."""
        _check(load().qmcx_batch_set_accept_band(self.h, float(band)))

    def set_fp64_rows(self, enable=True):
        """FP32 form: keep FP64 Slater rows for the FP64 refresh (evaluate_log fills them)."""
        _check(load().qmcx_batch_set_fp64_rows(self.h, int(bool(enable))))

    def arbitrations(self):
        a = np.zeros(self.nw, dtype=np.int64)
        _check(load().qmcx_get_arbitrations(self.h, a))
        return a

    def inverse(self, walker, det):
        p = np.zeros((self.n, self.n))
        _check(load().qmcx_get_inverse(self.h, walker, det, p.reshape(-1)))
        return p

    # ---- flex_* stage calls (miniqmc_sync_move.cpp:459-513)
    def flex_ratio_grad(self, iel, displ):
        """setActive + evalGrad + makeMove + ratioGrad for electron iel of every walker."""
        ratio, gnow, gnew = np.zeros(self.nw), np.zeros((self.nw, 3)), np.zeros((self.nw, 3))
        _check(load().qmcx_flex_ratio_grad(self.h, iel, _c(displ).reshape(-1), ratio.ctypes.data, gnow.ctypes.data,
                                           gnew.ctypes.data))
        return ratio, gnow, gnew

    def flex_accept_restore(self, iel, accepted):
        _check(load().qmcx_flex_accept_restore(self.h, iel, _c(accepted, np.int32)))

    def flex_complete_updates(self):
        _check(load().qmcx_flex_complete_updates(self.h))

    def flex_evaluate_gl(self):
        out = np.zeros(self.nw)
        _check(load().qmcx_flex_evaluate_gl(self.h, out.ctypes.data))
        return out

    def flex_evaluate_ratios(self, iel, vpos):
        """WaveFunction::evaluateRatios for electron iel of every walker at the virtual positions
        vpos[nw][nvp][3]; returns ratios[nw][nvp]."""
        vpos = _c(vpos).reshape(self.nw, -1, 3)
        out = np.zeros((self.nw, vpos.shape[1]))
        _check(load().qmcx_flex_evaluate_ratios(self.h, iel, vpos.shape[1], vpos.reshape(-1), out.ctypes.data))
        return out

    def set_nlpp_unit(self, knots_per_unit):
        """Knots per NonLocalPP work unit (1, 2, 4, 6; 0 = automatic)."""
        _check(load().qmcx_batch_set_nlpp_unit(self.h, knots_per_unit))

    def nlpp_evaluate(self, rmax, spheres, max_pairs=4, copy_out=True, ratios=True):
        """NonLocalPP::evaluate for every electron of every walker on the GPU, each walker with its
        own knots spheres[nw][12][3] (one [12][3] set is given to every walker): ratios
        [nw][N][max_pairs][12] and npairs [nw][N] (None, None without copy_out; None, npairs
        without ratios)."""
        sphere = _c(np.broadcast_to(_c(spheres).reshape(-1, 12, 3), (self.nw, 12, 3)))
        if not copy_out:
            _check(load().qmcx_nlpp_evaluate(self.h, float(rmax), sphere.reshape(-1), max_pairs, None, None))
            return None, None
        out = np.zeros((self.nw, self.N, max_pairs, 12)) if ratios else None
        cnt = np.zeros((self.nw, self.N), dtype=np.int32)
        _check(load().qmcx_nlpp_evaluate(self.h, float(rmax), sphere.reshape(-1), max_pairs,
                                         out.ctypes.data if ratios else None, cnt.ctypes.data))
        return out, cnt

    def set_kernel_timing(self, on=True):
        _check(load().qmcx_set_kernel_timing(self.h, 1 if on else 0))

    def kernel_timing(self):
        """{stage: (ms, launches)} accumulated since timing was enabled, for the stages
        move, flush, recompute (FP64 inverse), gl (evaluateGL) and proposals (SPO rows at the
        sweep's proposals)."""
        ms, n = np.zeros(5), np.zeros(5, dtype=np.int64)
        _check(load().qmcx_get_kernel_timing(self.h, ms, n))
        return {k: (float(ms[i]), int(n[i]))
                for i, k in enumerate(("move", "flush", "recompute", "gl", "proposals"))}

    def last_sweep_ms(self):
        ms = C.c_double(0)
        _check(load().qmcx_last_sweep_ms(self.h, C.byref(ms)))
        return ms.value


class Group:
    """Walkers of several batches (systems of different sizes) advanced together: one launch per move
    stage and per proposals stage over all of them (qmcx_group_*)."""

    def __init__(self, batches):
        self.batches = list(batches)
        arr = (_vp * len(self.batches))(*[b.h for b in self.batches])
        h = _vp()
        _check(load().qmcx_group_create(C.byref(h), arr, len(self.batches)))
        self.h = h
        self.nw = sum(b.nw for b in self.batches)

    def close(self):
        if getattr(self, "h", None):
            load().qmcx_group_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_graph(self, enable=True):
        _check(load().qmcx_group_set_graph(self.h, int(bool(enable))))

    def run(self, nsweeps, seed, tau, mode=ACCEPT_METROPOLIS, recompute_every=0):
        """Returns the log|Psi| history per batch: a list of [nsweeps][nw_b] arrays."""
        hist = np.zeros((max(nsweeps, 1), self.nw))
        _check(load().qmcx_group_run_sweeps(self.h, nsweeps, seed, tau, mode, recompute_every, hist.ctypes.data))
        out, o = [], 0
        for b in self.batches:
            out.append(hist[:nsweeps, o:o + b.nw].copy())
            o += b.nw
        return out

    def last_ms(self):
        ms = C.c_double(0)
        _check(load().qmcx_group_last_ms(self.h, C.byref(ms)))
        return ms.value
