"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the CPU baseline, never
as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", _HERE, "-B"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_philox4x64_10.argtypes = [_up, _up, _up]
        L.orc_move_randoms.argtypes = [C.c_uint64] * 4 + [_dp, _dp]
        L.orc_lattice_info.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_to_unit_floor.argtypes = [_dp, C.c_int, _dp, _dp]
        L.orc_distance_row.argtypes = [_dp, _dp, C.c_int, _dp, C.c_int, _dp, _dp]
        L.orc_functor_vgl.argtypes = [C.c_int, C.c_double, C.c_double, _dp, C.c_int, _dp, _dp, _dp, _dp, C.c_void_p]
        L.orc_spline_vgh_unit.argtypes = [_ip, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, _dp, _dp]
        L.orc_invert_transpose.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_invert_transpose.restype = C.c_double
        L.orc_du_create.argtypes = [C.c_int, C.c_int]
        L.orc_du_create.restype = C.c_void_p
        L.orc_du_destroy.argtypes = [C.c_void_p]
        L.orc_du_get_inv_row.argtypes = [C.c_void_p, _dp, C.c_int, _dp]
        L.orc_du_accept_row.argtypes = [C.c_void_p, _dp, C.c_int, _dp]
        L.orc_du_update_inv.argtypes = [C.c_void_p, _dp]
        L.orc_du_delay_count.argtypes = [C.c_void_p]
        L.orc_du_delay_count.restype = C.c_int
        L.orc_system_create.argtypes = [_dp, C.c_int, _dp, _ip, C.c_int, C.c_int, C.c_int]
        L.orc_system_create.restype = C.c_void_p
        L.orc_system_destroy.argtypes = [C.c_void_p]
        L.orc_system_set_j1.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, _dp]
        L.orc_system_set_j2.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, _dp]
        L.orc_system_set_spline.argtypes = [C.c_void_p, _ip, C.c_int, C.c_int, _fp, _dp, C.c_int]
        L.orc_system_ws_radius.argtypes = [C.c_void_p]
        L.orc_system_ws_radius.restype = C.c_double
        L.orc_spo_evaluate.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_run_vmc.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, C.c_int, C.c_uint64, C.c_double, C.c_int,
                                  C.c_int, _dp, _dp, C.c_void_p, C.c_void_p, C.c_void_p,
                                  np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"), C.c_int, C.c_void_p]
        L.orc_run_vmc.restype = C.c_double
        L.orc_evaluate_log.argtypes = [C.c_void_p, _dp, C.c_void_p, C.c_void_p]
        L.orc_evaluate_log.restype = C.c_double
        L.orc_evaluate_ratios.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, _dp, _dp]
        L.orc_evaluate_log_parts.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_evaluate_ratios_many.argtypes = [C.c_void_p, _dp, C.c_int, _ip, C.c_int, _dp, _dp]
        L.orc_system_set_j3.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int]
        L.orc_system_set_j3.restype = C.c_int
        L.orc_poly3d.argtypes = [C.c_int, C.c_int, C.c_double, _dp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_poly3d.restype = C.c_int
        _lib = L
    return _lib


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def philox4x64_10(ctr, key):
    out = np.zeros(4, dtype=np.uint64)
    lib().orc_philox4x64_10(_c(ctr, np.uint64), _c(key, np.uint64), out)
    return out


def move_randoms(seed, walker, sweep, iel):
    n = np.zeros(3)
    u = np.zeros(1)
    lib().orc_move_randoms(seed, walker, sweep, iel, n, u)
    return n, float(u[0])


def lattice_info(R):
    G, rb, g, corners, ws = np.zeros(9), np.zeros(9), np.zeros(9), np.zeros(24), np.zeros(1)
    lib().orc_lattice_info(_c(R).reshape(-1), G, rb, g, corners, ws)
    return dict(G=G.reshape(3, 3), rb=rb.reshape(3, 3), g=g.reshape(3, 3), corners=corners.reshape(8, 3), ws=ws[0])


def to_unit_floor(R, r):
    r = _c(r).reshape(-1, 3)
    u = np.zeros_like(r)
    lib().orc_to_unit_floor(_c(R).reshape(-1), len(r), r, u)
    return u


def distance_row(R, pos, targets, flip_ind=0):
    targets = _c(targets).reshape(-1, 3)
    n = len(targets)
    r, dr = np.zeros(n), np.zeros((n, 3))
    lib().orc_distance_row(_c(R).reshape(-1), _c(pos), n, targets, flip_ind, r, dr)
    return r, dr


def functor_vgl(nparams, rcut, cusp, params, r):
    r = _c(r).reshape(-1)
    u, du, d2u = np.zeros_like(r), np.zeros_like(r), np.zeros_like(r)
    coefs = np.zeros(nparams + 4)
    lib().orc_functor_vgl(nparams, rcut, cusp, _c(params), len(r), r, u, du, d2u, coefs.ctypes.data)
    return u, du, d2u, coefs


def spline_vgh_unit(ng, norb, npad, coefs, u):
    """coefs float32 or float64 table; returns [npts][10][norb]"""
    u = _c(u).reshape(-1, 3)
    prec = 4 if coefs.dtype == np.float32 else 8
    coefs = np.ascontiguousarray(coefs)
    out = np.zeros((len(u), 10, norb))
    lib().orc_spline_vgh_unit(_c(ng, np.int32), norb, npad, prec, coefs.ctypes.data, len(u), u, out)
    return out


def poly3d(neI, nee, rcut, params, r12, r1, r2):
    """The eeI polynomial functor alone: gamma[(neI+1),(neI+1),(nee+1)], and at the points
    (r12, r1, r2): out[npts][9] = F, F_12/r12, F_1/r1, F_2/r2, F_12,12, F_1,1, F_2,2,
    F_12,1/(r12 r1), F_12,2/(r12 r2) (zero where r1 or r2 reaches L), and the value alone."""
    r12, r1, r2 = (_c(x).reshape(-1) for x in (r12, r1, r2))
    params = _c(params)
    gamma = np.zeros((neI + 1) * (neI + 1) * (nee + 1))
    out, v = np.zeros((len(r12), 9)), np.zeros(len(r12))
    rc = lib().orc_poly3d(neI, nee, rcut, params, len(params), len(r12), r12, r1, r2, gamma, out.reshape(-1), v)
    if rc < 0:
        raise ValueError(f"shape ({neI}, {nee}) does not take {len(params)} parameters")
    return gamma.reshape(neI + 1, neI + 1, nee + 1), out, v


def invert_transpose(A):
    A = _c(A)
    n = A.shape[0]
    psiM = np.zeros_like(A)
    ph = np.zeros(1)
    ld = lib().orc_invert_transpose(n, A, psiM, ph)
    return psiM, ld, ph[0]


class DelayedUpdate:
    """The oracle's DelayedUpdate engine (DelayedUpdate.h:29-231)."""

    def __init__(self, norb, delay):
        self.h = lib().orc_du_create(norb, delay)
        self.n = norb

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_du_destroy(self.h)
            self.h = None

    def get_inv_row(self, A, row):
        out = np.zeros(self.n)
        lib().orc_du_get_inv_row(self.h, A, row, out)
        return out

    def accept_row(self, A, row, psiv):
        lib().orc_du_accept_row(self.h, A, row, _c(psiv))

    def update_inv(self, A):
        lib().orc_du_update_inv(self.h, A)

    @property
    def delay_count(self):
        return lib().orc_du_delay_count(self.h)


class OracleSystem:
    """Whole-system oracle: the FP64 restatement of the VMC sweep."""

    def __init__(self, sysd, spline_prec=8):
        L = lib()
        self.sys = sysd
        self.h = L.orc_system_create(_c(sysd.lattice).reshape(-1), sysd.nions, _c(sysd.ion_pos).reshape(-1),
                                     _c(sysd.ion_species, np.int32), sysd.nspecies, sysd.nup, sysd.ndn)
        for s, f in enumerate(sysd.j1):
            L.orc_system_set_j1(self.h, s, f.nparams, f.rcut, f.cusp, _c(f.params))
        for p, f in enumerate(sysd.j2):
            L.orc_system_set_j2(self.h, p, f.nparams, f.rcut, f.cusp, _c(f.params))
        L.orc_system_set_spline(self.h, _c(sysd.ng, np.int32), sysd.norb, sysd.npad, sysd.coefs,
                                _c(sysd.spo_lattice).reshape(-1), spline_prec)
        for f in getattr(sysd, "j3", []) or []:
            if L.orc_system_set_j3(self.h, f.species, f.pair, f.neI, f.nee, f.rcut, _c(f.params), len(f.params)) < 0:
                raise ValueError("J3 functor shape and parameter count disagree")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_system_destroy(self.h)
            self.h = None

    def spo_evaluate(self, r):
        r = _c(r).reshape(-1, 3)
        n = self.sys.norb
        psi, dpsi, d2 = np.zeros((len(r), n)), np.zeros((len(r), n, 3)), np.zeros((len(r), n))
        lib().orc_spo_evaluate(self.h, len(r), r, psi, dpsi, d2)
        return psi, dpsi, d2

    def evaluate_log(self, R):
        N = self.sys.nel
        G, L = np.zeros((N, 3)), np.zeros(N)
        v = lib().orc_evaluate_log(self.h, _c(R).reshape(-1), G.ctypes.data, L.ctypes.data)
        return v, G, L

    def evaluate_log_parts(self, R):
        """evaluateLog per component: logs[5], G[5][N][3], L[5][N] for the up and down
        determinants (SPO-lattice fractional gradients, h_xx), J1, J2 and J3 (Cartesian; zero
        without a J3)."""
        N = self.sys.nel
        logs, G, L = np.zeros(5), np.zeros((5, N, 3)), np.zeros((5, N))
        lib().orc_evaluate_log_parts(self.h, _c(R).reshape(-1), logs, G.reshape(-1), L.reshape(-1))
        return logs, G, L

    def evaluate_ratios(self, R, iel, vpos):
        """WaveFunction::evaluateRatios of electron iel at vpos[nvp][3] for configuration R[N][3]."""
        vpos = _c(vpos).reshape(-1, 3)
        out = np.zeros(len(vpos))
        lib().orc_evaluate_ratios(self.h, _c(R).reshape(-1), iel, len(vpos), vpos.reshape(-1), out)
        return out

    def evaluate_ratios_many(self, R, iels, vpos):
        """evaluate_ratios for several electrons of configuration R with one evaluateLog:
        electron iels[q] at vpos[q][nvp][3] -> ratios[q][nvp]."""
        iels = _c(iels, np.int32).reshape(-1)
        vpos = _c(vpos).reshape(len(iels), -1, 3)
        out = np.zeros((len(iels), vpos.shape[1]))
        lib().orc_evaluate_ratios_many(self.h, _c(R).reshape(-1), len(iels), iels, vpos.shape[1], vpos.reshape(-1),
                                       out.reshape(-1))
        return out

    def run_vmc(self, R, nsweeps, seed, tau, mode=1, delay=1, walker0=0, trace=False, nthreads=0,
                sensitivity=False):
        """sensitivity (delay 1): per move, the relative sensitivity of the determinant ratio to
        relative perturbations of the Slater matrix and the new row, [nw][nsweeps][N]."""
        R = _c(R).copy()
        nw, N = R.shape[0], self.sys.nel
        log_init = np.zeros(nw)
        log_out = np.zeros((nw, nsweeps))
        tr = np.zeros((nw, nsweeps, N, 9)) if trace else None
        sens = np.zeros((nw, nsweeps, N)) if sensitivity else None
        G, L = np.zeros((nw, N, 3)), np.zeros((nw, N))
        nacc = np.zeros(nw, dtype=np.int64)
        secs = lib().orc_run_vmc(self.h, nw, walker0, R.reshape(-1), nsweeps, seed, tau, mode, delay, log_init,
                                 log_out, tr.ctypes.data if trace else None, G.ctypes.data, L.ctypes.data, nacc,
                                 nthreads, sens.ctypes.data if sensitivity else None)
        return dict(R=R, log_init=log_init, log=log_out, trace=tr, G=G, L=L, naccept=nacc, seconds=secs,
                    sensitivity=sens)
