"""Non-local pseudopotential quadrature over a walker batch (src/Drivers/NonLocalPP.hpp).

For every electron and every ion closer than ``rmax`` the reference places the twelve
quadrature knots, randomly rotated, on the sphere through the electron around that ion and
asks the wave function for the twelve virtual-move ratios (NonLocalPP::evaluate,
NonLocalPP.hpp:115-140; driver miniqmc.cpp:470-495).  Here one call per electron covers every
walker: the ratios run on the GPU through ``qmcx_flex_evaluate_ratios``, and walkers with fewer
near ions than the batch's largest count are padded with the electron's own position, whose
ratio is one up to rounding and is dropped.  Every walker rotates its own knots for each
evaluation.  ``NonLocalPP.evaluate`` is this host loop; the batch's ``nlpp_evaluate`` does the same
work in a few device launches.
"""
from __future__ import annotations

import numpy as np

NKNOTS = 12


def quadrature_knots() -> np.ndarray:
    """The twelve vertices of the icosahedron: (+-1, 0, 0) and ten points at x = +-1/sqrt(5) on the
    circle of radius 2/sqrt(5), 36 degrees apart with alternating sign of x (the knot set and
    order of NonLocalPP.hpp:70-81).  Equal weights 1/12."""
    pts = [(1.0, 0.0, 0.0), (-1.0, 0.0, 0.0)]
    for k in range(10):
        a = k * np.pi / 5
        pts.append(((1.0 if k % 2 == 0 else -1.0) / np.sqrt(5), 2 / np.sqrt(5) * np.cos(a),
                    2 / np.sqrt(5) * np.sin(a)))
    return np.array(pts)


def random_rotation(u: np.ndarray) -> np.ndarray:
    """Uniform random rotation from three uniforms in [0, 1) (Shoemake's unit-quaternion method,
    Graphics Gems III, 1992).  Plays the part of NonLocalPP::randomize (NonLocalPP.hpp:100-113),
    whose Euler-angle draw from the reference's own generator is not reproduced."""
    u1, u2, u3 = (float(x) for x in u)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    w, x = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2)
    y, z = b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def electron_ion_displacements(lattice: np.ndarray, ion_pos: np.ndarray, r: np.ndarray):
    """Minimum-image ion - electron displacements [M][3] and distances [M] for one electron,
    searched over the 27 neighbouring images of the wrapped fractional difference."""
    inv = np.linalg.inv(lattice)
    f = (ion_pos - r) @ inv
    f -= np.round(f)
    shifts = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], dtype=float)
    cand = (f[:, None, :] + shifts[None, :, :]) @ lattice          # [M][27][3]
    d2 = np.einsum("mki,mki->mk", cand, cand)
    best = np.argmin(d2, axis=1)
    disp = cand[np.arange(len(f)), best]
    return disp, np.sqrt(d2[np.arange(len(f)), best])


class NonLocalPP:
    """NonLocalPP<T> (NonLocalPP.hpp:30-150) for a batch: knots, random sphere, evaluate."""

    def __init__(self, sysd, rmax: float = 1.7, seed: int = 11):
        self.sys = sysd
        self.rmax = float(rmax)
        self.knots = quadrature_knots()
        self.weight = np.full(NKNOTS, 1.0 / NKNOTS)
        self.rng = np.random.default_rng(seed)

    def randomize(self) -> np.ndarray:
        """Rotated knots [12][3] from one random rotation."""
        return self.knots @ random_rotation(self.rng.random(3)).T

    def randomize_walkers(self, nw: int) -> np.ndarray:
        """Rotated knots [nw][12][3]: one rotation per walker for one evaluation, as each of the
        reference's movers rotates its own quadrature (miniqmc.cpp:470, NonLocalPP.hpp:100-113)."""
        return np.stack([self.randomize() for _ in range(nw)])

    def virtual_positions(self, R: np.ndarray, jel: int, spheres: np.ndarray):
        """For electron jel of every walker (R [nw][N][3]), with walker w's knots spheres[w] (or one
        [12][3] set for all): vpos [nw][npair*12][3] and the number of real (electron, ion) pairs per
        walker; padding points sit at the electron itself."""
        nw = R.shape[0]
        spheres = np.broadcast_to(np.asarray(spheres).reshape(-1, NKNOTS, 3), (nw, NKNOTS, 3))
        rows = []
        for w in range(nw):
            disp, dist = electron_ion_displacements(self.sys.lattice, self.sys.ion_pos, R[w, jel])
            near = np.nonzero(dist < self.rmax)[0]
            # virtualPos[k] = dist * rOnSphere[k] + displ + R[jel]  (NonLocalPP.hpp:133)
            rows.append([dist[c] * spheres[w] + disp[c] + R[w, jel] for c in near])
        npair = np.array([len(r) for r in rows])
        width = max(1, int(npair.max())) * NKNOTS
        vpos = np.repeat(R[:, jel][:, None, :], width, axis=1).copy()
        for w, r in enumerate(rows):
            if r:
                vpos[w, :len(r) * NKNOTS] = np.concatenate(r)
        return vpos, npair

    def evaluate(self, wb, R: np.ndarray | None = None):
        """NonLocalPP::evaluate (NonLocalPP.hpp:115-140) for every walker of ``wb``.  Returns, per
        walker, the list of (electron, ratios[12]) for every electron-ion pair inside rmax, and the
        walkers' rotated knots [nw][12][3]."""
        R = wb.positions() if R is None else R
        sphere = self.randomize_walkers(R.shape[0])
        out = [[] for _ in range(R.shape[0])]
        for jel in range(self.sys.nel):
            vpos, npair = self.virtual_positions(R, jel, sphere)
            if npair.max() == 0:
                continue
            rat = wb.flex_evaluate_ratios(jel, vpos)
            for w in range(R.shape[0]):
                for p in range(npair[w]):
                    out[w].append((jel, rat[w, p * NKNOTS:(p + 1) * NKNOTS]))
        return out, sphere
