"""Seeded synthetic systems for the VMC particle-move path.

The reference builds its NiO systems in ``src/Input/nio.hpp`` (``tile_cell``
:23-79, ``buildJ1`` :127-168, ``buildJ2`` :82-124) and places electrons
uniformly in the cell (``build_els``, ``src/Particle/ParticleSet_builder.cpp:33-53``).
Its orbital generator makes every orbital of a block the same random array
times ``cos(2*pi*n/N)`` (``src/Numerics/Spline2/BsplineAllocator.hpp:161-182``),
which gives a rank-1 Slater matrix.  Here every orbital and grid point gets an
independent FP32 coefficient drawn from U(-1, 1); the three ghost planes per
axis repeat the first three (periodic), and the padded orbitals are zero.

All randomness comes from numpy's Philox4x64 with an explicit key, so a
system is a pure function of its name and seed.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# nio.hpp:25 — the 32-atom NiO cell (rows are the cell vectors, bohr)
NIO_CELL = np.array([[7.8811, 7.8811, 0.0], [-7.8811, 7.8811, 0.0], [0.0, 0.0, 15.7622]])
# nio.hpp:36-67 — fractional positions, species 0 = O (first 16), 1 = Ni (last 16)
NIO_FRAC = np.array([
    [0.5, 0.0, 0.25], [0.0, 0.0, 0.75], [0.25, 0.25, 0.5], [0.5, 0.5, 0.25],
    [0.75, 0.25, 0.0], [0.0, 0.5, 0.75], [0.25, 0.75, 0.5], [0.75, 0.75, 0.0],
    [0.0, 0.0, 0.25], [0.25, 0.25, 0.0], [0.5, 0.0, 0.75], [0.75, 0.25, 0.5],
    [0.0, 0.5, 0.25], [0.25, 0.75, 0.0], [0.5, 0.5, 0.75], [0.75, 0.75, 0.5],
    [0.0, 0.0, 0.0], [0.0, 0.5, 0.0], [0.25, 0.25, 0.75], [0.5, 0.0, 0.5],
    [0.5, 0.5, 0.5], [0.75, 0.25, 0.25], [0.25, 0.75, 0.75], [0.75, 0.75, 0.25],
    [0.5, 0.0, 0.0], [0.0, 0.0, 0.5], [0.25, 0.25, 0.25], [0.5, 0.5, 0.0],
    [0.75, 0.25, 0.75], [0.0, 0.5, 0.5], [0.25, 0.75, 0.25], [0.75, 0.75, 0.75],
])
NIO_SPECIES = np.array([0] * 16 + [1] * 16, dtype=np.int32)
SPECIES_NAMES = ("O", "Ni")

# nio.hpp:91-121 and :135-166 — the reference's own J2 (uu, ud) and J1 (O, Ni) parameters
NIO_J2_PARAMS = (
    [0.3062333936, 0.2068358231, 0.1324267459, 0.08673342451, 0.05641590391, 0.03626533946,
     0.02269217959, 0.01331043612, 0.006750919735, 0.002827476478, 0.0],
    [0.4397856796, 0.2589475889, 0.1516634434, 0.08962897215, 0.05676614496, 0.03675758961,
     0.02294244601, 0.01343860721, 0.00681803725, 0.002864695148, 0.0],
)
NIO_J1_PARAMS = (
    [-1.060468469, -0.9741652012, -0.8785027166, -0.6350075366, -0.3276892194, -0.08646817952,
     -0.07340983171, -0.0377479213, -0.01507835571, -0.001101691898, 0.0],
    [-3.058259288, -3.077898004, -2.51188807, -1.570208063, -0.6204423266, 0.06272826611,
     0.06228214399, 0.06665099211, 0.03230352443, 0.009254036239, 0.0],
)
J2_CUSP = 0.0    # nio.hpp:104,121
J1_CUSP = -0.25  # nio.hpp:149,166

# nio.hpp:170-244 (buildJeeI) — the reference's three-body (eeI) PolynomialFunctor3D parameters,
# shape (N_eI, N_ee) = (3, 3), 26 free parameters each, keyed by (ion species, spin pair) as the
# reference's addFunc(ion species, 0, 0 | 1) assigns them; cutoff min(Wigner-Seitz radius, 4.8261684030)
NIO_J3_SHAPE = (3, 3)
NIO_J3_RCUT_MAX = 4.8261684030
NIO_J3_PARAMS = {
    (0, 0): [-0.003356164484, 0.002412623253, 0.01653623839, 0.0008341346169, -0.002808360734, 0.000710697475,
             0.01076942152, 0.0009228283355, 0.01576022161, -0.003585259096, 0.003323106938, -0.02282975998,
             -0.002246144403, -0.007196992871, -0.00404316239, 0.001465337212, 0.02026982926, -0.03528735393,
             0.04594087928, -0.008776410679, -0.001552528476, -0.005554407743, 0.001858594451, 0.002001634408,
             0.0009302256139, -0.0006304447229],
    (0, 1): [-0.006861627197, 0.003278047306, 0.03324006545, 0.003097361067, -0.004710623571, 9.652180317e-06,
             0.02212708787, -0.003718893286, 0.03390124932, -0.00710566395, 0.008807743592, -0.04281661568,
             -0.008463011294, -0.01269994613, -0.002005229447, 0.002186590944, 0.03350196472, -0.05677253817,
             0.07810604648, -0.009629896208, -0.006372643712, -0.01056861605, 0.002485188615, 0.008392442289,
             1.073423014e-05, -0.0004812466328],
    (1, 0): [-0.003775082438, -0.00169971229, 0.02162925441, 0.005674020544, -0.0008296047161, 0.00128057705,
             0.005487203215, 0.001637322446, 0.02976838198, -0.0003207100945, 0.01143855436, -0.05336741304,
             -0.00732359381, -0.01556942626, 0.0001149478453, 0.001838601199, 0.02570154203, -0.0675325214,
             0.1080671614, -0.01258358969, 0.001839834045, -0.02422400426, 0.005154953014, 0.003510582598,
             0.007464427016, -0.002454817757],
    (1, 1): [-0.009590393593, 0.002498010871, 0.04225872633, 0.00460311261, -0.01071033503, 0.001253155062,
             0.02934351285, -0.01823794726, 0.07224890393, -0.01020046849, 0.006310807929, -0.05655009412,
             -0.0363775247, 0.002062411388, 0.02037856173, 0.003372676617, 0.03915277249, -0.02680556816,
             0.08648136635, 0.01499088063, -0.02231984329, -0.02399792096, 0.001105720128, 0.02196005181,
             0.003162638982, -0.00119645772],
}


@dataclasses.dataclass
class Functor:
    nparams: int
    rcut: float
    cusp: float
    params: np.ndarray  # float64 [nparams]


@dataclasses.dataclass
class Functor3D:
    """A three-body (eeI) polynomial functor of one ion species and spin pair (0 same, 1 opposite)."""
    species: int
    pair: int
    neI: int
    nee: int
    rcut: float
    params: np.ndarray  # float64, the free parameters


@dataclasses.dataclass
class System:
    name: str
    lattice: np.ndarray        # [3,3] rows = cell vectors
    ion_pos: np.ndarray        # [M,3] Cartesian, grouped by species
    ion_species: np.ndarray    # [M] int32
    nspecies: int
    nup: int
    ndn: int
    ng: tuple                  # spline grid (nx, ny, nz)
    norb: int
    npad: int
    spo_lattice: np.ndarray    # lattice the SPO maps positions with
    coefs: np.ndarray          # float32 [(nx+3)*(ny+3)*(nz+3)*npad]
    j1: list                   # [nspecies] Functor
    j2: list                   # [uu, ud] Functor
    j3: list = dataclasses.field(default_factory=list)   # Functor3D per (ion species, spin pair); empty: no J3

    @property
    def nel(self) -> int:
        return self.nup + self.ndn

    @property
    def nions(self) -> int:
        return len(self.ion_species)

    def table_bytes(self, elem: int = 4) -> int:
        nx, ny, nz = self.ng
        return (nx + 3) * (ny + 3) * (nz + 3) * self.npad * elem


def _rng(seed: int, stream: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=[int(seed) & (2**64 - 1), int(stream)]))


def wigner_seitz_radius(lattice: np.ndarray) -> float:
    """Half the shortest non-zero translation over {-1,0,1}^3 (LatticeAnalyzer.h:86-100)."""
    best = np.inf
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in (-1, 0, 1):
                if i or j or k:
                    v = i * lattice[0] + j * lattice[1] + k * lattice[2]
                    best = min(best, float(v @ v))
    return 0.5 * math.sqrt(best)


def tile_ions(cell: np.ndarray, frac: np.ndarray, species: np.ndarray, tiles=(1, 1, 1)):
    """Supercell expansion by a diagonal tile matrix, grouped by species.

    Follows expandSuperCell (src/Particle/ParticleIOUtility.h:30-117): for each
    species, images i0,i1,i2 of every primitive atom are kept when their
    fractional supercell coordinate lies in [-1e-6, 0.9999).
    """
    tmat = np.diag(tiles).astype(float)
    sup = tmat @ cell
    if tuple(tiles) == (1, 1, 1):
        return cell.copy(), frac @ cell, species.copy()
    inv = np.linalg.inv(sup)
    pos, spc = [], []
    rng = range(-10, 11)
    for s in range(int(species.max()) + 1):
        for i0 in rng:
            for i1 in rng:
                for i2 in rng:
                    for ia in np.nonzero(species == s)[0]:
                        u = frac[ia] - np.floor(frac[ia])
                        r = u @ cell + i0 * cell[0] + i1 * cell[1] + i2 * cell[2]
                        us = r @ inv
                        if np.all(us >= -1.0e-6) and np.all(us < 0.9999):
                            pos.append(r)
                            spc.append(s)
    return sup, np.array(pos), np.array(spc, dtype=np.int32)


def random_functor(rng: np.random.Generator, nparams: int, rcut: float, cusp: float) -> Functor:
    p = rng.uniform(0.0, 0.5, size=nparams)
    p[-1] = 0.0
    return Functor(nparams, rcut, cusp, p)


def spline_coefficients(ng, norb: int, npad: int, seed: int) -> np.ndarray:
    """Independent U(-1,1) FP32 coefficients [nx+3][ny+3][nz+3][npad], periodic ghosts."""
    nx, ny, nz = ng
    rng = _rng(seed, 0x5B1)
    core = rng.random((nx, ny, nz, norb), dtype=np.float32)
    core *= 2.0
    core -= 1.0
    full = np.zeros((nx + 3, ny + 3, nz + 3, npad), dtype=np.float32)
    full[:nx, :ny, :nz, :norb] = core
    del core
    full[nx:, :ny, :nz, :norb] = full[:3, :ny, :nz, :norb]
    full[:, ny:, :nz, :norb] = full[:, :3, :nz, :norb]
    full[:, :, nz:, :norb] = full[:, :, :3, :norb]
    return full.reshape(-1)


def nio_j3(lattice: np.ndarray) -> list:
    """The reference's eeI Jastrow (buildJeeI, nio.hpp:170-244, called with the electrons' Wigner-Seitz
    radius, WaveFunction.cpp:100,137)."""
    rcut = min(wigner_seitz_radius(lattice), NIO_J3_RCUT_MAX)
    neI, nee = NIO_J3_SHAPE
    return [Functor3D(sp, pair, neI, nee, rcut, np.array(p, dtype=float)) for (sp, pair), p in NIO_J3_PARAMS.items()]


def make_system(name: str, seed: int = 11, j_mode: str = "random", j3: bool = False) -> System:
    """Build one of the named workloads.

    name: 'nio8-cubic' (96 e-), 'nio16' (192 e-), 'nio32' (384 e-), 'nio64' (768 e-),
          or 'tiny' (small test system).
    j_mode: 'random' — 8 B-spline parameters U(0,0.5) (last 0), rcut 3.9;
            'nio'    — the reference's own nio.hpp parameter sets and cutoffs.
    j3:     add the reference's three-body eeI Jastrow (miniqmc's -j option).
    """
    if name == "nio8-cubic":
        a = 7.9
        lattice = np.eye(3) * a
        fcc = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
        frac = np.vstack([fcc + 0.5, fcc]) % 1.0   # 4 O then 4 Ni
        species = np.array([0] * 4 + [1] * 4, dtype=np.int32)
        ion_pos = frac @ lattice
        ng, norb, npad = (24, 24, 24), 48, 48
        spo_lattice = lattice.copy()
    elif name in ("nio32", "nio64"):
        tiles = (1, 1, 1) if name == "nio32" else (1, 1, 2)
        lattice, ion_pos, species = tile_ions(NIO_CELL, NIO_FRAC, NIO_SPECIES, tiles)
        if name == "nio32":
            ng, norb, npad = (48, 48, 48), 192, 208
        else:
            ng, norb, npad = (64, 64, 64), 384, 384
        # build_SPOSet receives the primitive cell (miniqmc.cpp:358, nio.hpp:79)
        spo_lattice = NIO_CELL.copy()
    elif name == "nio16":
        lattice = NIO_CELL.copy()
        lattice[2] *= 0.5
        keep = NIO_FRAC[:, 2] < 0.5
        frac = NIO_FRAC[keep].copy()
        frac[:, 2] *= 2.0
        species = NIO_SPECIES[keep]
        order = np.argsort(species, kind="stable")
        frac, species = frac[order], species[order]
        ion_pos = frac @ lattice
        ng, norb, npad = (32, 32, 32), 96, 96
        spo_lattice = lattice.copy()
    elif name == "tiny":
        a = 6.0
        lattice = np.array([[a, 0.3, 0.0], [0.2, a, 0.1], [0.0, 0.4, a]])
        frac = np.array([[0.1, 0.2, 0.3], [0.6, 0.5, 0.7]])
        species = np.array([0, 1], dtype=np.int32)
        ion_pos = frac @ lattice
        ng, norb, npad = (8, 9, 10), 8, 16
        spo_lattice = lattice.copy()
    else:
        raise ValueError(f"unknown system {name!r}")

    nions = len(species)
    nel = 16 if name == "tiny" else nions * 12   # count_electrons: 12 per atom (nio.hpp:17-20)
    nup = nel // 2
    ndn = nel - nup
    if name == "tiny":
        norb = nup
    rng = _rng(seed, 0x1A5)
    if j_mode == "random":
        rcut = 3.9 if name != "tiny" else 2.9
        j1 = [random_functor(rng, 8, rcut, J1_CUSP) for _ in range(2)]
        j2 = [random_functor(rng, 8, rcut, J2_CUSP) for _ in range(2)]
    elif j_mode == "nio":
        ws = wigner_seitz_radius(lattice)
        r2 = min(ws, 5.5727792532)
        r1 = min(ws, 4.8261684030)
        j2 = [Functor(10, r2, J2_CUSP, np.array(p[:10], dtype=float)) for p in NIO_J2_PARAMS]
        j1 = [Functor(10, r1, J1_CUSP, np.array(p[:10], dtype=float)) for p in NIO_J1_PARAMS]
    else:
        raise ValueError(j_mode)
    coefs = spline_coefficients(ng, norb, npad, seed)
    return System(name, lattice, ion_pos, species, 2, nup, ndn, ng, norb, npad, spo_lattice, coefs, j1, j2,
                  nio_j3(lattice) if j3 else [])


def initial_positions(system: System, nwalkers: int, seed: int = 11, walker0: int = 0) -> np.ndarray:
    """Electrons uniform in the cell, one Philox stream per walker: [nw][N][3] (build_els)."""
    out = np.empty((nwalkers, system.nel, 3))
    for w in range(nwalkers):
        u = _rng(seed, 0xE1000 + walker0 + w).random((system.nel, 3))
        out[w] = u @ system.lattice
    return out
