"""Seeded synthetic structures and potentials for the five benchmark workloads.

Conventions follow the reference: Cartesian positions in Angstrom, lattice
vectors as the COLUMNS of the cell matrix (src/format/xyz_struc.cpp:123-131).
Basis parameters are given in the reference's own units: a radial GAUSSIAN
width eta in 1/Angstrom (exp(-(eta (r-rs))^2), src/nnp/basis_radial.cpp:267-272)
and an angular GAUSS width eta in Angstrom (exp(-(rij^2+rik^2)/eta^2),
src/nnp/basis_angular.cpp:376-382); Behler-style widths quoted in 1/A^2 map as
eta_radial = sqrt(eta) and eta_angular = 1/sqrt(eta).

Workloads (sizes as in the benchmark definitions):
  ar256      Ar fcc 4x4x4 (256 atoms), L = 22.98 A, jitter 0.15 A, COS rc 6,
             8 radial + 8 angular, 16-20-20-1 tanh.
  water64    64 H2O (192 atoms), L = 12.42 A, TANH rc 6, 80 inputs per element,
             80-32-32-1 softplus; batches of such structures.
  uneven     N in [32,1024] rounded to whole molecules, 0.9 g/cm^3, cubic cell
             strained by a random symmetric matrix with entries U(-0.1,0.1).
  lips       3 species (Li/P/S masses), jittered rock-salt (a = 5.6 A),
             96-256 atoms, 96 inputs, 96-48-48-1 tanh.
  water_md   16x16x16 replica of a 256-atom H/O cell (85 H2O + 1 H), L = 198.8 A.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional

import numpy as np

from .model import ANG, CUT, NEURON, RAD, Model, init_network, pair_index


@dataclasses.dataclass
class Batch:
    """A batch of periodic structures, concatenated atom-major."""
    natoms: np.ndarray      # (S,) int
    pos: np.ndarray         # (N, 3) float64
    cells: np.ndarray       # (S, 3, 3) float64, columns = lattice vectors
    species: np.ndarray     # (N,) int32

    @property
    def nstruc(self) -> int:
        return len(self.natoms)

    @property
    def atom_off(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.natoms)]).astype(np.int32)

    @property
    def ntotal(self) -> int:
        return int(self.natoms.sum())

    def struc(self, k: int):
        o = self.atom_off
        return self.pos[o[k]:o[k + 1]], self.cells[k], self.species[o[k]:o[k + 1]]

    def subset(self, idx) -> "Batch":
        idx = list(idx)
        o = self.atom_off
        sl = [np.arange(o[k], o[k + 1]) for k in idx]
        sel = np.concatenate(sl) if sl else np.zeros(0, int)
        return Batch(self.natoms[idx].copy(), self.pos[sel].copy(), self.cells[idx].copy(),
                     self.species[sel].copy())

    @staticmethod
    def concat(batches: List["Batch"]) -> "Batch":
        return Batch(np.concatenate([b.natoms for b in batches]),
                     np.concatenate([b.pos for b in batches]),
                     np.concatenate([b.cells for b in batches]),
                     np.concatenate([b.species for b in batches]).astype(np.int32))


def single(pos, cell, species) -> Batch:
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    return Batch(np.array([len(pos)]), pos, np.asarray(cell, np.float64).reshape(1, 3, 3),
                 np.asarray(species, np.int32))


# ---------------------------------------------------------------- potentials

def _uniform_basis(S, rad, ang):
    return [[rad for _ in range(S)] for _ in range(S)], \
           [[ang for _ in range(S * (S + 1) // 2)] for _ in range(S)]


def model_ar256(seed: int = 1) -> Model:
    rng = np.random.default_rng(seed)
    rs = 1.0 + np.arange(8) * 5.0 / 7.0
    rad = (RAD["GAUSSIAN"], rs, np.full(8, np.sqrt(0.5)))
    zeta = np.repeat([1.0, 2.0, 4.0, 8.0], 2)
    lam = np.tile([1, -1], 4)
    ang = (ANG["GAUSS"], np.full(8, 1.0 / np.sqrt(0.01)), zeta, lam)
    radial, angular = _uniform_basis(1, rad, ang)
    net = init_network(16, [20, 20], NEURON["TANH"], rng)
    return Model(["Ar"], np.array([39.948]), np.array([0.0]), 6.0, CUT["COS"], radial, angular, [net])


def model_ar_zeta5(seed: int = 5) -> Model:
    """Ar with the angular terms of the reference's Ar example
    (examples/Ar/basis_Ar: GAUSS, zeta 5, lambda -1, three widths) beside zeta 1
    (lambda 1), 2 (lambda -1) and 5 (lambda 1) on the same widths, and its COS
    cutoff; 10 radial GAUSSIAN."""
    rng = np.random.default_rng(seed)
    rs = np.linspace(1.0, 5.5, 10)
    rad = (RAD["GAUSSIAN"], rs, np.full(10, 1.0 / (rs[1] - rs[0])))
    eta = np.repeat([1.5, 2.5, 4.0], 4)
    zeta = np.tile([1.0, 2.0, 5.0, 5.0], 3)
    lam = np.tile([1, -1, -1, 1], 3)
    ang = (ANG["GAUSS"], eta, zeta, lam)
    radial, angular = _uniform_basis(1, rad, ang)
    net = init_network(22, [24, 24], NEURON["TANH"], rng)
    return Model(["Ar"], np.array([39.948]), np.array([0.0]), 6.0, CUT["COS"], radial, angular, [net])


def water_basis(nr: int = 16, zmax: float = 4.0):
    """16 radial GAUSSIAN per neighbour species (rs 0.8..6.0 A, eta = 1/spacing)
    and 16 angular GAUSS per pair: eta in {1.5,2.5,4,6} A x zeta in {1,zmax} x lambda +-1."""
    rs = np.linspace(0.8, 6.0, nr)
    rad = (RAD["GAUSSIAN"], rs, np.full(nr, 1.0 / (rs[1] - rs[0])))
    eta = np.repeat([1.5, 2.5, 4.0, 6.0], 4)
    zeta = np.tile([1.0, 1.0, zmax, zmax], 4)
    lam = np.tile([1, -1], 8)
    ang = (ANG["GAUSS"], eta, zeta, lam)
    return rad, ang


def model_water(seed: int = 2, neuron: str = "SOFTPLUS", hidden=(32, 32), zmax: float = 4.0) -> Model:
    rng = np.random.default_rng(seed)
    rad, ang = water_basis(zmax=zmax)
    radial, angular = _uniform_basis(2, rad, ang)
    nets = [init_network(80, list(hidden), NEURON[neuron], rng) for _ in range(2)]
    return Model(["H", "O"], np.array([1.008, 15.999]), np.array([0.0, 0.0]), 6.0, CUT["TANH"],
                 radial, angular, nets)


def model_lips(seed: int = 3) -> Model:
    rng = np.random.default_rng(seed)
    rs = np.linspace(1.0, 6.0, 8)
    rad = (RAD["GAUSSIAN"], rs, np.full(8, 1.0 / (rs[1] - rs[0])))
    eta = np.repeat([2.0, 3.5, 5.0], 4)
    zeta = np.tile([1.0, 1.0, 2.0, 2.0], 3)
    lam = np.tile([1, -1], 6)
    ang = (ANG["GAUSS"], eta, zeta, lam)
    radial, angular = _uniform_basis(3, rad, ang)
    nets = [init_network(96, [48, 48], NEURON["TANH"], rng) for _ in range(3)]
    return Model(["Li", "P", "S"], np.array([6.94, 30.97, 32.06]), np.zeros(3), 6.0, CUT["TANH"],
                 radial, angular, nets)


# ---------------------------------------------------------------- structures

def struc_ar256(seed: int = 11) -> Batch:
    rng = np.random.default_rng(seed)
    a = 22.98 / 4
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.array([[i, j, k] for i in range(4) for j in range(4) for k in range(4)], float)
    pos = ((g[:, None, :] + basis[None]) * a).reshape(-1, 3)
    pos = pos + rng.normal(0.0, 0.15, size=pos.shape)
    return single(pos, np.eye(3) * 22.98, np.zeros(256, np.int32))


def _rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _water_sites(rng, nmol: int, L: float, jitter: float, extra_h: int = 0):
    """Oxygens on a jittered grid (randomly chosen sites when the grid has more
    sites than molecules), H at 0.96 A with a 104.5 degree angle, random
    orientation.  Species: H = 0, O = 1.  Atom order per molecule: O, H, H."""
    ng = int(np.ceil(nmol ** (1.0 / 3.0) - 1e-9))
    sites = np.array([[i, j, k] for i in range(ng) for j in range(ng) for k in range(ng)], float)
    sites = sites[np.sort(rng.choice(len(sites), size=nmol + extra_h, replace=False))]
    sites = (sites + 0.5) * (L / ng) + rng.normal(0.0, jitter, size=sites.shape)
    th = np.deg2rad(104.5) / 2
    local = np.array([[0, 0, 0], [0.96 * np.sin(th), 0, 0.96 * np.cos(th)],
                      [-0.96 * np.sin(th), 0, 0.96 * np.cos(th)]])
    pos, spc = [], []
    for m in range(nmol):
        R = _rotation(rng)
        pos.append(sites[m] + local @ R.T)
        spc += [1, 0, 0]
    for e in range(extra_h):
        pos.append(sites[nmol + e][None])
        spc.append(0)
    return np.concatenate(pos), np.array(spc, np.int32)


def batch_water(nstruc: int = 64, seed: int = 12, nmol: int = 64, L: float = 12.42,
                jitter: float = 0.1) -> Batch:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nstruc):
        pos, spc = _water_sites(rng, nmol, L, jitter)
        out.append(single(pos, np.eye(3) * L, spc))
    return Batch.concat(out)


def batch_uneven(nstruc: int = 512, seed: int = 13, nmin: int = 32, nmax: int = 1024,
                 density: float = 0.9) -> Batch:
    rng = np.random.default_rng(seed)
    out = []
    NA = 6.02214076e23
    for _ in range(nstruc):
        n = int(rng.integers(nmin, nmax + 1))
        nmol = max(1, int(round(n / 3)))
        vol = nmol * 18.015 / NA / density * 1e24   # A^3
        L = vol ** (1.0 / 3.0)
        pos, spc = _water_sites(rng, nmol, L, 0.1)
        e = rng.uniform(-0.1, 0.1, size=(3, 3))
        strain = np.eye(3) + 0.5 * (e + e.T)
        cell = strain @ (np.eye(3) * L)
        out.append(single(pos @ strain.T, cell, spc))
    return Batch.concat(out)


def batch_lips(nstruc: int = 128, seed: int = 14, a: float = 5.6, jitter: float = 0.1) -> Batch:
    rng = np.random.default_rng(seed)
    fcc = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    shapes = [(x, y, z) for x in (2, 3, 4) for y in (2, 3, 4) for z in (2, 3, 4) if 12 <= x * y * z <= 32]
    out = []
    for _ in range(nstruc):
        nx, ny, nz = shapes[rng.integers(len(shapes))]
        g = np.array([[i, j, k] for i in range(nx) for j in range(ny) for k in range(nz)], float)
        cat = (g[:, None, :] + fcc[None]).reshape(-1, 3)
        ani = cat + 0.5 * np.array([1, 0, 0])
        pos = np.concatenate([cat, ani]) * a
        spc = np.concatenate([np.zeros(len(cat), np.int32),
                              np.where(rng.random(len(ani)) < 0.25, 1, 2).astype(np.int32)])
        pos = pos + rng.normal(0.0, jitter, size=pos.shape)
        out.append(single(pos, np.diag([nx * a, ny * a, nz * a]), spc))
    return Batch.concat(out)


def struc_water_md(seed: int = 15, reps: int = 16, L: float = 12.425, jitter: float = 0.05) -> Batch:
    """reps^3 replica of a 256-atom H/O cell (85 H2O + 1 extra H; H 171, O 85),
    each replica with independent Gaussian jitter."""
    rng = np.random.default_rng(seed)
    pos0, spc0 = _water_sites(rng, 85, L, 0.1, extra_h=1)
    g = np.array([[i, j, k] for i in range(reps) for j in range(reps) for k in range(reps)], float)
    pos = (g[:, None, :] * L + pos0[None]).reshape(-1, 3)
    pos = pos + rng.normal(0.0, jitter, size=pos.shape)
    spc = np.tile(spc0, len(g)).astype(np.int32)
    return single(pos, np.eye(3) * (L * reps), spc)


def small_random(seed: int, natoms: int = 20, nspecies: int = 3, a0: float = 20.0,
                 cell: Optional[np.ndarray] = None) -> Batch:
    """Random atoms in a cubic box with positions in [-a0, a0] (not wrapped),
    the generator shape of src/nnp/test_nnp_force.cpp:11-45."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-a0, a0, size=(natoms, 3))
    spc = rng.integers(0, nspecies, size=natoms).astype(np.int32)
    return single(pos, np.eye(3) * a0 if cell is None else cell, spc)


def model_family(rad: str = "GAUSSIAN", ang: str = "GAUSS", neuron: str = "TANH", cutoff: str = "TANH",
                 seed: int = 4, nr: int = 8, hidden=(16, 16)) -> Model:
    """Two-species (H/O) potential on the water structures with one radial and
    one angular family throughout: nr radial functions per neighbour species
    (rs 0.8..5.5 A, eta = 1/spacing) and 8 angular per pair, widths {1.5, 3.0} A
    x zeta {1, 2} x lambda +-1 (a grid: GAUSS runs the moment form)."""
    rng = np.random.default_rng(seed)
    rs = np.linspace(0.8, 5.5, nr)
    radb = (RAD[rad], rs, np.full(nr, 1.0 / (rs[1] - rs[0])))
    eta = np.repeat([1.5, 3.0], 4)
    zeta = np.tile([1.0, 1.0, 2.0, 2.0], 2)
    lam = np.tile([1, -1], 4)
    angb = (ANG[ang], eta, zeta, lam)
    radial, angular = _uniform_basis(2, radb, angb)
    nin = 2 * nr + 3 * 8
    nets = [init_network(nin, list(hidden), NEURON[neuron], rng) for _ in range(2)]
    return Model(["H", "O"], np.array([1.008, 15.999]), np.array([-0.5, -2.0]), 6.0, CUT[cutoff],
                 radial, angular, nets)
