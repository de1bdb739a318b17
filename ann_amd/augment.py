"""Taylor augmentation of a force-labelled training set (host data preparation).

Restates the structure augmentation of the reference trainer,
src/nnp/nnptefr2.cpp:2383-2482: every structure whose mean per-atom force norm
exceeds 1e-12 gets ``nadd`` displaced copies.  Each atom j of a copy moves by
dr_j = v_j (sin(theta) cos(phi), sin(theta) sin(phi), cos(theta)) with
cos(theta) ~ U(-1, 1), phi ~ pi U(-1, 1) and the length v_j drawn from the
chosen distribution of scale ``rdelta`` (:2386-2406, :2447-2462; the COVALENT
variant multiplies v_j by the species' covalent radius, here ``scale[species]``),
and the copy's energy is the first-order Taylor estimate
E' = E - sum_j dr_j . F_j (:2466-2467).  The copies carry no forces (:2470),
so they train the energy term only.  Originals and copies keep the reference's
order: each original followed by its copies (:2435-2473).

Two random streams:

* ``rng="reference"`` draws exactly what the reference trainer draws: the
  std::mt19937 seeded through the trainer's chain (:2160-2168, one rank) and
  the reference's distribution objects (ann_amd.refrng), atom by atom in the
  order length, cos(theta), phi, with the position and energy updates done in
  the reference's order and rounding (E' -= dr . F one atom at a time).  For
  the same seed, structures and labels the displaced structures and energies
  equal the reference's bit for bit.  A :class:`TrainerDraws` passed as
  ``draws`` continues one stream over several sets, as the reference does for
  its training, validation and test sets.
* ``rng="numpy"`` (default) draws from a seeded numpy Generator, vectorised;
  every relation above holds, the displacement sequence differs.
"""
from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np

from .configs import Batch
from .refrng import Distribution, sincos, trainer_generator, uniform_real

_DISTS = ("uniform", "normal", "laplace", "logistic")
_PI = 3.141592653589793


def _lengths(rng: np.random.Generator, dist: str, rdelta: float, n: int) -> np.ndarray:
    # rng::dist::Uniform(0, rdelta), Normal / Laplace / Logistic(0, rdelta) (:2386-2399)
    if dist == "uniform":
        return rng.uniform(0.0, rdelta, size=n)
    if dist == "normal":
        return rng.normal(0.0, rdelta, size=n)
    if dist == "laplace":
        return rng.laplace(0.0, rdelta, size=n)
    if dist == "logistic":
        return rng.logistic(0.0, rdelta, size=n)
    raise ValueError(f"unsupported displacement distribution {dist!r} (supported: {', '.join(_DISTS)})")


class TrainerDraws:
    """The reference trainer's displacement stream: the rank's std::mt19937
    (seed chain of src/nnp/nnptefr2.cpp:2160-2168 from ``seed_global`` > 0) and
    one distribution object of scale ``rdelta`` (:2385-2406), created once and
    shared by every set, so a Normal keeps its cached Box-Muller partner."""

    def __init__(self, seed_global: int, dist: str, rdelta: float):
        if not 0 < int(seed_global) < 2**32:
            raise ValueError("the reference stream needs a seed in [1, 2^32): "
                             "the trainer seeds from the clock otherwise (:2161)")
        self.gen = trainer_generator(int(seed_global))
        self.dist = Distribution(dist, rdelta)

    def atom(self) -> Tuple[float, float, float]:
        """(length, cos(theta), phi / pi) of one atom, in the reference's draw order (:2450-2458)."""
        vr = self.dist.rand(self.gen)
        ct = uniform_real(self.gen, -1.0, 1.0)
        ph = uniform_real(self.gen, -1.0, 1.0)
        return vr, ct, ph


def _displace_reference(draws: TrainerDraws, p: np.ndarray, f: np.ndarray, s: np.ndarray, e: float,
                        scale: Optional[np.ndarray]) -> Tuple[np.ndarray, float]:
    # one copy, atom by atom as src/nnp/nnptefr2.cpp:2447-2467 updates it
    q = np.array(p, dtype=np.float64, copy=True)
    for j in range(len(p)):
        vr, vct, u = draws.atom()
        if scale is not None:
            vr = vr * float(scale[s[j]])
        vst = math.sqrt(1.0 - vct * vct)
        sn, cs = sincos(u * _PI)
        d0 = vr * vst * cs
        d1 = vr * vst * sn
        d2 = vr * vct
        q[j, 0] += d0
        q[j, 1] += d1
        q[j, 2] += d2
        fj = f[j]
        e -= d0 * float(fj[0]) + d1 * float(fj[1]) + d2 * float(fj[2])
    return q, e


def taylor_augment(batch: Batch, energy: np.ndarray, forces: np.ndarray, nadd: int, rdelta: float,
                   dist: str = "uniform", seed: int = 0, scale: Optional[np.ndarray] = None,
                   rng: str = "numpy", draws: Optional[TrainerDraws] = None
                   ) -> Tuple[Batch, np.ndarray, np.ndarray, np.ndarray]:
    """Return (batch', energy', forces', has_force) with the displaced copies inserted.

    ``energy`` (S,) and ``forces`` (N, 3) are the reference labels of ``batch``;
    ``forces'`` holds the originals' forces and zeros for the copies, whose
    ``has_force`` entry is False.  ``scale`` (per species) selects the
    covalent-radius variant of the displacement length.  ``rng`` picks the
    stream: "numpy" (``dist`` one of uniform, normal, laplace, logistic) or
    "reference" (``dist`` also sech or cosine; ``seed`` is the trainer's
    global seed, or ``draws`` continues an existing stream).
    """
    if nadd < 0:
        raise ValueError("nadd must be >= 0")
    if rng not in ("numpy", "reference"):
        raise ValueError(f"unknown random stream {rng!r} (numpy or reference)")
    energy = np.asarray(energy, dtype=np.float64).reshape(-1)
    forces = np.asarray(forces, dtype=np.float64).reshape(-1, 3)
    if len(energy) != batch.nstruc or len(forces) != batch.ntotal:
        raise ValueError("labels do not match the batch")
    if scale is not None:
        scale = np.asarray(scale, dtype=np.float64)
    if rng == "reference":
        if draws is None:
            draws = TrainerDraws(seed, dist, rdelta)
        gen = None
    else:
        if draws is not None:
            raise ValueError("draws continues the reference stream only")
        if dist not in _DISTS:
            raise ValueError(f"unsupported displacement distribution {dist!r} (supported: {', '.join(_DISTS)})")
        gen = np.random.default_rng(seed)
    off = batch.atom_off
    natoms, pos, cells, species, E, F, hasf = [], [], [], [], [], [], []
    for k in range(batch.nstruc):
        p, c, s = batch.struc(k)
        f = forces[off[k]:off[k + 1]]
        natoms.append(len(p)); pos.append(p); cells.append(c); species.append(s)
        E.append(energy[k]); F.append(f); hasf.append(True)
        favg = np.linalg.norm(f, axis=1).mean() if len(p) else 0.0
        if not favg > 1.0e-12:
            continue
        for _ in range(nadd):
            if gen is None:
                q, e = _displace_reference(draws, p, f, s, float(energy[k]), scale)
            else:
                v = _lengths(gen, dist, rdelta, len(p))
                if scale is not None:
                    v = v * scale[s]
                ct = gen.uniform(-1.0, 1.0, size=len(p))
                st = np.sqrt(1.0 - ct * ct)
                ph = gen.uniform(-1.0, 1.0, size=len(p)) * np.pi
                dr = v[:, None] * np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)
                q, e = p + dr, energy[k] - float(np.sum(dr * f))
            natoms.append(len(p)); pos.append(q); cells.append(c); species.append(s)
            E.append(e); F.append(np.zeros_like(f)); hasf.append(False)
    out = Batch(np.asarray(natoms, dtype=np.int64), np.concatenate(pos) if pos else np.zeros((0, 3)),
                np.stack(cells) if cells else np.zeros((0, 3, 3)),
                np.concatenate(species).astype(np.int32) if species else np.zeros(0, np.int32))
    return out, np.asarray(E), (np.concatenate(F) if F else np.zeros((0, 3))), np.asarray(hasf)
