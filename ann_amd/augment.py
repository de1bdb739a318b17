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

The reference draws from std::mt19937 through its own distribution classes;
this restatement uses a seeded numpy Generator, so the sequence of
displacements differs while every relation above holds exactly.
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

from .configs import Batch

_DISTS = ("uniform", "normal", "laplace", "logistic")


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


def taylor_augment(batch: Batch, energy: np.ndarray, forces: np.ndarray, nadd: int, rdelta: float,
                   dist: str = "uniform", seed: int = 0, scale: Optional[np.ndarray] = None
                   ) -> Tuple[Batch, np.ndarray, np.ndarray, np.ndarray]:
    """Return (batch', energy', forces', has_force) with the displaced copies inserted.

    ``energy`` (S,) and ``forces`` (N, 3) are the reference labels of ``batch``;
    ``forces'`` holds the originals' forces and zeros for the copies, whose
    ``has_force`` entry is False.  ``scale`` (per species) selects the
    covalent-radius variant of the displacement length.
    """
    if nadd < 0:
        raise ValueError("nadd must be >= 0")
    energy = np.asarray(energy, dtype=np.float64).reshape(-1)
    forces = np.asarray(forces, dtype=np.float64).reshape(-1, 3)
    if len(energy) != batch.nstruc or len(forces) != batch.ntotal:
        raise ValueError("labels do not match the batch")
    rng = np.random.default_rng(seed)
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
            v = _lengths(rng, dist, rdelta, len(p))
            if scale is not None:
                v = v * np.asarray(scale, dtype=np.float64)[s]
            ct = rng.uniform(-1.0, 1.0, size=len(p))
            st = np.sqrt(1.0 - ct * ct)
            ph = rng.uniform(-1.0, 1.0, size=len(p)) * np.pi
            dr = v[:, None] * np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)
            natoms.append(len(p)); pos.append(p + dr); cells.append(c); species.append(s)
            E.append(energy[k] - float(np.sum(dr * f))); F.append(np.zeros_like(f)); hasf.append(False)
    out = Batch(np.asarray(natoms, dtype=np.int64), np.concatenate(pos) if pos else np.zeros((0, 3)),
                np.stack(cells) if cells else np.zeros((0, 3, 3)),
                np.concatenate(species).astype(np.int32) if species else np.zeros(0, np.int32))
    return out, np.asarray(E), (np.concatenate(F) if F else np.zeros((0, 3))), np.asarray(hasf)
