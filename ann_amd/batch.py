"""Training-batch sampler of the reference trainer, and a device-resident
training set from which batches are gathered on the GPU.

``BatchSampler`` restates ``Batch`` (src/ml/batch.hpp:18-55,
src/ml/batch.cpp) as NNPTEFR uses it: at set-up the index list of the
training set is shuffled once (src/nnp/nnptefr2.cpp:676-678); every step takes
the next ``size`` indices cyclically, ``elements[i] = data[count++ % capacity]``,
sorts them, and once ``count`` reaches ``capacity`` reshuffles the list and
restarts the count (:1309-1315).  A step therefore never repeats an index,
and when ``size`` divides ``capacity`` every epoch visits each structure once.
The permutation comes from a seeded numpy Generator, not from std::shuffle
over std::mt19937, so the order of the draws differs from the reference's
while every relation above holds.

``ResidentSet`` keeps a whole training set (positions, cells, species, labels)
in HBM and gathers the structures of a step into contiguous batch arrays with
device index operations, so that a rotating-batch training step moves only
the index lists across PCIe.

``validation_error`` is the trainer's error report on a set of structures
(the validation loss and RMSE of NNPTEFR::error_cost_o1,
src/nnp/nnptefr2.cpp:1375-1411, and their averages at :986-991).
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np

from .configs import Batch


class BatchSampler:
    def __init__(self, size: int, capacity: int, seed: int = 0):
        if size < 0:
            raise ValueError("Batch::resize(int,int): size cannot be negative.")
        if capacity < 0:
            raise ValueError("Batch::resize(int,int): capacity cannot be negative.")
        if size > capacity:
            raise ValueError("Batch::resize(int,int): size must be smaller than capacity")
        self.size = size
        self.capacity = capacity
        self.count = 0
        self.rng = np.random.default_rng(seed)
        self.data = np.arange(capacity)
        self.rng.shuffle(self.data)
        self.elements = np.arange(size)

    def next(self) -> np.ndarray:
        """Indices of the next batch (sorted)."""
        # elements[i] = data[count++ % capacity] for i < size, in one step
        self.elements[:] = self.data[(self.count + np.arange(self.size)) % self.capacity]
        self.count += self.size
        self.elements.sort()
        if self.count >= self.capacity:
            self.rng.shuffle(self.data)
            self.count = 0
        return self.elements.copy()


class ResidentSet:
    """A training set in device memory and the gather of one batch from it."""

    def __init__(self, batch: Batch, energy: np.ndarray, forces: np.ndarray, device: str = "cuda:0",
                 wE: Optional[np.ndarray] = None, wF: Optional[np.ndarray] = None):
        import torch
        self.batch = batch
        self.device = torch.device(device)
        self.off = batch.atom_off.astype(np.int64)
        self.natoms = batch.natoms.astype(np.int64)
        dev = self.device
        self.pos = torch.from_numpy(np.ascontiguousarray(batch.pos, np.float64)).to(dev)
        self.cells = torch.from_numpy(np.ascontiguousarray(batch.cells, np.float64).reshape(-1, 9)).to(dev)
        self.species = torch.from_numpy(np.ascontiguousarray(batch.species, np.int32)).to(dev)
        self.energy = torch.as_tensor(np.asarray(energy, np.float64), device=dev)
        self.forces = torch.as_tensor(np.asarray(forces, np.float64).reshape(-1, 3), device=dev)
        self.wE = None if wE is None else torch.as_tensor(np.asarray(wE, np.float64), device=dev)
        self.wF = None if wF is None else torch.as_tensor(np.asarray(wF, np.float64), device=dev)

    def gather(self, idx: np.ndarray, device_meta: bool = True):
        """Contiguous batch arrays of structures idx (host atom offsets; device
        positions, energies, forces and weights; device cells and species unless
        device_meta is false, as Engine.set_batch_host takes the host copies)."""
        import torch
        idx = np.asarray(idx, dtype=np.int64)
        n = self.natoms[idx]
        atom_off = np.concatenate([[0], np.cumsum(n)]).astype(np.int32)
        # atom a of batch structure p is atom off[idx[p]] + (a - atom_off[p]) of the set
        atoms = np.repeat(self.off[idx] - atom_off[:-1], n) + np.arange(int(atom_off[-1]), dtype=np.int64)
        ai = torch.from_numpy(atoms).to(self.device, non_blocking=True)
        si = torch.from_numpy(idx).to(self.device, non_blocking=True)
        out = {"atom_off": atom_off,
               "pos": self.pos.index_select(0, ai),
               "cells": self.cells.index_select(0, si) if device_meta else None,
               "species": self.species.index_select(0, ai) if device_meta else None,
               "energy": self.energy.index_select(0, si), "forces": self.forces.index_select(0, ai),
               # host copies for Engine.set_batch_host (no device readback in the set-up)
               "cells_host": self.batch.cells[idx], "species_host": self.batch.species[atoms]}
        out["wE"] = None if self.wE is None else self.wE.index_select(0, si)
        out["wF"] = None if self.wF is None else self.wF.index_select(0, si)
        return out


def _energy_norm(norm: str, n: np.ndarray) -> np.ndarray:
    """Per-structure energy normalisation (src/nnp/nnptefr2.cpp:600-624)."""
    norm = norm.upper()
    if norm == "IDENTITY":
        return np.ones_like(n)
    if norm == "LINEAR":
        return 1.0 / n
    if norm == "SQRT":
        return 1.0 / np.sqrt(n)
    if norm == "CBRT":
        return 1.0 / np.cbrt(n)
    if norm == "LOG":
        return 1.0 / (np.log(n) + 1.0)
    raise ValueError("Invalid normalization method.")


def _energy_loss(loss: str, delta: float, dU: np.ndarray) -> np.ndarray:
    """Loss of a normalised energy error (src/nnp/nnptefr2.cpp:1392-1410)."""
    loss = loss.upper()
    if loss == "MSE":
        return 0.5 * dU * dU
    if loss == "MAE":
        return np.abs(dU)
    a = dU / delta
    q = np.sqrt(1.0 + a * a)
    if loss == "HUBER":
        return delta * delta * (q - 1.0)
    if loss == "ASINH":
        return delta * delta * (1.0 - q + a * np.arcsinh(a))
    raise ValueError("Invalid loss function.")


def validation_error(eng, rs: ResidentSet, idx, loss: str = "MSE", delta: float = 1.0, norm: str = "LINEAR",
                     form: int = 1, chunk: int = 512) -> dict:   # form 1: FAST
    """Validation loss and RMSE of structures ``idx`` of a resident set with the
    engine's current parameters, as the reference trainer reports them
    (src/nnp/nnptefr2.cpp:1375-1411, averaged at :986-991): for each structure
    the network energy pe (no vacuum energies, like the labels), dE = pe - E_ref,
    dU = dE * norm(n); loss = sum loss(dU) / n_val and
    rmse = sqrt(sum (dE / n)^2 / n_val).  The energies are computed on the GPU
    (energy only: no adjoint) in chunks of ``chunk`` structures; the engine is
    left holding the last chunk as its batch."""
    idx = np.asarray(idx, dtype=np.int64)
    if idx.size == 0:
        raise ValueError("NNPTEFR::train(int): No validation data provided.")
    lsum, rsum = 0.0, 0.0
    for c0 in range(0, idx.size, chunk):
        b = rs.gather(idx[c0:c0 + chunk], device_meta=False)
        eng.set_batch_host(b["atom_off"], b["pos"], b["cells_host"], b["species_host"])
        eng.neighbors()
        E = eng.energies(form, with_evac=False).cpu().numpy()
        n = np.diff(b["atom_off"]).astype(np.float64)
        dE = E - b["energy"].cpu().numpy()
        lsum += float(_energy_loss(loss, delta, dE * _energy_norm(norm, n)).sum())
        rsum += float(((dE / n) ** 2).sum())
    return {"loss": lsum / idx.size, "rmse": math.sqrt(rsum / idx.size), "structures": int(idx.size)}
