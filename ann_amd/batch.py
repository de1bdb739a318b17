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
"""
from __future__ import annotations

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

    def gather(self, idx: np.ndarray):
        """Contiguous batch arrays of structures idx (host atom offsets; device
        positions, cells, species, energies, forces and weights)."""
        import torch
        idx = np.asarray(idx, dtype=np.int64)
        n = self.natoms[idx]
        atom_off = np.concatenate([[0], np.cumsum(n)]).astype(np.int32)
        # atom a of batch structure p is atom off[idx[p]] + (a - atom_off[p]) of the set
        atoms = np.repeat(self.off[idx] - atom_off[:-1], n) + np.arange(int(atom_off[-1]), dtype=np.int64)
        ai = torch.from_numpy(atoms).to(self.device, non_blocking=True)
        si = torch.from_numpy(idx).to(self.device, non_blocking=True)
        out = {"atom_off": atom_off,
               "pos": self.pos.index_select(0, ai), "cells": self.cells.index_select(0, si),
               "species": self.species.index_select(0, ai),
               "energy": self.energy.index_select(0, si), "forces": self.forces.index_select(0, ai)}
        out["wE"] = None if self.wE is None else self.wE.index_select(0, si)
        out["wF"] = None if self.wF is None else self.wF.index_select(0, si)
        return out
