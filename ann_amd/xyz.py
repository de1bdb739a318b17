"""Extended-XYZ structure files (host I/O), restating XYZ::read of the reference,
src/format/xyz_struc.cpp:38-186:

* line 1: the atom count (> 0);
* line 2, matched case-insensitively: ``Properties=species:S:1:pos:R:3[:forces:R:3]``
  (forces are read only when declared, :84-117, :131-146), ``potential_energy=E``
  (:118-121) and ``Lattice="a1 a2 a3"`` with the nine numbers filling the lattice
  vectors as columns, a1 first (:122-134);
* one line per atom: name, x, y, z[, fx, fy, fz];
* with a lattice, positions are returned to the cell (Cell::returnToCell, :168-173).

Units are the reference's METAL system (Angstrom, eV).  A file may hold several
frames one after another; each becomes one structure of the returned batch.
Species are mapped by name through ``names`` (the potential's species order).
"""
from __future__ import annotations

import re
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .configs import Batch


def _return_to_cell(pos: np.ndarray, cell: np.ndarray) -> np.ndarray:
    # fractional coordinates in [0, 1), back to Cartesian (returnToCell)
    f = np.linalg.solve(cell, pos.T).T
    f -= np.floor(f)
    return f @ cell.T


def _parse_header(line: str):
    up = line.upper()
    forces = False
    m = re.search(r"PROPERTIES\s*=\s*(\S+)", up)
    if m:
        tok = m.group(1).split(":")
        for q in range(0, len(tok) - 2, 3):
            tag, kind, n = tok[q], tok[q + 1], tok[q + 2]
            if tag == "SPECIES" and (kind != "S" or int(n) != 1):
                raise ValueError("invalid species property")
            if tag == "POS" and (kind != "R" or int(n) != 3):
                raise ValueError("invalid position property")
            if tag == "FORCES":
                if kind != "R" or int(n) != 3:
                    raise ValueError("invalid force property")
                forces = True
    pe = 0.0
    m = re.search(r"POTENTIAL_ENERGY\s*=\s*([-+0-9.EeDd]+)", up)
    if m:
        pe = float(m.group(1).replace("D", "E"))
    cell = None
    m = re.search(r'LATTICE\s*=\s*"([^"]*)"', up)
    if m:
        v = [float(x) for x in m.group(1).split()]
        if len(v) != 9:
            raise ValueError("a lattice needs nine numbers")
        cell = np.array(v, dtype=np.float64).reshape(3, 3).T   # columns = lattice vectors
    return forces, pe, cell


def read_xyz(path: str, names: Sequence[str]) -> Tuple[Batch, np.ndarray, Optional[np.ndarray]]:
    """Read every frame of an extended-XYZ file.

    Returns (batch, energies (S,), forces (N, 3) or None when no frame declares them).
    A frame without a lattice gets a zero cell (a free cluster) and keeps its positions.
    """
    index = {n.upper(): k for k, n in enumerate(names)}
    with open(path) as fh:
        lines = fh.read().splitlines()
    natoms: List[int] = []
    pos, cells, species, pes, frc = [], [], [], [], []
    any_forces = False
    q = 0
    while q < len(lines):
        if not lines[q].strip():
            q += 1
            continue
        n = int(lines[q].split()[0])
        if n <= 0:
            raise ValueError("found zero atoms")
        forces, pe, cell = _parse_header(lines[q + 1] if q + 1 < len(lines) else "")
        p = np.zeros((n, 3))
        f = np.zeros((n, 3))
        s = np.zeros(n, dtype=np.int32)
        for a in range(n):
            tok = lines[q + 2 + a].split()
            if tok[0].upper() not in index:
                raise ValueError(f"unknown species {tok[0]!r}")
            s[a] = index[tok[0].upper()]
            p[a] = [float(x) for x in tok[1:4]]
            if forces:
                f[a] = [float(x) for x in tok[4:7]]
        if cell is not None and np.linalg.norm(cell) > 0:
            p = _return_to_cell(p, cell)
        else:
            cell = np.zeros((3, 3))
        natoms.append(n); pos.append(p); cells.append(cell); species.append(s); pes.append(pe); frc.append(f)
        any_forces |= forces
        q += 2 + n
    batch = Batch(np.asarray(natoms, dtype=np.int64), np.concatenate(pos), np.stack(cells),
                  np.concatenate(species).astype(np.int32))
    return batch, np.asarray(pes), (np.concatenate(frc) if any_forces else None)


def write_xyz(path: str, batch: Batch, names: Sequence[str], energies=None, forces=None) -> None:
    """Write every structure of ``batch`` as one extended-XYZ frame."""
    off = batch.atom_off
    with open(path, "w") as fh:
        for k in range(batch.nstruc):
            p, c, s = batch.struc(k)
            props = "species:S:1:pos:R:3" + (":forces:R:3" if forces is not None else "")
            lat = " ".join(repr(float(x)) for x in c.T.reshape(-1))
            head = f'Lattice="{lat}" Properties={props}'
            if energies is not None:
                head += f" potential_energy={float(energies[k])!r}"
            fh.write(f"{len(p)}\n{head}\n")
            for a in range(len(p)):
                row = [names[s[a]]] + [repr(float(x)) for x in p[a]]
                if forces is not None:
                    row += [repr(float(x)) for x in forces[off[k] + a]]
                fh.write(" ".join(row) + "\n")
