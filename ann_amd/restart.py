"""Training checkpoints: what the trainer's restart file holds for the path
built here (NNPTEFR restart, src/nnp/nnptefr2.cpp:73-195 written at :1012-1023,
read at :533-566): the FP64 parameters, the optimiser state (Adam moments and
steps taken), the input / output scaling of every species network and the
iteration counter.

The file is a numpy ``.npz`` archive of plain arrays (read with
``allow_pickle=False``), not the reference's serialised object: the potential
itself travels in the reference's text format (``ann_amd/potfile.py``).
"""
from __future__ import annotations

from typing import Dict

import numpy as np

_VERSION = 1


def save_restart(path: str, engine, iteration: int = 0) -> None:
    """Write engine's parameters, Adam state and scaling to ``path`` (.npz)."""
    m, v, step = engine.get_optimizer_state()
    arrays: Dict[str, np.ndarray] = {
        "version": np.array([_VERSION]),
        "params": engine.get_params(),
        "adam_m": m,
        "adam_v": v,
        "adam_step": np.array([step]),
        "iteration": np.array([int(iteration)]),
        "nspecies": np.array([engine.nspecies]),
    }
    for a in range(engine.nspecies):
        w, b, ow, ob = engine.scaling(a)
        arrays[f"inpw_{a}"] = np.asarray(w, dtype=np.float64)
        arrays[f"inpb_{a}"] = np.asarray(b, dtype=np.float64)
        arrays[f"out_{a}"] = np.array([ow, ob], dtype=np.float64)
    with open(path, "wb") as fh:
        np.savez(fh, **arrays)


def load_restart(path: str, engine) -> int:
    """Restore a checkpoint into ``engine``; returns the stored iteration.

    Parameters first (this restarts Adam from zero moments), then the Adam state,
    then the scaling of every species.
    """
    with np.load(path, allow_pickle=False) as z:
        if int(z["version"][0]) != _VERSION:
            raise ValueError(f"unsupported restart version {int(z['version'][0])}")
        if int(z["nspecies"][0]) != engine.nspecies:
            raise ValueError("restart file was written for a different number of species")
        params = z["params"]
        if len(params) != engine.nparams:
            raise ValueError("restart file was written for a different network")
        engine.set_params(params)
        engine.set_optimizer_state(z["adam_m"], z["adam_v"], int(z["adam_step"][0]))
        for a in range(engine.nspecies):
            ow, ob = z[f"out_{a}"]
            engine.set_scaling(a, z[f"inpw_{a}"], z[f"inpb_{a}"], float(ow), float(ob))
        return int(z["iteration"][0])
