"""Input preconditioning of the per-species networks (host set-up of training).

Restates the choice of input bias and scale in NNPTEFR::train,
src/nnp/nnptefr2.cpp:776-867, from the statistics that ``Engine.input_stats``
(C ABI ``ann_input_stats``, :690-775) gathers on the GPU:

* bias (:783-820): MEAN -mean, HAR -harmonic mean, MID -(max + min) / 2,
  MIN -min, RMS -rms, IDENTITY 0; then + ``inbias`` (:821-823);
* scale (:824-856): DEV 1 / deviation (1 where the deviation is 0),
  MINMAX 1 / (max - min) (1 unless positive), MAX 1 / max (1 unless positive),
  IDENTITY 1; then x ``inscale`` (:857-859);
* the network input becomes inpw * (G + inpb) (src/ml/nn.cpp:1346) and the
  output scaling is reset to outw = 1, outb = 0 (:862-867).
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

BIAS = ("MEAN", "HAR", "MID", "MIN", "RMS", "IDENTITY")
SCALE = ("DEV", "MINMAX", "MAX", "IDENTITY")

# rows of the statistics array of Engine.input_stats / ann_input_stats
ST_MIN, ST_MAX, ST_HAR, ST_MEAN, ST_DEV, ST_RMS = range(6)


def precondition(nin: Sequence[int], stats: np.ndarray, bias: str = "MEAN", scale: str = "DEV",
                 inbias: float = 0.0, inscale: float = 1.0) -> List[Tuple[np.ndarray, np.ndarray]]:
    """Return [(inpw, inpb)] per species from stats (S, 6, max_ninput)."""
    bias, scale = bias.upper(), scale.upper()
    if bias not in BIAS:
        raise ValueError(f"Invalid input bias {bias!r}.")
    if scale not in SCALE:
        raise ValueError(f"Invalid input scaling {scale!r}.")
    out = []
    for a, n in enumerate(nin):
        st = np.asarray(stats[a], dtype=np.float64)[:, :n]
        if bias == "MEAN":
            b = -st[ST_MEAN]
        elif bias == "HAR":
            b = -st[ST_HAR]
        elif bias == "MID":
            b = -0.5 * (st[ST_MAX] + st[ST_MIN])
        elif bias == "MIN":
            b = -st[ST_MIN]
        elif bias == "RMS":
            b = -st[ST_RMS]
        else:
            b = np.zeros(n)
        b = b + inbias
        if scale == "DEV":
            d = st[ST_DEV]
            w = np.where(d == 0, 1.0, 1.0 / np.where(d == 0, 1.0, d))
        elif scale == "MINMAX":
            s = st[ST_MAX] - st[ST_MIN]
            w = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 1.0)
        elif scale == "MAX":
            m = st[ST_MAX]
            w = np.where(m > 0, 1.0 / np.where(m > 0, m, 1.0), 1.0)
        else:
            w = np.ones(n)
        out.append((w * inscale, b))
    return out


def apply_preconditioning(engine, bias: str = "MEAN", scale: str = "DEV", inbias: float = 0.0,
                          inscale: float = 1.0, G=None) -> List[Tuple[np.ndarray, np.ndarray]]:
    """Statistics of the engine's current batch -> input scaling of every species network."""
    _, stats = engine.input_stats(G)
    nin = [engine.model.ninput(a) for a in range(engine.nspecies)]
    pre = precondition(nin, stats, bias, scale, inbias, inscale)
    for a, (w, b) in enumerate(pre):
        engine.set_scaling(a, w, b, 1.0, 0.0)
    return pre
