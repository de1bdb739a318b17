"""Neural-network potential description (host side).

Holds what the reference keeps in ``NNP`` / ``NNH`` / ``BasisR`` / ``BasisA`` /
``NN::ANN`` (src/nnp/nnp.hpp:67-123, src/nnp/nnh.hpp:49-105,
src/nnp/basis_radial.hpp:23-89, src/nnp/basis_angular.hpp:23-103,
src/ml/nn.hpp:242-371) as numpy arrays in the reference's own units and
conventions:

* radial basis of centre species ``a`` for neighbour species ``b``:
  ``radial[a][b] = (name, rs, eta)``;
* angular basis of centre ``a`` for neighbour pair ``(j <= k)``, stored in
  TensorTri order ``p = k(k+1)/2 + j`` (src/mem/tensortri.hpp:196-201):
  ``angular[a][p] = (name, eta, zeta, lambda)``;
* one network per species: layer sizes (hidden + output), neuron, sharpness,
  packed parameters in the reference order (all biases layer by layer, then
  every weight matrix column-major; src/ml/nn.cpp:846-865), input scaling
  ``inpw``/``inpb`` and output scaling ``outw``/``outb``
  (src/ml/nn.cpp:1343-1365), vacuum energy per species (src/nnp/nnh.cpp:118-121).

``flat()`` lays the tables out as the flat arrays both C interfaces take.
"""
from __future__ import annotations

import dataclasses
from typing import List, Tuple

import numpy as np

# enumerations use the reference numbering
CUT = {"NONE": 0, "STEP": 1, "COS": 2, "TANH": 3, "POLY3": 4}
RAD = {"NONE": 0, "GAUSSIAN": 1, "SECH": 2, "LOGISTIC": 3, "TANH": 4, "LOGCOSH": 5, "LOGCOSH2": 6}
ANG = {"NONE": 0, "GAUSS": 1, "GAUSS2": 2, "SECH": 3, "STUDENT3": 4, "STUDENT4": 5, "STUDENT5": 6}
NEURON = {"LINEAR": 1, "SIGMOID": 2, "TANH": 3, "ISRU": 4, "ARCTAN": 5, "RELU": 6, "ELU": 7,
          "TANHRE": 8, "SQRE": 9, "SWISH": 10, "GELU": 11, "MISH": 12, "PFLU": 13, "SERF": 14,
          "LOGISH": 15, "SOFTPLUS": 16, "SQPLUS": 17, "ATISH": 18}


def pair_index(j: int, k: int) -> int:
    """TensorTri rank-2 storage index (src/mem/tensortri.hpp:196-201)."""
    if j > k:
        j, k = k, j
    return k * (k + 1) // 2 + j


@dataclasses.dataclass
class Network:
    nin: int
    nodes: List[int]            # hidden layers + output layer (output size 1)
    neuron: int
    c: float
    params: np.ndarray          # reference packing order
    inpw: np.ndarray
    inpb: np.ndarray
    outw: float = 1.0
    outb: float = 0.0

    @property
    def nparams(self) -> int:
        n, prev = 0, self.nin
        for h in self.nodes:
            n += h + h * prev
            prev = h
        return n

    def layers(self):
        """(W, b) per layer, W of shape (n_out, n_in) (column-major source)."""
        out, cnt, prev = [], 0, self.nin
        bs = []
        for h in self.nodes:
            bs.append(self.params[cnt:cnt + h])
            cnt += h
        for l, h in enumerate(self.nodes):
            w = self.params[cnt:cnt + h * prev].reshape(prev, h).T
            cnt += h * prev
            out.append((w, bs[l]))
            prev = h
        return out


@dataclasses.dataclass
class Model:
    species: List[str]
    masses: np.ndarray
    evac: np.ndarray
    rc: float
    cutoff: int
    radial: List[List[Tuple[int, np.ndarray, np.ndarray]]]
    angular: List[List[Tuple[int, np.ndarray, np.ndarray, np.ndarray]]]
    nets: List[Network]

    @property
    def nspecies(self) -> int:
        return len(self.species)

    def ninput(self, a: int) -> int:
        return sum(len(r[1]) for r in self.radial[a]) + sum(len(g[1]) for g in self.angular[a])

    @property
    def max_ninput(self) -> int:
        return max(self.ninput(a) for a in range(self.nspecies))

    def offsets(self, a: int):
        """Radial offsets per neighbour species and angular offsets per pair
        (angular offsets include nInputR), NNH::init_input src/nnp/nnh.cpp:94-112."""
        offr, o = [], 0
        for r in self.radial[a]:
            offr.append(o)
            o += len(r[1])
        offa = []
        for g in self.angular[a]:
            offa.append(o)
            o += len(g[1])
        return offr, offa

    def param_vector(self) -> np.ndarray:
        return np.concatenate([n.params for n in self.nets])

    def set_param_vector(self, p: np.ndarray) -> None:
        cnt = 0
        for n in self.nets:
            k = n.nparams
            n.params = np.array(p[cnt:cnt + k], dtype=np.float64)
            cnt += k

    def copy(self) -> "Model":
        return dataclasses.replace(
            self, masses=self.masses.copy(), evac=self.evac.copy(),
            radial=[[tuple(np.copy(x) if isinstance(x, np.ndarray) else x for x in t) for t in row]
                    for row in self.radial],
            angular=[[tuple(np.copy(x) if isinstance(x, np.ndarray) else x for x in t) for t in row]
                     for row in self.angular],
            nets=[dataclasses.replace(n, params=n.params.copy(), inpw=n.inpw.copy(),
                                      inpb=n.inpb.copy()) for n in self.nets])

    def flat(self) -> dict:
        """Flat arrays shared by the GPU descriptor and the oracle descriptor."""
        S = self.nspecies
        NP = S * (S + 1) // 2
        d = {}
        rn, rnm, roff, rs, reta = [], [], [], [], []
        for a in range(S):
            if len(self.radial[a]) != S or len(self.angular[a]) != NP:
                raise ValueError("basis tables must cover every species / species pair")
            for b in range(S):
                name, r_s, eta = self.radial[a][b]
                rnm.append(name)
                rn.append(len(r_s))
                roff.append(len(rs))
                rs.extend(np.asarray(r_s, float).tolist())
                reta.extend(np.asarray(eta, float).tolist())
        an, anm, aoff, aeta, azeta, alam = [], [], [], [], [], []
        for a in range(S):
            for p in range(NP):
                name, eta, zeta, lam = self.angular[a][p]
                anm.append(name)
                an.append(len(eta))
                aoff.append(len(aeta))
                aeta.extend(np.asarray(eta, float).tolist())
                azeta.extend(np.asarray(zeta, float).tolist())
                alam.extend(np.asarray(lam, int).tolist())
        nl, noff, nodes, neu, cc, poff, params, ioff, inpw, inpb, outw, outb = ([] for _ in range(12))
        for a, n in enumerate(self.nets):
            if n.nin != self.ninput(a):
                raise ValueError(f"network {a} has {n.nin} inputs, basis gives {self.ninput(a)}")
            if len(n.params) != n.nparams:
                raise ValueError(f"network {a} parameter count mismatch")
            nl.append(len(n.nodes))
            noff.append(len(nodes))
            nodes.extend(n.nodes)
            neu.append(n.neuron)
            cc.append(n.c)
            poff.append(len(params))
            params.extend(np.asarray(n.params, float).tolist())
            ioff.append(len(inpw))
            inpw.extend(np.asarray(n.inpw, float).tolist())
            inpb.extend(np.asarray(n.inpb, float).tolist())
            outw.append(float(n.outw))
            outb.append(float(n.outb))
        i32 = lambda x: np.ascontiguousarray(np.asarray(x, dtype=np.int32).reshape(-1))
        f64 = lambda x: np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
        d.update(
            rad_name=i32(rnm), rad_n=i32(rn), rad_off=i32(roff), rad_rs=f64(rs or [0.0]),
            rad_eta=f64(reta or [0.0]),
            ang_name=i32(anm), ang_n=i32(an), ang_off=i32(aoff), ang_eta=f64(aeta or [0.0]),
            ang_zeta=f64(azeta or [0.0]), ang_lambda=i32(alam or [0]),
            nn_nlayer=i32(nl), nn_nodes_off=i32(noff), nn_nodes=i32(nodes), nn_neuron=i32(neu),
            nn_c=f64(cc), nn_par_off=i32(poff), params=f64(params), nn_inp_off=i32(ioff),
            inpw=f64(inpw), inpb=f64(inpb), outw=f64(outw), outb=f64(outb), evac=f64(self.evac),
        )
        return d


def init_network(nin: int, hidden: List[int], neuron: int, rng: np.random.Generator,
                 c: float = 1.0, sigma_b: float = 0.001) -> Network:
    """Xavier initialisation as ANN::resize does it (src/ml/nn.cpp:1070-1156):
    biases normal with s = sigma_b*sqrt(2/(ni+no)), weights uniform on [-s, s]
    with s = sqrt(3*sigma_w)*sqrt(2/(ni+no)), sigma_w = 1, drawn biases first.
    The generator is numpy's, so values are seeded but not mt19937-identical."""
    nodes = list(hidden) + [1]
    bs, ws, prev = [], [], nin
    for h in nodes:
        bs.append(rng.normal(0.0, sigma_b * np.sqrt(2.0 / (prev + h)), size=h))
        prev = h
    prev = nin
    for h in nodes:
        s = np.sqrt(3.0) * np.sqrt(2.0 / (prev + h))
        ws.append(rng.uniform(-s, s, size=h * prev))   # column-major (n_out x n_in)
        prev = h
    params = np.concatenate(bs + ws)
    return Network(nin=nin, nodes=nodes, neuron=neuron, c=c, params=params,
                   inpw=np.ones(nin), inpb=np.zeros(nin))


def precondition(model: Model, G_rows: List[np.ndarray], prebias: str = "MEAN",
                 prescale: str = "DEV") -> None:
    """Input preconditioning from per-species input statistics
    (NNPTEFR::train, src/nnp/nnptefr2.cpp:681-869): inpb = -mean (MEAN),
    -0.5(max+min) (MID) or 0 (IDENTITY); inpw = 1/stddev (DEV, N-1 divisor),
    1/(max-min) (MINMAX) or 1 (IDENTITY); zero spreads give weight 1.  Output
    scaling is reset to outw=1, outb=0 as the trainer does (:872-878).
    ``G_rows[a]`` holds the inputs of every atom of species ``a`` (rows)."""
    for a, net in enumerate(model.nets):
        X = np.asarray(G_rows[a], dtype=np.float64)
        if X.shape[0] < 2:
            continue
        X = X[:, :net.nin]
        avg = X.mean(axis=0)
        dev = np.sqrt(((X - avg) ** 2).sum(axis=0) / (X.shape[0] - 1.0))
        mx, mn = X.max(axis=0), X.min(axis=0)
        if prebias == "MEAN":
            net.inpb = -avg
        elif prebias == "MID":
            net.inpb = -0.5 * (mx + mn)
        elif prebias == "IDENTITY":
            net.inpb = np.zeros(net.nin)
        else:
            raise ValueError("unsupported prebias")
        if prescale == "DEV":
            net.inpw = np.where(dev == 0, 1.0, 1.0 / np.where(dev == 0, 1.0, dev))
        elif prescale == "MINMAX":
            s = mx - mn
            net.inpw = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 1.0)
        elif prescale == "IDENTITY":
            net.inpw = np.ones(net.nin)
        else:
            raise ValueError("unsupported prescale")
        net.outw, net.outb = 1.0, 0.0
