"""The reference's NNP potential text format (read and write).

Layout (NNP::write / NNP::read, src/nnp/nnp.cpp:197-307):

    ann
    ntypes S
    name <X> mass <m> energy <E_vac>            (Type::write, src/nnp/type.cpp:149-164), S lines
    rc <rc>
    basis <X>                                     for every centre species X
      basis_radial <Y>                            for every neighbour species Y
      BasisR <CUT> <rc> <NAME> <n>                (BasisR::write, src/nnp/basis_radial.cpp:108-116)
        <rs> <eta>                                n lines
      basis_angular <Y> <Z>                       for Y <= Z in FILE order (0,0),(0,1),...,(1,1),...
      BasisA <CUT> <rc> <NAME> <n>                (BasisA::write, src/nnp/basis_angular.cpp:101-109)
        <eta> <zeta> <lambda> <rflag>             n lines
    nn <X>                                        for every species
    nn <nin> <h1> ... <nout>                      (ANN::write, src/ml/nn.cpp:1436-1477)
    neuron <NAME> <c>
    inpw ..., outw ..., inpb ..., outb ...
    bias[l] ...                                   one line per layer
    weight[l,l+1] ...                             row-major (output neuron, input)

The angular blocks are written in file order but stored in TensorTri order
(pair index k(k+1)/2 + j); the two differ for three or more species.
"""
from __future__ import annotations

from typing import List

import numpy as np

from .model import ANG, CUT, NEURON, RAD, Model, Network, pair_index

_CUT = {v: k for k, v in CUT.items()}
_RAD = {v: k for k, v in RAD.items()}
_ANG = {v: k for k, v in ANG.items()}
_NEU = {v: k for k, v in NEURON.items()}


def _f(x: float) -> str:
    return f"{x:.15f}"


def write(path: str, m: Model) -> None:
    S = m.nspecies
    out: List[str] = ["ann", f"ntypes {S}"]
    for a in range(S):
        out.append(f"name {m.species[a]} mass {m.masses[a]:f} energy {m.evac[a]:f} ")
    out.append(f"rc {m.rc:f}")
    for a in range(S):
        out.append(f"basis {m.species[a]}")
        for b in range(S):
            name, rs, eta = m.radial[a][b]
            out.append(f"basis_radial {m.species[b]}")
            out.append(f"BasisR {_CUT[m.cutoff]} {m.rc:f} {_RAD[name]} {len(rs)}")
            out += [f"\t{_f(r)} {_f(e)}" for r, e in zip(rs, eta)]
        for j in range(S):
            for k in range(j, S):
                name, eta, zeta, lam = m.angular[a][pair_index(j, k)]
                out.append(f"basis_angular {m.species[j]} {m.species[k]}")
                out.append(f"BasisA {_CUT[m.cutoff]} {m.rc:f} {_ANG[name]} {len(eta)}")
                # rflag: index of the width group (consecutive equal widths share one)
                flag, flags = 0, []
                for n in range(len(eta)):
                    if n > 0 and abs(eta[n] - eta[n - 1]) >= 1e-6:
                        flag += 1
                    flags.append(flag)
                out += [f"\t{_f(e)} {_f(z)} {int(l)} {f}" for e, z, l, f in zip(eta, zeta, lam, flags)]
    for a in range(S):
        n = m.nets[a]
        out.append(f"nn {m.species[a]}")
        out.append("nn " + " ".join(str(x) for x in [n.nin] + list(n.nodes)) + " ")
        out.append(f"neuron {_NEU[n.neuron]} {n.c:f}")
        out.append("inpw " + " ".join(_f(x) for x in n.inpw) + " ")
        out.append(f"outw {_f(n.outw)} ")
        out.append("inpb " + " ".join(_f(x) for x in n.inpb) + " ")
        out.append(f"outb {_f(n.outb)} ")
        layers = n.layers()
        for l, (W, b) in enumerate(layers):
            out.append(f"bias[{l + 1}] " + " ".join(_f(x) for x in b) + " ")
        for l, (W, b) in enumerate(layers):
            out.append(f"weight[{l},{l + 1}] " + " ".join(_f(x) for x in W.reshape(-1)) + " ")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def read(path: str) -> Model:
    lines = [l.split() for l in open(path) if l.strip()]
    it = iter(lines)
    next(it)                                    # header
    S = int(next(it)[1])
    names, masses, evac = [], [], []
    for _ in range(S):
        tok = next(it)
        d = {tok[q].upper(): tok[q + 1] for q in range(0, len(tok) - 1, 2)}
        names.append(d["NAME"])
        masses.append(float(d.get("MASS", 0.0)))
        evac.append(float(d.get("ENERGY", 0.0)))
    index = {n: i for i, n in enumerate(names)}
    rc = float(next(it)[1])
    radial = [[None] * S for _ in range(S)]
    angular = [[None] * (S * (S + 1) // 2) for _ in range(S)]
    cut = None
    for _ in range(S):
        a = index[next(it)[1]]
        for _ in range(S):
            b = index[next(it)[1]]
            h = next(it)
            cut = CUT[h[1]]
            nf = int(h[4])
            vals = np.array([[float(x) for x in next(it)[:2]] for _ in range(nf)]).reshape(nf, 2)
            radial[a][b] = (RAD[h[3]], vals[:, 0].copy(), vals[:, 1].copy())
        for _ in range(S * (S + 1) // 2):
            tok = next(it)
            j, k = index[tok[1]], index[tok[2]]
            h = next(it)
            nf = int(h[4])
            rows = [next(it) for _ in range(nf)]
            eta = np.array([float(r[0]) for r in rows])
            zeta = np.array([float(r[1]) for r in rows])
            lam = np.array([int(r[2]) for r in rows], dtype=int)
            angular[a][pair_index(j, k)] = (ANG[h[3]], eta, zeta, lam)
    nets: List[Network] = [None] * S
    for _ in range(S):
        a = index[next(it)[1]]
        dims = [int(x) for x in next(it)[1:]]
        nin, nodes = dims[0], dims[1:]
        tok = next(it)
        neuron, c = NEURON[tok[1]], float(tok[2])
        inpw = np.array([float(x) for x in next(it)[1:]])
        outw = float(next(it)[1])
        inpb = np.array([float(x) for x in next(it)[1:]])
        outb = float(next(it)[1])
        bs = [np.array([float(x) for x in next(it)[1:]]) for _ in nodes]
        ws, prev = [], nin
        for h in nodes:
            W = np.array([float(x) for x in next(it)[1:]]).reshape(h, prev)
            ws.append(W.T.reshape(-1))          # column-major packing
            prev = h
        nets[a] = Network(nin=nin, nodes=nodes, neuron=neuron, c=c, params=np.concatenate(bs + ws),
                          inpw=inpw, inpb=inpb, outw=outw, outb=outb)
    return Model(names, np.array(masses), np.array(evac), rc, cut, radial, angular, nets)
