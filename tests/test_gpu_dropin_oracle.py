"""The C++ drop-in classes against the FP64 oracle: every NNP entry point
(src/nnp/nnp.cpp:333-665) with a NeighborList and with a verlet::List
(list built at rc + skin on positions 0, evaluated at displaced positions 1,
the reuse pattern of src/torch/engine.cpp:64-125), and NNP::read_basis
(src/nnp/nnp.cpp:96-156) with the same and with a changed input count.

Bars as in the C-ABI parity tests: symmetry functions max |dG| / max |G|
<= 1e-5 (inputs evaluated in FP32 unless their scaling needs FP64),
energies <= 1e-6 eV/atom, forces relative RMS <= 1e-5."""
import os
import subprocess

import numpy as np
import pytest

from ann_amd import configs, model as M, potfile

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "test_nnp_gpu")
SKIN = 1.0


def _structure(seed):
    b = configs.batch_water(1, seed=seed, nmol=16, L=8.0)
    pos, _, sp = b.struc(0)
    L = 8.0
    cell = np.array([[L, 1.5, 0.7], [0.0, L, -1.1], [0.0, 0.0, L]])     # columns = lattice vectors
    rng = np.random.default_rng(seed + 100)
    pos1 = pos + rng.uniform(-0.2, 0.2, size=pos.shape)                 # |d| < skin / 2
    return np.asarray(pos, float), pos1, cell, np.asarray(sp, np.int32)


def _write_struc(path, pos0, pos1, cell, sp):
    with open(path, "w") as fh:
        fh.write(f"{len(sp)}\n" + " ".join(f"{x:.17g}" for x in cell.reshape(-1)) + "\n")
        for i in range(len(sp)):
            fh.write(f"{sp[i]} " + " ".join(f"{x:.17g}" for x in pos0[i]) + " "
                     + " ".join(f"{x:.17g}" for x in pos1[i]) + "\n")


def _write_basis(path, m, a):
    """The basis block of species a in the layout NNP::read_basis reads."""
    lines = [l.rstrip("\n") for l in open(path + ".full")]
    start = lines.index(f"basis {m.species[a]}")
    end = start + 1
    while end < len(lines) and not lines[end].startswith("basis ") and not lines[end].startswith("nn "):
        end += 1
    with open(path, "w") as fh:
        fh.write(f"rc {m.rc:f}\nntypes {m.nspecies}\n" + "\n".join(lines[start:end]) + "\n")


def _read_dump(path):
    out, tag = {}, None
    for line in open(path):
        t = line.split()
        if len(t) == 1 and not t[0][0].isdigit() and t[0][0] != "-":
            tag = t[0]
            out[tag] = []
        elif t:
            out[tag].append([float(x) for x in t])
    res = {}
    for k, v in out.items():
        if len(v) == 1 and len(v[0]) == 1:
            res[k] = v[0][0]
        else:                               # rows of species with fewer inputs padded with zeros
            w = max(len(r) for r in v)
            res[k] = np.array([r + [0.0] * (w - len(r)) for r in v])
    return res


def _run(tmp_path, pot_model, pos0, pos1, cell, sp, basis=None):
    pot = tmp_path / "pot.nnp"
    potfile.write(str(pot), pot_model)
    st = tmp_path / "struc.txt"
    _write_struc(str(st), pos0, pos1, cell, sp)
    out = tmp_path / "out.txt"
    cmd = [DRIVER, "dump", str(pot), str(st), str(out), str(SKIN)] + (basis or [])
    print(" ".join(cmd[1:2]) + f" natoms {len(sp)}", flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    return _read_dump(str(out))


def _check(orc, m, d, pos, cell, sp, suffix):
    N = len(sp)
    Go = orc.symm(pos, cell, sp)
    G = d["G" + suffix]
    assert G.shape == (N, max(m.ninput(a) for a in set(sp.tolist())))
    errg = np.abs(G - Go[:, :G.shape[1]]).max() / np.abs(Go).max()
    assert errg <= 1e-5, errg
    Eo, _, Fo, _ = orc.compute(pos, cell, sp)
    evac = float(sum(m.evac[s] for s in sp))
    rel = lambda F: float(np.sqrt(((F - Fo) ** 2).sum() / (Fo ** 2).sum()))
    print(f"  {suffix}: G {errg:.2e} E {abs(d['E' + suffix] - Eo) / N:.2e} "
          f"Fforce {rel(d['Fforce' + suffix]):.2e} Fcompute {rel(d['Fcompute' + suffix]):.2e}", flush=True)
    assert abs(d["E" + suffix] - Eo) / N <= 1e-6
    assert abs(d["pe" + suffix] + evac - Eo) / N <= 1e-6
    assert rel(d["Fforce" + suffix]) <= 1e-5
    assert rel(d["Fcompute" + suffix]) <= 1e-5


@pytest.mark.parametrize("seed", [41, 42])
def test_entry_points_match_oracle(tmp_path, oracle_mod, seed):
    m = configs.model_water()
    m.evac = m.evac + np.array([0.3, -0.7])
    pos0, pos1, cell, sp = _structure(seed)
    d = _run(tmp_path, m, pos0, pos1, cell, sp)
    orc = oracle_mod.Oracle(m)
    _check(orc, m, d, pos0, cell, sp, "0")
    _check(orc, m, d, pos1, cell, sp, "1")
    Eo0 = orc.compute(pos0, cell, sp)[0]
    assert abs(d["pev0"] + float(sum(m.evac[s] for s in sp)) - Eo0) / len(sp) <= 1e-6
    # the Verlet list holds the pairs within rc + skin
    nl = orc.nlist(pos0, cell, sp, rc=m.rc + SKIN)
    assert int(d["vpairs"]) == nl.npairs


@pytest.mark.parametrize("grow", [False, True])
def test_read_basis(tmp_path, oracle_mod, grow):
    m = configs.model_water()
    m2 = m.copy()
    a = 1
    rng = np.random.default_rng(7)
    for b in range(m.nspecies):
        name, rs, eta = m2.radial[a][b]
        if grow:
            rs, eta = np.append(rs, rs[-1] + 0.4), np.append(eta, eta[-1])
        m2.radial[a][b] = (name, rs * (1 + 0.02 * rng.standard_normal(len(rs))), eta * 1.1)
    if grow:
        # the species' network follows the new input count: zero parameters, identity scaling
        n = m2.nets[a]
        nin = m2.ninput(a)
        net = M.Network(nin, list(n.nodes), n.neuron, n.c, np.zeros(0), np.ones(nin), np.zeros(nin), n.outw, n.outb)
        net.params = np.zeros(net.nparams)
        m2.nets[a] = net
    bpath = str(tmp_path / "basis.txt")
    potfile.write(bpath + ".full", m2)
    _write_basis(bpath, m2, a)
    pos0, pos1, cell, sp = _structure(43)
    d = _run(tmp_path, m, pos0, pos1, cell, sp, [bpath, m.species[a]])
    orc = oracle_mod.Oracle(m2)
    _check(orc, m2, d, pos0, cell, sp, "0")
    _check(orc, m2, d, pos1, cell, sp, "1")
