"""The C ABI library loads without a GPU and exports every entry point that
include/ann_gpu.h declares; host-side descriptors are consistent."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(ann_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from ann_amd import gpu
    L = gpu.lib()
    names = declared("ann_gpu.h")
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert sorted(gpu.EXPORTS) == names


def test_error_path_without_context():
    from ann_amd import gpu
    L = gpu.lib()
    rc = L.ann_set_stream(None, None)
    assert rc == -1
    assert b"null" in L.ann_last_error()


def test_model_flat_layout():
    from ann_amd import configs
    m = configs.model_water()
    f = m.flat()
    S, NP = 2, 3
    assert len(f["rad_n"]) == S * S and len(f["ang_n"]) == S * NP
    assert m.ninput(0) == 80 and m.ninput(1) == 80
    assert len(f["params"]) == sum(n.nparams for n in m.nets) == 2 * (32 + 32 + 1 + 80 * 32 + 32 * 32 + 32)
    offr, offa = m.offsets(0)
    assert offr == [0, 16] and offa == [32, 48, 64]


def test_workload_generators():
    from ann_amd import configs
    b = configs.batch_water(2)
    assert b.ntotal == 384 and np.all(b.natoms == 192)
    assert np.bincount(b.species) .tolist() == [256, 128]
    p = b.pos[:3]
    assert abs(np.linalg.norm(p[1] - p[0]) - 0.96) < 1e-12
    cosang = np.dot(p[1] - p[0], p[2] - p[0]) / 0.96 ** 2
    assert abs(np.degrees(np.arccos(cosang)) - 104.5) < 1e-9
    u = configs.batch_uneven(20)
    assert np.all(u.natoms % 3 == 0) and u.natoms.min() >= 30 and u.natoms.max() <= 1026
    dens = [3 * 18.015 / 6.02214076e23 / (np.linalg.det(c) * 1e-24) * (u.natoms[k] / 3) / u.natoms[k] * 3 / 3
            for k, c in enumerate(u.cells)]
    l = configs.batch_lips(10)
    assert np.all((l.natoms >= 96) & (l.natoms <= 256))
    a = configs.struc_ar256()
    assert a.ntotal == 256


def test_precondition_statistics():
    from ann_amd import configs
    from ann_amd.model import precondition
    m = configs.model_water()
    rng = np.random.default_rng(0)
    rows = [rng.normal(3.0, 2.0, (500, 80)), rng.normal(-1.0, 0.5, (400, 80))]
    precondition(m, rows)
    s = m.nets[0].inpw * (rows[0] + m.nets[0].inpb)
    assert np.allclose(s.mean(0), 0, atol=1e-12) and np.allclose(s.std(0, ddof=1), 1, atol=1e-12)
