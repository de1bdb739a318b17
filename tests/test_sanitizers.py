"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only):
the FP64 oracle (neighbour lists, symmetry functions, forces, the training
gradient, Adam) and the C++ drop-in's file handling (NNP::read / write /
read_basis through the driver's roundtrip and basis modes, which make no GPU
call),
each built with -fsanitize=address,undefined and run with the sanitizer
runtime preloaded.  Device code is not instrumented (no GPU sanitizer here)."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _asan_runtime():
    p = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = p.stdout.strip()
    if p.returncode != 0 or not os.path.isabs(path) or not os.path.exists(path):
        pytest.skip("no AddressSanitizer runtime")
    return path


def _env(extra):
    env = dict(os.environ, LD_PRELOAD=_asan_runtime(),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    env.update(extra)
    return env


def _clean(out):
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-3000:]


ORACLE_RUN = r"""
import sys
sys.path.insert(0, {root!r})
sys.path.insert(0, {oracle!r})
import numpy as np
import pyoracle as P
from ann_amd import configs
for m, b in ((configs.model_water(), configs.batch_water(2, seed=3, nmol=16, L=8.0)),
             (configs.model_lips(), configs.batch_lips(1, seed=4)),
             (configs.model_family("SECH", "STUDENT4", cutoff="POLY3"), configs.batch_water(1, seed=5, nmol=12, L=7.0))):
    orc = P.Oracle(m)
    for k in range(b.nstruc):
        pos, cell, sp = b.struc(k)
        nl = orc.nlist(pos, cell, sp)
        G = orc.symm(pos, cell, sp, nl)
        E, dE = orc.net(sp, G)
        F, V, Pp = orc.force(pos, cell, sp, dE, nl=nl)
        orc.tangent(pos, cell, sp, np.ones((len(sp), 3)) * 0.01, nl=nl)
    Eref = np.zeros(b.nstruc)
    Fref = np.zeros((b.ntotal, 3))
    for loss in (1, 2, 3, 4):
        orc.train_loss_grad(b, Eref, Fref, loss=loss)
    g = np.ones(8)
    P.adam_step(np.zeros(8), np.zeros(8), np.zeros(8), g, 1e-3, 1e-8, 1)
print("oracle ok")
"""


def test_oracle_under_sanitizers(tmp_path):
    _asan_runtime()
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    lib = os.path.join(ROOT, "oracle", "_san", "liboracle.so")
    script = tmp_path / "run.py"
    script.write_text(ORACLE_RUN.format(root=ROOT, oracle=os.path.join(ROOT, "oracle")))
    p = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=600,
                       env=_env({"ORACLE_LIB": lib}))
    print(p.stdout[-2000:], p.stderr[-3000:])
    assert p.returncode == 0 and "oracle ok" in p.stdout
    _clean(p.stdout + p.stderr)


def test_cpp_file_handling_under_sanitizers(tmp_path):
    """NNP::read / write / read_basis of the drop-in classes, built with the
    sanitizers: two potentials round trip, one basis grows an input count."""
    _asan_runtime()
    if not shutil.which("g++") or not os.path.exists(os.path.join(ROOT, "ann_amd", "libann_amd.so")):
        pytest.skip("needs g++ and the built library")
    cpp = os.path.join(ROOT, "ann_amd", "cpp")
    srcs = [os.path.join(cpp, f) for f in sorted(os.listdir(cpp)) if f.endswith(".cpp")]
    exe = str(tmp_path / "drv")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-D__HIP_PLATFORM_AMD__", "-I" + cpp, "-I" + os.path.join(ROOT, "include"),
           "-I/opt/rocm/include", "-o", exe, os.path.join(ROOT, "tests", "cpp", "test_nnp_gpu.cpp")] + srcs + \
          ["-L" + os.path.join(ROOT, "ann_amd"), "-Wl,-rpath," + os.path.join(ROOT, "ann_amd"), "-lann_amd",
           "-L/opt/rocm/lib", "-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    from ann_amd import configs, potfile
    for make in (configs.model_water, configs.model_lips):
        a, b = tmp_path / "a.nnp", tmp_path / "b.nnp"
        potfile.write(str(a), make())
        p = subprocess.run([exe, "roundtrip", str(a), str(b)], capture_output=True, text=True, timeout=120,
                           env=_env({}))
        print(p.stdout, p.stderr[-3000:])
        assert p.returncode == 0
        _clean(p.stdout + p.stderr)
        assert potfile.read(str(b)).nspecies == make().nspecies
    # NNP::read_basis: the O basis of a grown water model into the water potential
    m = configs.model_water()
    m2 = m.copy()
    name, rs, eta = m2.radial[1][0]
    m2.radial[1][0] = (name, list(rs) + [rs[-1] + 0.3], list(eta) + [eta[-1]])
    full = tmp_path / "full.nnp"
    potfile.write(str(full), m2)
    lines = [l.rstrip("\n") for l in open(full)]
    s0 = lines.index("basis O")
    s1 = next(q for q in range(s0 + 1, len(lines)) if lines[q].startswith("nn "))
    basis = tmp_path / "basis.txt"
    basis.write_text(f"rc {m.rc:f}\nntypes 2\n" + "\n".join(lines[s0:s1]) + "\n")
    a, c = tmp_path / "w.nnp", tmp_path / "w2.nnp"
    potfile.write(str(a), m)
    p = subprocess.run([exe, "basis", str(a), str(basis), "O", str(c)], capture_output=True, text=True,
                       timeout=120, env=_env({}))
    print(p.stdout, p.stderr[-3000:])
    assert p.returncode == 0
    _clean(p.stdout + p.stderr)
    out = potfile.read(str(c))
    assert out.ninput(1) == m.ninput(1) + 1 and out.ninput(0) == m.ninput(0)
