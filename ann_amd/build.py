"""Build step of the package: compiles the HIP sources of ann_amd/csrc into
ann_amd/libann_amd.so for gfx950 (one object per source, in parallel, then
linked) and the C++ host shim ann_amd/cpp into ann_amd/libann_nnp.so.

Usage: python -m ann_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
CPP = os.path.join(HERE, "cpp")
LIB = os.path.join(HERE, "libann_amd.so")
SHIM = os.path.join(HERE, "libann_nnp.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed-FP32 instructions are disabled for all device code
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-value", "-Wno-unused-result", "-I" + os.path.join(ROOT, "include")]


# FP64 pair vectors and the inclusion test must round like the oracle (no
# fused multiply-add); the intrinsics' own bodies are outside any pragma scope
EXTRA = {"neighbor.hip": ["-ffp-contract=off"], "md.hip": ["-ffp-contract=off"]}


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def build_hip(jobs: int = 8) -> str:
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    with tempfile.TemporaryDirectory(prefix="ann_build_") as tmp:
        objs = [os.path.join(tmp, s[:-4] + ".o") for s in srcs]
        cmds = [[HIPCC] + HIPFLAGS + EXTRA.get(s, []) + ["-c", os.path.join(CSRC, s), "-o", o]
                for s, o in zip(srcs, objs)]
        with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
            list(ex.map(_run, cmds))
        tmp_lib = LIB + ".tmp"
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp_lib] + objs)
        os.replace(tmp_lib, LIB)
    return LIB


def build_shim() -> str:
    """C++ host classes mirroring the reference interface (NNP, NeighborList,
    Structure) over the C ABI; plain g++ plus the HIP runtime for streams."""
    srcs = [os.path.join(CPP, f) for f in sorted(os.listdir(CPP)) if f.endswith(".cpp")] if os.path.isdir(CPP) else []
    if not srcs:
        return ""
    tmp_lib = SHIM + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__",
          "-I" + os.path.join(ROOT, "include"), "-I" + CPP, "-I/opt/rocm/include", "-o", tmp_lib] + srcs +
         ["-L" + HERE, "-Wl,-rpath,$ORIGIN", "-lann_amd", "-L/opt/rocm/lib", "-lamdhip64"])
    os.replace(tmp_lib, SHIM)
    # the C++ test driver of the drop-in classes (tests/cpp)
    tdir = os.path.join(ROOT, "tests", "cpp")
    for name in ("test_nnp_gpu", "test_classes_gpu"):
        tsrc = os.path.join(tdir, name + ".cpp")
        if os.path.exists(tsrc):
            _run(["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I" + CPP, "-I" + os.path.join(ROOT, "include"),
                  "-I/opt/rocm/include", "-o", os.path.join(tdir, name), tsrc,
                  "-L" + HERE, "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/../../ann_amd", "-lann_nnp", "-lann_amd",
                  "-L/opt/rocm/lib", "-lamdhip64"])
    return SHIM


def main() -> None:
    print(build_hip(), flush=True)
    s = build_shim()
    if s:
        print(s, flush=True)


if __name__ == "__main__":
    sys.exit(main())
