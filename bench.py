"""Benchmark of the Behler-Parrinello energy/force path on one MI355X.

Workload (one step = one pass of the hot path over one batch, inputs resident
in HBM): 64 periodic water structures of 64 H2O (192 atoms, L = 12.42 A,
12,288 atoms), TANH cutoff 6 A, 80 inputs per element (16 radial x 2 + 16
angular x 3), 80-32-32-1 softplus networks, random init (no input
preconditioning).  A step is the neighbour build + symmetry functions + network +
force adjoint + force/energy/virial reduction (the E+F definition of
nnpce: atoms / wall time of neighbour build + G + E + F).

Reported: E+F atoms/s of the FAST form (``value``) and of the PLAIN FP32 form,
training structures/s of the energy+force training step on 128 Li/P/S
structures (96 inputs, 96-48-48-1 tanh, loss = energy MSE + 0.1 force MSE,
second backward, FP64 Adam), the roofline of the dominant kernel measured with
HIP events over the timed steps, and the FP64 CPU oracle timed on a bounded
sample of the same workload with every host core (with the GPU force error
on that sample).

    python bench.py [--gpus 1] [--steps 20] [--warmup 5]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "E+F atoms/s & train structs/s, mixed vs plain FP32 GPU at ≤1e-5 rel force err"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def flop_counts(model, batch, nl):
    """Algorithmic FLOPs of the basis kernels from the actual list:
    symm_fwd = P(10 + 6 nR) + T(8 + 4 nA), symm_bwd = P(20 + 6 nR) + T(52 + 8 nA)
    (together the E+F convention P(30 + 12 nR) + T(60 + 12 nA)), nR / nA the
    functions of the pair's / triplet's basis block; network 4 MAC per atom."""
    off = nl["off"]
    nsp = batch.species[nl["idx"]]
    S = model.nspecies
    fwd = bwd = 0.0
    P = T = 0
    for i in range(batch.ntotal):
        a = batch.species[i]
        cnt = np.bincount(nsp[off[i]:off[i + 1]], minlength=S)
        for b in range(S):
            nR = len(model.radial[a][b][1])
            fwd += cnt[b] * (10 + 6 * nR)
            bwd += cnt[b] * (20 + 6 * nR)
            P += cnt[b]
            for c in range(b, S):
                t = cnt[b] * (cnt[b] - 1) // 2 if b == c else cnt[b] * cnt[c]
                nA = len(model.angular[a][c * (c + 1) // 2 + b][1])
                fwd += t * (8 + 4 * nA)
                bwd += t * (52 + 8 * nA)
                T += t
    mac = 0
    for i in range(batch.ntotal):
        n = model.nets[batch.species[i]]
        prev = n.nin
        for h in n.nodes:
            mac += prev * h
            prev = h
    return {"symm_fwd": fwd, "symm_bwd": bwd, "net": 4.0 * mac, "pairs": P, "triplets": T}


def time_ef(eng, form, steps, warmup, out):
    import torch
    for _ in range(warmup):
        eng.neighbors()
        eng.energy_forces(form, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.neighbors()
        eng.energy_forces(form, out=out)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def run_train(steps, warmup, form):
    import torch
    from ann_amd import configs
    from ann_amd.gpu import Engine
    model = configs.model_lips()
    teacher = configs.model_lips(seed=33)
    batch = configs.batch_lips(128, seed=34)
    te = Engine(teacher)
    te.set_batch(batch)
    te.neighbors()
    Eref, Fref, _ = te.energy_forces(form, with_evac=False)
    Eref, Fref = Eref.clone(), Fref.clone()
    del te
    eng = Engine(model)
    eng.set_batch(batch)
    eng.neighbors()
    loss0 = None
    for _ in range(warmup):
        loss, _ = eng.train_step(Eref, Fref, 1.0, 0.1, 1e-3, form)
        if loss0 is None:
            loss0 = loss.cpu().numpy().copy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss, _ = eng.train_step(Eref, Fref, 1.0, 0.1, 1e-3, form)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return {"structs_per_s": batch.nstruc / dt, "ms_per_step": dt * 1e3, "batch": batch.nstruc,
            "atoms": batch.ntotal, "loss_first": (loss0.tolist() if loss0 is not None else None),
            "loss_last": loss.cpu().numpy().tolist()}


def cpu_baseline(model, batch, gpu_forces):
    """FP64 oracle (the port of the reference algorithm) on a bounded sample:
    the first structures of the batch, one per host thread (about 1 s of CPU
    work per structure), every host core.  Also returns the force error of each
    GPU form on that sample (the oracle as checker)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from ann_amd.configs import Batch
    cores = os.cpu_count() or 1
    nsample = int(min(batch.nstruc, max(8, cores)))
    sub = batch.subset(range(nsample))
    orc = pyoracle.Oracle(model)
    t0 = time.perf_counter()
    Eo, Fo, _ = orc.compute_batch(sub, nthreads=cores)
    dt = time.perf_counter() - t0
    errs = {}
    na = sub.ntotal
    for k, (E, F) in gpu_forces.items():
        Fs = F[:na]
        errs[k] = {"force_rel_rms": float(np.sqrt(((Fs - Fo) ** 2).sum() / (Fo ** 2).sum())),
                   "energy_per_atom": float(np.max(np.abs(E[:nsample] - Eo) / sub.natoms))}
    return {"value": sub.ntotal / dt, "unit": "atoms/s", "cores": cores, "kind": "port",
            "sample": f"{nsample} of the 64 water structures ({sub.ntotal} atoms), neighbour build + G + E + F, "
                      f"FP64, {cores} threads, {dt:.2f} s wall"}, errs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-train", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy three times (calibrates HBM counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 4}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (structures shard "
              "independently; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch
    from ann_amd import configs
    from ann_amd.gpu import Engine, FAST, PLAIN

    model = configs.model_water()
    batch = configs.batch_water(64, seed=12)
    eng = Engine(model)
    eng.set_batch(batch)
    eng.neighbors()
    nl = eng.nlist()
    flops = flop_counts(model, batch, nl)
    dev = eng.device
    out = (torch.empty(batch.nstruc, dtype=torch.float64, device=dev),
           torch.empty((batch.ntotal, 3), dtype=torch.float64, device=dev),
           torch.empty((batch.nstruc, 3, 3), dtype=torch.float64, device=dev))
    res = {}
    forces = {}
    for form, name in ((PLAIN, "plain"), (FAST, "fast")):
        eng.profile(True)
        dt = time_ef(eng, form, args.steps, args.warmup, out)
        kern = {}
        for k in ("nb_bin", "nb_search", "nb_compact", "symm_fwd_plain", "symm_fwd_fast", "symm_bwd_plain",
                  "symm_bwd_fast", "net_plain", "net_mfma", "force_gather", "struc_sums"):
            ms, n = eng.profile_get(k)
            if n:
                kern[k] = {"avg_us": ms / n * 1e3, "launches": n}
        eng.profile(False)
        res[name] = {"ms_per_step": dt * 1e3, "atoms_per_s": batch.ntotal / dt, "kernels": kern}
        E, F, _ = eng.energy_forces(form)
        forces[name] = (E.cpu().numpy(), F.cpu().numpy())
    # dominant kernel of the FAST form
    fk = res["fast"]["kernels"]
    dom = max((k for k in fk if k.startswith("symm")), key=lambda k: fk[k]["avg_us"])
    fl = flops["symm_bwd"] if "bwd" in dom else flops["symm_fwd"]
    achieved = fl / (fk[dom]["avg_us"] * 1e-6) / 1e12
    roof = {"bound": "valu", "kernel": dom, "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
            "algorithmic_flop_per_launch": fl, "avg_launch_us": fk[dom]["avg_us"],
            "note": "FP32 vector ALU bound; peak is the MI355X FP32 figure (packed FP32 disabled in this build, "
                    "so the attainable non-packed VALU rate is half of it)"}
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        tj = json.load(open(traffic_file))
        if dom in tj:
            roof["traffic"] = tj[dom]["bytes_per_launch"]
    train = None
    if not args.no_train:
        train = run_train(max(3, args.steps // 2), max(1, args.warmup // 2), FAST)
        train_plain = run_train(max(3, args.steps // 4), 1, PLAIN)
        train["plain_structs_per_s"] = train_plain["structs_per_s"]
    cpu, errs = (None, {}) if args.no_cpu else cpu_baseline(model, batch, forces)
    line = {
        "metric": METRIC,
        "value": res["fast"]["atoms_per_s"],
        "unit": "atoms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["fast"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "water 64 x 192 atoms E+F (neighbour build + G + E + F), FAST form",
                   "global_batch": batch.nstruc, "atoms": batch.ntotal, "pairs": flops["pairs"],
                   "triplets": flops["triplets"], "parallelism": "single GPU"},
        "plain": {"atoms_per_s": res["plain"]["atoms_per_s"], "ms_per_step": res["plain"]["ms_per_step"]},
        "fast_over_plain": res["fast"]["atoms_per_s"] / res["plain"]["atoms_per_s"],
        "kernels": {"fast": res["fast"]["kernels"], "plain": res["plain"]["kernels"]},
        "roofline": roof,
        "train": train,
        "cpu_baseline": cpu,
        "errors_vs_fp64_oracle": errs,
    }
    def _np(o):
        if isinstance(o, np.integer):
            return int(o)
        if isinstance(o, np.floating):
            return float(o)
        raise TypeError(type(o))
    print(json.dumps(line, default=_np), flush=True)


if __name__ == "__main__":
    main()
