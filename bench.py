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


def work_model(model, batch, nl):
    """Algorithmic work of one launch of each kernel family, from the actual
    neighbour list (P slots, T = sum_i n_i(n_i-1)/2 triplets, nR / nA the
    functions of a pair's / triplet's block, FMA = 2 FLOP):

    triplet tiles and PLAIN form (the E+F convention P(30 + 12 nR) + T(60 + 12 nA)):
      symm_fwd = P(10 + 6 nR) + T(8 + 4 nA),  symm_bwd = P(20 + 6 nR) + T(52 + 8 nA)
    moment form (M = (Z+1)(Z+2)(Z+3)/6 monomials, U widths of the centre species,
    NP species pairs, nk (lambda, zeta) combinations, S species):
      symm_fwd_mom = P(10 + 6 nR) + sum_i [n_i (M(2 + 2U) + 2U) + NP U 3M + sum_f 2(zeta_f + 1)]
      symm_bwd_mom = P(20 + 6 nR) + sum_i [n_i (M(4U + 16) + 20) + NP U (Z+1) 2 nk + S M U 2S]
                     (the moments themselves come from the G kernel)
      symm_tan_mom = P(20 + 8 nR) + sum_i [n_i M (9 + 6U) + NP U 6M + sum_f 2(zeta_f + 1)]
      symm_tan (triplet tiles / PLAIN) = symm_fwd + symm_bwd
    networks: 4 MAC per atom and parameter-matrix entry (forward + input backward);
    second backward (net_bwd2) 8 MAC, parameter-gradient GEMMs (grad_reduce) 4 MAC;
    HBM kernels in bytes: nb_search 24 B/atom + 8 B/slot, nb_compact (CSR fill) 28 B/atom
    + 37 B/slot, nb_rev (reverse slots) 20 B/slot, force_gather 96 B/atom + 44 B/slot,
    struc_sums 76 B/atom."""
    off = nl["off"]
    nsp = batch.species[nl["idx"]]
    S = model.nspecies
    NP = S * (S + 1) // 2
    fwd = bwd = fwd_m = bwd_m = tan_m = 0.0
    P = T = 0
    for i in range(batch.ntotal):
        a = batch.species[i]
        cnt = np.bincount(nsp[off[i]:off[i + 1]], minlength=S)
        n_i = int(cnt.sum())
        pair = 0.0
        for b in range(S):
            nR = len(model.radial[a][b][1])
            fwd += cnt[b] * (10 + 6 * nR)
            bwd += cnt[b] * (20 + 6 * nR)
            pair += cnt[b] * nR
            P += cnt[b]
            for c in range(b, S):
                t = cnt[b] * (cnt[b] - 1) // 2 if b == c else cnt[b] * cnt[c]
                nA = len(model.angular[a][c * (c + 1) // 2 + b][1])
                fwd += t * (8 + 4 * nA)
                bwd += t * (52 + 8 * nA)
                T += t
        blocks = [model.angular[a][p] for p in range(NP)]
        Z = int(max(max(bl[2]) for bl in blocks if len(bl[2])))
        M = (Z + 1) * (Z + 2) * (Z + 3) // 6
        U = len(set(np.concatenate([bl[1] for bl in blocks]).tolist()))
        nk = max(len(set(zip(bl[2].tolist(), bl[3].tolist()))) for bl in blocks)
        zsum = sum(float((bl[2] + 1).sum()) for bl in blocks)
        fwd_m += n_i * 10 + 6 * pair + n_i * (M * (2 + 2 * U) + 2 * U) + NP * U * 3 * M + 2 * zsum
        # the adjoint reads the moments the G kernel stored (no moment sums of its own)
        bwd_m += n_i * 20 + 6 * pair + n_i * (M * (4 * U + 16) + 20) \
            + NP * U * (Z + 1) * 2 * nk + S * M * U * 2 * S
        # tangent (forward-mode G): per (neighbour, monomial) the monomial, its
        # directional derivative and U value / derivative accumulations
        tan_m += n_i * 20 + 8 * pair + n_i * M * (9 + 6 * U) + NP * U * 6 * M + 2 * zsum
    mac = 0
    for i in range(batch.ntotal):
        n = model.nets[batch.species[i]]
        prev = n.nin
        for h in n.nodes:
            mac += prev * h
            prev = h
    N = batch.ntotal
    return {"symm_fwd": fwd, "symm_bwd": bwd, "symm_fwd_mom": fwd_m, "symm_bwd_mom": bwd_m,
            "symm_tan": fwd + bwd, "symm_tan_mom": tan_m, "net": 4.0 * mac,
            # fused E+F: G, network (forward + dE/dG) and adjoint of each atom in one kernel
            "ef_fused": fwd_m + bwd_m + 4.0 * mac,
            "net_bwd2": 8.0 * mac, "grad_reduce": 4.0 * mac,
            "pairs": P, "triplets": T,
            "bytes": {"nb_search": 24.0 * N + 8.0 * P, "nb_compact": 28.0 * N + 37.0 * P, "nb_rev": 20.0 * P,
                      "force_gather": 96.0 * N + 44.0 * P, "struc_sums": 76.0 * N}}


MFMA_BF16_PEAK_TFLOPS = 2500.0


def kernel_rooflines(kern, work):
    """Achieved rate and fraction of the MI355X peak for each timed kernel."""
    out = {}
    for k, v in kern.items():
        sec = v["avg_us"] * 1e-6
        if k.startswith("symm_"):
            base = k.rsplit("_", 1)[0]            # symm_fwd / symm_bwd
            fl = work[base + "_mom"] if k.endswith("_mom") else work[base]
            ach = fl / sec / 1e12
            out[k] = {"bound": "valu", "work": fl, "unit": "TFLOP/s", "achieved": ach, "peak": FP32_PEAK_TFLOPS,
                      "frac": ach / FP32_PEAK_TFLOPS}
        elif k == "ef_fused":
            ach = work[k] / sec / 1e12
            out[k] = {"bound": "valu", "work": work[k], "unit": "TFLOP/s", "achieved": ach, "peak": FP32_PEAK_TFLOPS,
                      "frac": ach / FP32_PEAK_TFLOPS}
        elif k in ("net_bwd2", "grad_reduce"):
            # FP32 matrix-core products (v_mfma_f32_16x16x4_f32); peak = the FP32 figure
            ach = work[k] / sec / 1e12
            out[k] = {"bound": "mfma", "work": work[k], "unit": "TFLOP/s", "achieved": ach,
                      "peak": FP32_PEAK_TFLOPS, "frac": ach / FP32_PEAK_TFLOPS}
        elif k.startswith("net_"):
            ach = work["net"] / sec / 1e12
            peak = MFMA_BF16_PEAK_TFLOPS if k == "net_mfma" else FP32_PEAK_TFLOPS
            out[k] = {"bound": "mfma" if k == "net_mfma" else "valu", "work": work["net"], "unit": "TFLOP/s",
                      "achieved": ach, "peak": peak, "frac": ach / peak}
        elif k in work["bytes"]:
            ach = work["bytes"][k] / sec / 1e9
            out[k] = {"bound": "hbm", "work": work["bytes"][k], "unit": "GB/s", "achieved": ach,
                      "peak": HBM_PEAK_GBS, "frac": ach / HBM_PEAK_GBS}
    return out


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


def train_setup(nstruc=128, seed=34):
    """Li/P/S training workload: model, seeded teacher labels (E without vacuum
    energies, F) from the teacher network on the GPU."""
    from ann_amd import configs
    from ann_amd.gpu import Engine, FAST
    model = configs.model_lips()
    teacher = configs.model_lips(seed=33)
    batch = configs.batch_lips(nstruc, seed=seed)
    te = Engine(teacher)
    te.set_batch(batch)
    te.neighbors()
    Eref, Fref, _ = te.energy_forces(FAST, with_evac=False)
    Eref, Fref = Eref.clone(), Fref.clone()
    del te
    return model, batch, Eref, Fref


def run_train(steps, warmup, form, graph=False, kernels=False):
    import torch
    from ann_amd.gpu import Engine, FAST
    model, batch, Eref, Fref = train_setup()
    eng = Engine(model)
    eng.set_batch(batch)
    eng.neighbors()
    if form == FAST:
        eng.autotune(FAST)
    if graph:
        eng.train_graph(True)
    out = (torch.empty(2, dtype=torch.float64, device=eng.device),
           torch.empty(eng.nparams, dtype=torch.float64, device=eng.device))
    loss0 = None
    for _ in range(max(warmup, 2)):
        loss, _ = eng.train_step(Eref, Fref, 1.0, 0.1, 1e-3, form, out=out)
        if loss0 is None:
            loss0 = loss.cpu().numpy().copy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss, _ = eng.train_step(Eref, Fref, 1.0, 0.1, 1e-3, form, out=out)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    res = {"structs_per_s": batch.nstruc / dt, "ms_per_step": dt * 1e3, "batch": batch.nstruc,
           "atoms": batch.ntotal, "loss_first": (loss0.tolist() if loss0 is not None else None),
           "loss_last": loss.cpu().numpy().tolist()}
    if kernels:
        # per-kernel table of the training step (events around every launch, a
        # separate pass), then the dominant kernel timed alone over the steps
        eng.profile(True)
        for _ in range(max(3, steps // 2)):
            eng.train_step(Eref, Fref, 1.0, 0.1, 1e-3, form, out=out)
        kern = {}
        for k in ("symm_fwd_mom", "symm_bwd_mom", "symm_tan_mom", "symm_fwd_fast", "symm_bwd_fast", "symm_tan_fast",
                  "net_vec", "net_mfma", "net_bwd2", "grad_reduce", "loss", "adam", "force_gather", "struc_sums"):
            ms, n = eng.profile_get(k)
            if n:
                kern[k] = {"avg_us": ms / n * 1e3, "launches": n}
        eng.profile(False)
        dom = max(kern, key=lambda k: kern[k]["avg_us"])
        eng.profile(True)
        eng.profile_only(dom)
        for _ in range(steps):
            eng.train_step(Eref, Fref, 1.0, 0.1, 1e-3, form, out=out)
        ms, n = eng.profile_get(dom)
        eng.profile(False)
        eng.profile_only("")
        nl = eng.nlist()
        work = work_model(model, batch, nl)
        res["kernels"] = kern
        res["dominant"] = {"kernel": dom, "avg_launch_us": ms / n * 1e3 if n else kern[dom]["avg_us"]}
        res["work"] = work
    return res


def run_uneven(steps, warmup):
    """512 uneven triclinic water structures (atom counts U[32, 1024], density
    0.9 g/cm^3, cells strained by a random symmetric matrix): neighbour build +
    E + F of the whole batch per step, FAST and PLAIN."""
    import torch
    from ann_amd import configs
    from ann_amd.gpu import Engine, FAST, PLAIN
    model = configs.model_water()
    batch = configs.batch_uneven(512, seed=13)
    eng = Engine(model)
    eng.set_batch(batch)
    eng.neighbors()
    tuning = eng.autotune(FAST)
    P, maxn = eng.nlist_size()
    dev = eng.device
    out = (torch.empty(batch.nstruc, dtype=torch.float64, device=dev),
           torch.empty((batch.ntotal, 3), dtype=torch.float64, device=dev),
           torch.empty((batch.nstruc, 3, 3), dtype=torch.float64, device=dev))
    res = {"structures": batch.nstruc, "atoms": batch.ntotal, "pairs": P, "tuning": tuning}
    for form, name, k in ((FAST, "fast", steps), (PLAIN, "plain", max(2, steps // 5))):
        dt = time_ef(eng, form, k, max(1, warmup // 2), out)
        res[name] = {"atoms_per_s": batch.ntotal / dt, "ms_per_step": dt * 1e3, "steps": k}
    res["fast_over_plain"] = res["fast"]["atoms_per_s"] / res["plain"]["atoms_per_s"]
    # the moment kernels under grid dispatch and under the persistent (structure,
    # atom-range) work queues, the rest of the tuning as chosen
    res["moment_dispatch"] = {}
    for q, name in ((0, "grid"), (1, "persistent_queues")):
        t = list(tuning) + [0] * (8 - len(tuning))
        t[5] = q
        t[7] = 0                                  # the queues apply to the separate moment kernels
        eng.set_tuning(t)
        dt = time_ef(eng, FAST, steps, max(1, warmup // 2), out)
        res["moment_dispatch"][name] = {"atoms_per_s": batch.ntotal / dt, "ms_per_step": dt * 1e3}
    eng.set_tuning(tuning)
    # correctness at full size: the 4 smallest and 4 largest structures of the
    # batch against the FP64 oracle (FAST form, the timed configuration)
    E, F, _ = eng.energy_forces(FAST, out=out)
    E, F = E.cpu().numpy(), F.cpu().numpy()
    order = np.argsort(batch.natoms, kind="stable")
    pick = np.concatenate([order[:4], order[-4:]])
    res["errors_vs_fp64_oracle"] = oracle_errors(model, batch, pick, E, F)
    # the same FAST step with the E+F kernels replayed from a captured HIP graph
    eng.ef_graph(True)
    dt = time_ef(eng, FAST, steps, max(2, warmup // 2), out)
    eng.ef_graph(False)
    res["fast_graph"] = {"atoms_per_s": batch.ntotal / dt, "ms_per_step": dt * 1e3, "steps": steps}
    return res


def run_md(nsteps, stride, plain_steps, skin=0.5, dt=0.02, reps=16):
    """Velocity-Verlet MD of a reps^3 replica of a 256-atom H/O cell (1,048,576
    atoms for reps = 16, L = 198.8 A): list at rc + skin rebuilt every
    `stride` steps (positions returned to the cell first), E + F + virial every
    step, FP64 energy / kinetic-energy sums.  Time unit A sqrt(amu/eV);
    initial velocities Maxwell-Boltzmann at 300 K."""
    import torch
    from ann_amd import configs
    from ann_amd.gpu import Engine, FAST, PLAIN
    model = configs.model_water()
    b = configs.struc_water_md(seed=15, reps=reps)
    rng = np.random.default_rng(16)
    mass = model.masses[b.species]
    kT = 8.617333262e-5 * 300.0
    vel0 = rng.normal(size=(b.ntotal, 3)) * np.sqrt(kT / mass)[:, None]
    vel0 -= (mass[:, None] * vel0).sum(0) / mass.sum()
    out = {"atoms": b.ntotal, "steps": nsteps, "rebuild_every": stride, "skin": skin, "dt": dt}
    for form, name, k in ((FAST, "fast", nsteps), (PLAIN, "plain", plain_steps)):
        eng = Engine(model)
        eng.set_batch(b)
        dev = eng.device
        m = torch.from_numpy(mass).to(dev)
        v = torch.from_numpy(vel0.copy()).to(dev)
        E = torch.empty(1, dtype=torch.float64, device=dev)
        KE = torch.empty(1, dtype=torch.float64, device=dev)
        V = torch.empty((1, 3, 3), dtype=torch.float64, device=dev)
        eng.neighbors(model.rc + skin)
        _, F, _ = eng.energy_forces(form, out=(E, torch.empty((b.ntotal, 3), dtype=torch.float64, device=dev), V))
        moved_any = False
        e0 = E.item() + 0.5 * float((mass[:, None] * vel0 ** 2).sum())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in range(k):
            if t > 0 and t % stride == 0:
                eng.wrap_positions()
                eng.neighbors(model.rc + skin)
            moved_any |= eng.md_step(dt, m, v, F, form=form, E=E, KE=KE, V=V,
                                     check=(t % stride == stride - 1 or t == k - 1))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        P, maxn = eng.nlist_size()
        etot = (E + KE).item()
        out[name] = {"atom_steps_per_s": b.ntotal * k / wall, "ms_per_step": wall / k * 1e3, "steps": k,
                     "skin_exceeded": bool(moved_any)}
        if name == "fast":
            out.update({"pairs_in_list": P, "max_neighbours": maxn,
                        "energy_per_atom_first": e0 / b.ntotal, "energy_per_atom_last": etot / b.ntotal,
                        "drift_per_atom": (etot - e0) / b.ntotal})
        del eng
        torch.cuda.empty_cache()
    out["fast_over_plain"] = out["fast"]["atom_steps_per_s"] / out["plain"]["atom_steps_per_s"]
    return out


def oracle_errors(model, batch, pick, E, F):
    """Energy (eV/atom) and force (relative RMS) errors of GPU results E, F on
    the structures `pick` of `batch` against the FP64 oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    sub = batch.subset(pick)
    Eo, Fo, _ = pyoracle.Oracle(model).compute_batch(sub, nthreads=os.cpu_count() or 1)
    off = batch.atom_off
    Fg = np.concatenate([F[off[k]:off[k + 1]] for k in pick])
    return {"structures": [int(k) for k in pick], "atoms": [int(batch.natoms[k]) for k in pick],
            "energy_per_atom": float(np.max(np.abs(E[pick] - Eo) / sub.natoms)),
            "force_rel_rms": float(np.sqrt(((Fg - Fo) ** 2).sum() / (Fo ** 2).sum()))}


def run_precond(steps, warmup):
    """The water E+F step with the reference trainer's MEAN / DEV input
    preconditioning (the scaling a trained reference potential carries): its
    cost, the number of inputs evaluated in FP64, and the errors against the
    FP64 oracle under the same scaling on 8 of the 64 structures, with and
    without the FP64 inputs."""
    import torch
    from ann_amd import configs
    from ann_amd.gpu import Engine, FAST
    from ann_amd.precond import apply_preconditioning
    model = configs.model_water()
    batch = configs.batch_water(64, seed=12)
    eng = Engine(model)
    eng.set_batch(batch)
    eng.neighbors()
    pre = apply_preconditioning(eng, "MEAN", "DEV")
    eng.autotune(FAST)
    m2 = model.copy()
    for a, (w, b) in enumerate(pre):
        m2.nets[a].inpw, m2.nets[a].inpb, m2.nets[a].outw, m2.nets[a].outb = w, b, 1.0, 0.0
    dev = eng.device
    out = (torch.empty(batch.nstruc, dtype=torch.float64, device=dev),
           torch.empty((batch.ntotal, 3), dtype=torch.float64, device=dev),
           torch.empty((batch.nstruc, 3, 3), dtype=torch.float64, device=dev))
    dt = time_ef(eng, FAST, steps, warmup, out)
    res = {"fp64_inputs": eng.precise_count(), "inputs": 2 * model.max_ninput,
           "max_abs_inpw_inpb": float(max(np.abs(w * b).max() for w, b in pre)),
           "atoms_per_s": batch.ntotal / dt, "ms_per_step": dt * 1e3}
    eng.profile(True)
    time_ef(eng, FAST, 3, 1, out)
    ms, n = eng.profile_get("precise")
    eng.profile(False)
    res["precise_kernel_us"] = ms / n * 1e3 if n else None
    pick = np.arange(8)
    E, F, _ = eng.energy_forces(FAST, out=out)
    res["errors_vs_fp64_oracle"] = oracle_errors(m2, batch, pick, E.cpu().numpy(), F.cpu().numpy())
    eng.set_precise_ratio(1e300)
    E, F, _ = eng.energy_forces(FAST, out=out)
    res["errors_fp32_inputs_only"] = oracle_errors(m2, batch, pick, E.cpu().numpy(), F.cpu().numpy())
    return res


def train_gradient_check(nstruc=2):
    """The training gradient of 2 of the 128 Li/P/S structures (GPU step
    without the update) against the oracle's analytic gradient."""
    import torch
    from ann_amd.gpu import Engine, FAST
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    model, batch, Eref, Fref = train_setup()
    sub = batch.subset(range(nstruc))
    na = sub.ntotal
    Er, Fr = Eref[:nstruc].clone(), Fref[:na].clone()
    eng = Engine(model)
    eng.set_batch(sub)
    eng.neighbors()
    loss, g = eng.train_step(Er, Fr, 1.0, 0.1, 0.0, FAST)
    _, go, _, _, parts = pyoracle.Oracle(model).train_loss_grad(sub, Er.cpu().numpy(), Fr.cpu().numpy(), 1.0, 0.1)
    return {"structures": nstruc, "atoms": na,
            "gradient_rel_err": float(np.linalg.norm(g.cpu().numpy() - go) / np.linalg.norm(go)),
            "loss_rel_err": float(np.max(np.abs(loss.cpu().numpy() - parts) / np.abs(parts)))}


def train_cpu_baseline():
    """The FP64 oracle's training step (E + F, loss, force adjoint, second
    backward; the port of the reference algorithm) on all 128 Li/P/S
    structures, structures split over every host core (one call per thread;
    the gradient sums are the step's, the Adam update is negligible)."""
    import concurrent.futures as cf
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    model, batch, Eref, Fref = train_setup()
    Er, Fr = Eref.cpu().numpy(), Fref.cpu().numpy()
    cores = os.cpu_count() or 1
    parts = [list(range(k, batch.nstruc, cores)) for k in range(cores)]
    parts = [p for p in parts if p]
    orc = pyoracle.Oracle(model)
    off = batch.atom_off

    def one(idx):
        sub = batch.subset(idx)
        Fs = np.concatenate([Fr[off[k]:off[k + 1]] for k in idx])
        return orc.train_loss_grad(sub, Er[idx], Fs, 1.0, 0.1)[1] * len(idx) / batch.nstruc

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=len(parts)) as ex:
        sum(ex.map(one, parts))
    dt = time.perf_counter() - t0
    return {"value": batch.nstruc / dt, "unit": "structs/s", "cores": len(parts), "kind": "port",
            "sample": f"all 128 Li/P/S structures ({batch.ntotal} atoms), one training step (E + F, loss, "
                      f"force adjoint, second backward) in FP64, {len(parts)} threads, {dt:.2f} s wall"}


def run_rotating(steps, warmup, nset=512, size=128):
    """Training with a new batch every step: the reference's sampler draws 128
    of 512 Li/P/S structures resident in HBM, the positions are gathered on the
    GPU (offsets, cells and species from the host copies), and the step
    includes ann_set_batch_host, the neighbour build, G and the training step
    (FAST, Adam)."""
    import torch
    from ann_amd.batch import BatchSampler, ResidentSet
    from ann_amd.gpu import Engine, FAST
    model, full, Eref, Fref = train_setup(nset, seed=35)
    rs = ResidentSet(full, Eref.cpu().numpy(), Fref.cpu().numpy())
    sampler = BatchSampler(size, nset, seed=7)
    eng = Engine(model)
    out = (torch.empty(2, dtype=torch.float64, device=eng.device),
           torch.empty(eng.nparams, dtype=torch.float64, device=eng.device))

    def step():
        b = rs.gather(sampler.next(), device_meta=False)
        eng.set_batch_host(b["atom_off"], b["pos"], b["cells_host"], b["species_host"])
        eng.neighbors()
        eng.train_step(b["energy"], b["forces"], 1.0, 0.1, 1e-3, FAST, out=out)
        return b

    for _ in range(max(warmup, 2)):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nstr = 0
    for _ in range(steps):
        nstr += len(step()["atom_off"]) - 1
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"structs_per_s": nstr / dt, "ms_per_step": dt / steps * 1e3, "steps": steps, "batch": size,
            "training_set": nset, "loss_last": out[0].cpu().numpy().tolist()}


def cpu_baseline(model, batch, gpu_forces):
    """FP64 oracle (the port of the reference algorithm) on a bounded sample:
    the 64 structures of the batch (about 0.4 s of CPU work each, 25-30 s in
    total), spread over every host core.  Also returns the force error of each
    GPU form on that sample (the oracle as checker)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from ann_amd.configs import Batch
    cores = os.cpu_count() or 1
    nsample = batch.nstruc
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
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-md", action="store_true")
    ap.add_argument("--md-steps", type=int, default=100)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy three times (calibrates HBM counters)")
    ap.add_argument("--ef-loop", type=int, default=0, metavar="N",
                    help="only run N back-to-back E+F steps (neighbour build + energy_forces, FAST, after the "
                         "autotuner and five untimed steps) on water 64 x 192, for a kernel-trace timeline")
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
        print("bench.py measures one GPU: the library has no multi-GPU path (batches shard by structure, "
              "each GPU running the path on its share). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch
    from ann_amd import configs
    from ann_amd.gpu import Engine, FAST, PLAIN, TRIPLET

    model = configs.model_water()
    batch = configs.batch_water(64, seed=12)
    eng = Engine(model)
    eng.set_batch(batch)
    eng.neighbors()
    tuning = eng.autotune(FAST)          # atoms per workgroup, kernel family, search blocking
    if args.ef_loop > 0:
        dev = eng.device
        o = (torch.empty(batch.nstruc, dtype=torch.float64, device=dev),
             torch.empty((batch.ntotal, 3), dtype=torch.float64, device=dev),
             torch.empty((batch.nstruc, 3, 3), dtype=torch.float64, device=dev))
        for _ in range(5):
            eng.neighbors()
            eng.energy_forces(FAST, out=o)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.ef_loop):
            eng.neighbors()
            eng.energy_forces(FAST, out=o)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.ef_loop
        print(json.dumps({"ef_loop_steps": args.ef_loop, "ms_per_step": dt * 1e3, "tuning": tuning}))
        return
    nl = eng.nlist()
    flops = work_model(model, batch, nl)
    dev = eng.device
    out = (torch.empty(batch.nstruc, dtype=torch.float64, device=dev),
           torch.empty((batch.ntotal, 3), dtype=torch.float64, device=dev),
           torch.empty((batch.nstruc, 3, 3), dtype=torch.float64, device=dev))
    res = {}
    forces = {}
    live = {}
    for form, name in ((PLAIN, "plain"), (TRIPLET, "triplet"), (FAST, "fast")):
        # kernel table: a pass of the same steps with HIP events around every launch
        eng.profile(True)
        time_ef(eng, form, max(3, args.steps // 2), args.warmup, out)
        kern = {}
        for k in ("nb_bin", "nb_search", "nb_compact", "nb_rev", "symm_fwd_plain", "symm_fwd_fast", "symm_fwd_mom",
                  "symm_bwd_plain", "symm_bwd_fast", "symm_bwd_mom", "net_plain", "net_vec", "net_mfma",
                  "ef_fused", "force_gather", "struc_sums"):
            ms, n = eng.profile_get(k)
            if n:
                kern[k] = {"avg_us": ms / n * 1e3, "launches": n}
        eng.profile(False)
        # timed region: events only around the dominant kernel (events around every launch
        # add a few microseconds per launch to the step)
        roofs_k = kernel_rooflines(kern, flops)
        dom_k = max((k for k in kern if k in roofs_k), key=lambda k: kern[k]["avg_us"])
        eng.profile(True)
        eng.profile_only(dom_k)
        dt = time_ef(eng, form, args.steps, args.warmup, out)
        ms, n = eng.profile_get(dom_k)
        eng.profile(False)
        eng.profile_only("")
        live[name] = (dom_k, ms / n * 1e3 if n else kern[dom_k]["avg_us"])
        res[name] = {"ms_per_step": dt * 1e3, "atoms_per_s": batch.ntotal / dt, "kernels": kern}
        E, F, _ = eng.energy_forces(form)
        forces[name] = (E.cpu().numpy(), F.cpu().numpy())
    # the E+F step with the fused kernel against the separate G / network /
    # adjoint kernels (the autotuner keeps the faster)
    ef_kernels = {}
    for q, key in ((0, "separate"), (1, "fused")):
        t = list(tuning) + [0] * (8 - len(tuning))
        t[7] = q
        eng.set_tuning(t)
        ef_kernels[key] = {"ms_per_step": time_ef(eng, FAST, max(3, args.steps // 2), 2, out) * 1e3}
    # the two network kernels on the same E+F steps (separate kernels): vector-ALU
    # tiles against matrix-core tiles (the autotuner keeps the faster)
    net_kernels = {}
    for q, key in ((0, "net_vec"), (1, "net_mfma")):
        t = list(tuning) + [0] * (8 - len(tuning))
        t[6] = q
        t[7] = 0
        eng.set_tuning(t)
        eng.profile(True)
        eng.profile_only(key)
        time_ef(eng, FAST, 5, 1, out)
        ms, n = eng.profile_get(key)
        eng.profile(False)
        eng.profile_only("")
        if n:
            net_kernels[key] = {"avg_us": ms / n * 1e3, "launches": n}
    eng.set_tuning(tuning)
    # dominant kernel of the FAST form, its duration measured inside the timed region
    dom, dom_us = live["fast"]
    roofs = {name: kernel_rooflines(res[name]["kernels"], flops) for name in res}
    rd = kernel_rooflines({dom: {"avg_us": dom_us}}, flops)[dom]
    roof = {"bound": rd["bound"], "kernel": dom, "achieved": rd["achieved"], "peak": rd["peak"], "unit": rd["unit"],
            "frac": rd["frac"], "traffic": None, "algorithmic_work_per_launch": rd["work"],
            "avg_launch_us": dom_us}
    if rd["bound"] == "valu":
        roof["note"] = ("FP32 vector ALU; peak is the MI355X FP32 figure (the kernels are compiled without packed FP32, "
                        "so the attainable non-packed VALU rate is half of it)")
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        tj = json.load(open(traffic_file))
        if dom in tj:
            roof["traffic"] = tj[dom]["bytes_per_launch"]
    train = None
    if not args.no_train:
        train = run_train(max(3, args.steps // 2), max(1, args.warmup // 2), FAST, kernels=True)
        tw = train.pop("work")
        tk = train["dominant"]["kernel"]
        tr = kernel_rooflines({tk: {"avg_us": train["dominant"]["avg_launch_us"]}}, tw).get(tk)
        if tr:
            train["roofline"] = {"bound": tr["bound"], "kernel": tk, "achieved": tr["achieved"], "peak": tr["peak"],
                                 "unit": tr["unit"], "frac": tr["frac"], "algorithmic_work_per_launch": tr["work"],
                                 "avg_launch_us": train["dominant"]["avg_launch_us"], "traffic": None}
            if os.path.exists(traffic_file):
                tj = json.load(open(traffic_file))
                if tk in tj and "train" in tj[tk].get("workload", ""):
                    train["roofline"]["traffic"] = tj[tk]["bytes_per_launch"]
        train["kernel_rooflines"] = kernel_rooflines(train["kernels"], tw)
        train["graph_structs_per_s"] = run_train(max(3, args.steps // 2), 1, FAST, graph=True)["structs_per_s"]
        train_plain = run_train(max(3, args.steps // 4), 1, PLAIN)
        train["plain_structs_per_s"] = train_plain["structs_per_s"]
        train["triplet_structs_per_s"] = run_train(max(3, args.steps // 2), 1, TRIPLET)["structs_per_s"]
        train["fast_over_plain"] = train["structs_per_s"] / train["plain_structs_per_s"]
        train["gradient_check"] = train_gradient_check()
        train["rotating_batch"] = run_rotating(max(3, args.steps // 2), max(1, args.warmup // 2))
        if not args.no_cpu:
            train["cpu_baseline"] = train_cpu_baseline()
    def note(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    note(f"E+F fast {res['fast']['atoms_per_s']:.4g} atoms/s, plain {res['plain']['atoms_per_s']:.4g}")
    if train:
        note(f"train {train['structs_per_s']:.4g} structs/s")
    uneven = None if args.no_uneven else run_uneven(max(3, args.steps // 2), args.warmup)
    if uneven:
        note(f"uneven batch fast {uneven['fast']['atoms_per_s']:.4g} atoms/s")
    precond = run_precond(max(3, args.steps // 2), max(1, args.warmup // 2))
    note(f"preconditioned E+F {precond['atoms_per_s']:.4g} atoms/s, {precond['fp64_inputs']} FP64 inputs")
    md = None if args.no_md else run_md(args.md_steps, 10, 2)
    if md:
        note(f"md fast {md['fast']['atom_steps_per_s']:.4g} atom-steps/s")
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
        "triplet": {"atoms_per_s": res["triplet"]["atoms_per_s"], "ms_per_step": res["triplet"]["ms_per_step"]},
        "fast_over_plain": res["fast"]["atoms_per_s"] / res["plain"]["atoms_per_s"],
        "kernels": {name: res[name]["kernels"] for name in res},
        "roofline": roof,
        "kernel_rooflines": roofs,
        "tuning": {"atoms_per_workgroup": {"G": tuning[0], "adjoint": tuning[1], "tangent": tuning[2]},
                   "angular": "triplet" if tuning[3] else "moment", "search_blocks": "single" if tuning[4] else "2x2x2",
                   "moment_dispatch": "persistent queues" if len(tuning) > 5 and tuning[5] else "grid",
                   "network": "matrix cores" if len(tuning) > 6 and tuning[6] else "vector ALU",
                   "ef_kernels": "fused" if len(tuning) > 7 and tuning[7] else "separate"},
        "ef_kernels": ef_kernels,
        "network_kernels": net_kernels,
        "train": train,
        "uneven_batch": uneven,
        "md": md,
        "cpu_baseline": cpu,
        "errors_vs_fp64_oracle": errs,
        "preconditioned": precond,
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
