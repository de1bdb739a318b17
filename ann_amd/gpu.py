"""Python binding of the C ABI (include/ann_gpu.h) with torch for device memory.

The binding is thin: torch allocates HBM and provides the stream, every
computation runs in libann_amd.so.  There is no CPU fallback; constructing an
``Engine`` without the built library or without a GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np
import torch

from .configs import Batch
from .model import Model

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libann_amd.so")
_LIB = None

PLAIN, FAST, TRIPLET = 0, 1, 2   # computation forms (include/ann_gpu.h)
P_D = C.POINTER(C.c_double)
P_I = C.POINTER(C.c_int)

ERRORS = {-1: "ANN_ERR_ARG", -2: "ANN_ERR_UNSUPPORTED", -3: "ANN_ERR_CAPACITY", -4: "ANN_ERR_HIP",
          -5: "ANN_ERR_STATE"}


class AnnError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class ModelDesc(C.Structure):
    _fields_ = [("nspecies", C.c_int), ("rc", C.c_double), ("cutoff", C.c_int),
                ("rad_name", P_I), ("rad_n", P_I), ("rad_off", P_I), ("rad_rs", P_D), ("rad_eta", P_D),
                ("ang_name", P_I), ("ang_n", P_I), ("ang_off", P_I), ("ang_eta", P_D),
                ("ang_zeta", P_D), ("ang_lambda", P_I),
                ("nn_nlayer", P_I), ("nn_nodes_off", P_I), ("nn_nodes", P_I), ("nn_neuron", P_I),
                ("nn_c", P_D), ("nn_par_off", P_I), ("params", P_D),
                ("nn_inp_off", P_I), ("inpw", P_D), ("inpb", P_D),
                ("outw", P_D), ("outb", P_D), ("evac", P_D)]


class BatchDesc(C.Structure):
    _fields_ = [("nstruc", C.c_int), ("atom_off", P_I), ("d_atom_off", C.c_void_p),
                ("d_pos", C.c_void_p), ("d_cells", C.c_void_p), ("d_species", C.c_void_p)]


EXPORTS = ["ann_last_error", "ann_create", "ann_destroy", "ann_set_stream", "ann_info", "ann_set_params",
           "ann_set_scaling", "ann_set_batch", "ann_set_batch_host", "ann_neighbors", "ann_nlist_size", "ann_nlist_copy",
           "ann_symm_fwd", "ann_net_fwd", "ann_symm_bwd", "ann_force_reduce", "ann_energy_forces",
           "ann_symm_tangent", "ann_net_bwd2", "ann_train_step", "ann_get_params", "ann_profile",
           "ann_profile_get", "ann_profile_only", "ann_refresh_geometry", "ann_wrap_positions", "ann_md_step", "ann_train_graph", "ann_ef_graph",
           "ann_autotune", "ann_get_tuning", "ann_set_tuning", "ann_input_stats",
           "ann_get_optimizer_state", "ann_set_optimizer_state", "ann_symm_fwd64", "ann_net_fwd64",
           "ann_set_precise_ratio", "ann_precise_count", "ann_force_reduce_noalloc", "ann_set_loss",
           "ann_set_struc_weights", "ann_set_neighbors", "ann_cutoff_eval", "ann_radial_eval",
           "ann_angular_eval", "ann_mlp_create", "ann_mlp_set_params", "ann_mlp_set_scaling", "ann_mlp_size",
           "ann_mlp_eval", "ann_mlp_grad", "ann_mlp_destroy", "ann_pair_compute"]


def lib():
    """Load libann_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m ann_amd.build` first")
        L = C.CDLL(LIB_PATH)
        L.ann_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def _check(rc: int) -> None:
    if rc != 0:
        raise AnnError(rc, lib().ann_last_error().decode())


def _ptr(t: Optional[torch.Tensor]):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


class Engine:
    """One context of the GPU path for a potential ``model`` on ``device``."""

    def __init__(self, model: Model, device: str = "cuda:0"):
        if not torch.cuda.is_available():
            raise RuntimeError("ann_amd.gpu.Engine needs a GPU")
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        self.model = model
        self.flat = model.flat()
        f = self.flat
        self.desc = ModelDesc(model.nspecies, float(model.rc), int(model.cutoff),
                              *[f[k].ctypes.data_as(P_I if f[k].dtype == np.int32 else P_D) for k in (
                                  "rad_name", "rad_n", "rad_off", "rad_rs", "rad_eta",
                                  "ang_name", "ang_n", "ang_off", "ang_eta", "ang_zeta", "ang_lambda",
                                  "nn_nlayer", "nn_nodes_off", "nn_nodes", "nn_neuron", "nn_c",
                                  "nn_par_off", "params", "nn_inp_off", "inpw", "inpb", "outw", "outb",
                                  "evac")])
        h = C.c_void_p()
        _check(lib().ann_create(C.byref(self.desc), C.byref(h)))
        self.h = h
        self.stream = torch.cuda.current_stream(self.device)
        _check(lib().ann_set_stream(self.h, C.c_void_p(self.stream.cuda_stream)))
        S, nin, ldg, npar = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(lib().ann_info(self.h, C.byref(S), C.byref(nin), C.byref(ldg), C.byref(npar)))
        self.ldg, self.max_nin, self.nparams = ldg.value, nin.value, npar.value
        self.nspecies = S.value
        self._scaling = {}
        self.batch: Optional[Batch] = None

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().ann_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # ------------------------------------------------------------ set up
    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float64)
        _check(lib().ann_set_params(self.h, p.ctypes.data_as(P_D), len(p)))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.nparams)
        _check(lib().ann_get_params(self.h, p.ctypes.data_as(P_D), self.nparams))
        return p

    def input_stats(self, G: Optional[torch.Tensor] = None):
        """Per-species input statistics of the current batch (ann_input_stats).

        Returns (count (S,), stats (S, 6, max_ninput)) with stats[:, q] = min, max,
        harmonic mean, mean, deviation, rms; G defaults to the FAST G of the batch.
        """
        if G is None:
            G = self.symm(FAST)
        G = G.contiguous()
        torch.cuda.synchronize(self.device)
        cnt = np.zeros(self.nspecies, dtype=np.int32)
        st = np.zeros((self.nspecies, 6, self.max_nin))
        _check(lib().ann_input_stats(self.h, _ptr(G), cnt.ctypes.data_as(P_I), st.ctypes.data_as(P_D), st.size))
        return cnt, st

    def get_optimizer_state(self):
        """(first moments, second moments, steps taken) of Adam."""
        m, v, t = np.zeros(self.nparams), np.zeros(self.nparams), C.c_int()
        _check(lib().ann_get_optimizer_state(self.h, m.ctypes.data_as(P_D), v.ctypes.data_as(P_D), C.byref(t),
                                             self.nparams))
        return m, v, t.value

    def set_optimizer_state(self, m, v, step: int) -> None:
        m = np.ascontiguousarray(m, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        _check(lib().ann_set_optimizer_state(self.h, m.ctypes.data_as(P_D), v.ctypes.data_as(P_D), int(step),
                                             len(m)))

    def scaling(self, species: int):
        """(inpw, inpb, outw, outb) last set for a species, else the model's own."""
        if species in self._scaling:
            return self._scaling[species]
        n = self.model.nets[species]
        return (np.asarray(n.inpw, dtype=np.float64), np.asarray(n.inpb, dtype=np.float64), float(n.outw),
                float(n.outb))

    def set_scaling(self, species: int, inpw, inpb, outw: float = 1.0, outb: float = 0.0) -> None:
        w = np.ascontiguousarray(inpw, dtype=np.float64)
        b = np.ascontiguousarray(inpb, dtype=np.float64)
        self._scaling[species] = (w.copy(), b.copy(), float(outw), float(outb))
        _check(lib().ann_set_scaling(self.h, species, w.ctypes.data_as(P_D), b.ctypes.data_as(P_D),
                                     C.c_double(outw), C.c_double(outb)))

    def set_batch(self, batch: Batch) -> None:
        dev = self.device
        self.batch = batch
        self._off_h = np.ascontiguousarray(batch.atom_off, dtype=np.int32)
        self.d_off = torch.from_numpy(self._off_h).to(dev)
        self.d_pos = torch.from_numpy(np.ascontiguousarray(batch.pos, dtype=np.float64)).to(dev)
        self.d_cells = torch.from_numpy(np.ascontiguousarray(batch.cells, dtype=np.float64).reshape(-1, 9)).to(dev)
        self.d_species = torch.from_numpy(np.ascontiguousarray(batch.species, dtype=np.int32)).to(dev)
        self.N = batch.ntotal
        self.S = batch.nstruc
        self._bdesc = BatchDesc(self.S, self._off_h.ctypes.data_as(P_I), _ptr(self.d_off), _ptr(self.d_pos),
                                _ptr(self.d_cells), _ptr(self.d_species))
        _check(lib().ann_set_batch(self.h, C.byref(self._bdesc)))

    def set_batch_tensors(self, atom_off: np.ndarray, pos: torch.Tensor, cells: torch.Tensor,
                          species: torch.Tensor) -> None:
        """ann_set_batch on arrays already in device memory (positions [N, 3] FP64,
        cells [S, 9] FP64, species [N] int32; host atom offsets [S + 1])."""
        self.batch = None
        self._off_h = np.ascontiguousarray(atom_off, dtype=np.int32)
        self.d_off = torch.from_numpy(self._off_h).to(self.device, non_blocking=True)
        self.d_pos = pos.contiguous()
        self.d_cells = cells.contiguous()
        self.d_species = species.to(torch.int32).contiguous()
        self.N = int(self._off_h[-1])
        self.S = len(self._off_h) - 1
        self._bdesc = BatchDesc(self.S, self._off_h.ctypes.data_as(P_I), _ptr(self.d_off), _ptr(self.d_pos),
                                _ptr(self.d_cells), _ptr(self.d_species))
        _check(lib().ann_set_batch(self.h, C.byref(self._bdesc)))

    def set_batch_host(self, atom_off: np.ndarray, pos: torch.Tensor, cells: np.ndarray, species: np.ndarray) -> None:
        """ann_set_batch_host: positions [N, 3] FP64 in device memory; atom
        offsets, cells and species from host arrays (copied by the library
        without a host wait, so the set-up queues behind earlier GPU work)."""
        self.batch = None
        self._off_h = np.ascontiguousarray(atom_off, dtype=np.int32)
        cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 9)
        species = np.ascontiguousarray(species, dtype=np.int32)
        self.d_pos = pos.contiguous()
        self.d_off = self.d_cells = self.d_species = None
        self.N = int(self._off_h[-1])
        self.S = len(self._off_h) - 1
        _check(lib().ann_set_batch_host(self.h, self.S, self._off_h.ctypes.data_as(P_I), _ptr(self.d_pos),
                                        cells.ctypes.data_as(P_D), species.ctypes.data_as(P_I)))

    def neighbors(self, rc: float = 0.0) -> None:
        _check(lib().ann_neighbors(self.h, C.c_double(rc)))

    def set_neighbors(self, off, idx, img, rc: float) -> None:
        """A neighbour list of the caller's (ann_set_neighbors): host CSR."""
        off = np.ascontiguousarray(off, dtype=np.int32)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        img = np.ascontiguousarray(img, dtype=np.int32).reshape(-1)
        _check(lib().ann_set_neighbors(self.h, off.ctypes.data_as(P_I), idx.ctypes.data_as(P_I),
                                       img.ctypes.data_as(P_I), C.c_double(rc)))

    def nlist_size(self):
        P, m = C.c_int64(), C.c_int()
        _check(lib().ann_nlist_size(self.h, C.byref(P), C.byref(m)))
        return P.value, m.value

    def nlist(self) -> dict:
        P, _ = self.nlist_size()
        dev = self.device
        off = torch.empty(self.N + 1, dtype=torch.int32, device=dev)
        idx = torch.empty(max(P, 1), dtype=torch.int32, device=dev)
        img = torch.empty((max(P, 1), 3), dtype=torch.int32, device=dev)
        rev = torch.empty(max(P, 1), dtype=torch.int32, device=dev)
        geo = torch.empty((max(P, 1), 4), dtype=torch.float32, device=dev)
        _check(lib().ann_nlist_copy(self.h, _ptr(off), _ptr(idx), _ptr(img), _ptr(rev), _ptr(geo)))
        torch.cuda.synchronize(dev)
        return {"off": off.cpu().numpy(), "idx": idx[:P].cpu().numpy(), "img": img[:P].cpu().numpy(),
                "rev": rev[:P].cpu().numpy(), "geo": geo[:P].cpu().numpy()}

    # ------------------------------------------------------------ forward / backward pairs
    def symm(self, form: int = FAST) -> torch.Tensor:
        G = torch.empty((self.N, self.ldg), dtype=torch.float32, device=self.device)
        _check(lib().ann_symm_fwd(self.h, form, _ptr(G)))
        return G

    def symm64(self, form: int = FAST) -> torch.Tensor:
        """G in FP64 (ann_symm_fwd64): FP64-evaluated inputs exact to FP64, others widened."""
        G = torch.empty((self.N, self.ldg), dtype=torch.float64, device=self.device)
        _check(lib().ann_symm_fwd64(self.h, form, _ptr(G)))
        return G

    def net64(self, G: torch.Tensor, form: int = FAST, grad: bool = True):
        """Network on FP64 rows (ann_net_fwd64)."""
        E = torch.empty(self.N, dtype=torch.float32, device=self.device)
        dE = torch.empty((self.N, self.ldg), dtype=torch.float32, device=self.device) if grad else None
        _check(lib().ann_net_fwd64(self.h, form, _ptr(G.to(torch.float64).contiguous()), _ptr(E), _ptr(dE)))
        return E, dE

    def set_precise_ratio(self, ratio: float) -> int:
        """Inputs with |inpw inpb| > ratio (all if ratio < 0) are evaluated in FP64;
        returns how many inputs that flags."""
        _check(lib().ann_set_precise_ratio(self.h, C.c_double(ratio)))
        return self.precise_count()

    def precise_count(self) -> int:
        n = C.c_int()
        _check(lib().ann_precise_count(self.h, C.byref(n)))
        return n.value

    def net(self, G: torch.Tensor, form: int = FAST, grad: bool = True):
        E = torch.empty(self.N, dtype=torch.float32, device=self.device)
        dE = torch.empty((self.N, self.ldg), dtype=torch.float32, device=self.device) if grad else None
        _check(lib().ann_net_fwd(self.h, form, _ptr(G.contiguous()), _ptr(E), _ptr(dE)))
        return E, dE

    def symm_bwd(self, dEdG: torch.Tensor, form: int = FAST) -> torch.Tensor:
        P, _ = self.nlist_size()
        out = torch.empty((max(P, 1), 3), dtype=torch.float32, device=self.device)
        _check(lib().ann_symm_bwd(self.h, form, _ptr(dEdG.contiguous()), _ptr(out)))
        return out[:P]

    def force_reduce(self, P: torch.Tensor, Eatom: torch.Tensor, with_evac: bool = True):
        F = torch.empty((self.N, 3), dtype=torch.float64, device=self.device)
        E = torch.empty(self.S, dtype=torch.float64, device=self.device)
        V = torch.empty((self.S, 3, 3), dtype=torch.float64, device=self.device)
        _check(lib().ann_force_reduce(self.h, _ptr(P.contiguous()), _ptr(Eatom), int(with_evac), _ptr(F),
                                      _ptr(E), _ptr(V)))
        return E, F, V

    def force_reduce_noalloc(self, P: torch.Tensor, Eatom: torch.Tensor, with_evac: bool = True, out=None,
                             virial: bool = True):
        """ann_force_reduce_noalloc into caller-owned buffers ``out`` = (E, F, V)."""
        if out is None:
            out = (torch.empty(self.S, dtype=torch.float64, device=self.device),
                   torch.empty((self.N, 3), dtype=torch.float64, device=self.device),
                   torch.empty((self.S, 3, 3), dtype=torch.float64, device=self.device))
        E, F, V = out
        _check(lib().ann_force_reduce_noalloc(self.h, _ptr(P.contiguous()), _ptr(Eatom), int(with_evac), _ptr(F),
                                              _ptr(E), _ptr(V if virial else None)))
        return E, F, V

    def energy_forces(self, form: int = FAST, with_evac: bool = True, out=None):
        if out is None:
            out = (torch.empty(self.S, dtype=torch.float64, device=self.device),
                   torch.empty((self.N, 3), dtype=torch.float64, device=self.device),
                   torch.empty((self.S, 3, 3), dtype=torch.float64, device=self.device))
        E, F, V = out
        _check(lib().ann_energy_forces(self.h, form, int(with_evac), _ptr(E), _ptr(F), _ptr(V)))
        return E, F, V

    def energies(self, form: int = FAST, with_evac: bool = True, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Structure energies only (ann_energy_forces without force or virial
        outputs: G, the networks and the per-structure sums, no adjoint)."""
        E = out if out is not None else torch.empty(self.S, dtype=torch.float64, device=self.device)
        _check(lib().ann_energy_forces(self.h, form, int(with_evac), _ptr(E), None, None))
        return E

    def tangent(self, u: torch.Tensor, form: int = FAST) -> torch.Tensor:
        v = torch.empty((self.N, self.ldg), dtype=torch.float32, device=self.device)
        _check(lib().ann_symm_tangent(self.h, form, _ptr(u.to(torch.float32).contiguous()), _ptr(v)))
        return v

    def net_bwd2(self, G: torch.Tensor, v: Optional[torch.Tensor], gE: torch.Tensor) -> torch.Tensor:
        g = torch.empty(self.nparams, dtype=torch.float64, device=self.device)
        _check(lib().ann_net_bwd2(self.h, _ptr(G.contiguous()), _ptr(v), _ptr(gE.to(torch.float64).contiguous()),
                                  _ptr(g)))
        return g

    def train_step(self, Eref: torch.Tensor, Fref: torch.Tensor, wE: float = 1.0, wF: float = 0.1,
                   gamma: float = 0.0, form: int = FAST, out=None):
        """One training step; ``out`` = (loss[2], grad[nparams]) FP64 buffers to reuse
        (a captured step is replayed only while its buffers stay the same)."""
        if out is None:
            out = (torch.empty(2, dtype=torch.float64, device=self.device),
                   torch.empty(self.nparams, dtype=torch.float64, device=self.device))
        loss, g = out
        _check(lib().ann_train_step(self.h, form, _ptr(Eref.contiguous()), _ptr(Fref.contiguous()),
                                    C.c_double(wE), C.c_double(wF), C.c_double(gamma), _ptr(loss), _ptr(g)))
        return loss, g

    LOSS = {"MSE": 1, "MAE": 2, "HUBER": 3, "ASINH": 4}
    NORM = {"IDENTITY": 0, "LINEAR": 1, "SQRT": 2, "CBRT": 3, "LOG": 4}

    def set_loss(self, loss: str = "MSE", delta: float = 1.0, norm: str = "LINEAR") -> None:
        """Energy loss of the training step (ann_set_loss)."""
        _check(lib().ann_set_loss(self.h, self.LOSS[loss.upper()], C.c_double(delta), self.NORM[norm.upper()]))

    def set_struc_weights(self, wE: Optional[torch.Tensor] = None, wF: Optional[torch.Tensor] = None) -> None:
        """Per-structure energy / force weights (FP64 device tensors [nstruc], kept alive here)."""
        self._wts = (None if wE is None else wE.to(torch.float64).contiguous(),
                     None if wF is None else wF.to(torch.float64).contiguous())
        _check(lib().ann_set_struc_weights(self.h, _ptr(self._wts[0]), _ptr(self._wts[1])))

    def train_graph(self, enable: bool = True) -> None:
        """Capture the training step as a HIP graph after its first call and replay it."""
        _check(lib().ann_train_graph(self.h, int(enable)))

    def ef_graph(self, enable: bool = True) -> None:
        """Capture energy_forces as a HIP graph after its first call and replay it
        while its buffers, the list's arrays and the tuning stay the same."""
        _check(lib().ann_ef_graph(self.h, int(enable)))

    # ------------------------------------------------------------ molecular dynamics
    def refresh_geometry(self) -> None:
        """Pair geometry from the current positions (self.d_pos) and the list's images."""
        _check(lib().ann_refresh_geometry(self.h))

    def wrap_positions(self) -> None:
        """Return self.d_pos to the cell in place; the list must be rebuilt afterwards."""
        _check(lib().ann_wrap_positions(self.h))

    def md_step(self, dt: float, mass: torch.Tensor, vel: torch.Tensor, F: torch.Tensor, form: int = FAST,
                E: Optional[torch.Tensor] = None, KE: Optional[torch.Tensor] = None,
                V: Optional[torch.Tensor] = None, check: bool = False) -> bool:
        """One velocity-Verlet step; updates self.d_pos, vel and F in place.
        Returns True when an atom moved more than skin/2 since the last build
        (only if check is set; that synchronises)."""
        moved = C.c_int(0)
        _check(lib().ann_md_step(self.h, form, C.c_double(dt), _ptr(mass), _ptr(vel), _ptr(F), _ptr(E), _ptr(KE),
                                 _ptr(V), C.byref(moved) if check else None))
        return bool(moved.value)

    # ------------------------------------------------------------ LAMMPS-style pair interface
    def pair_compute(self, x, species, cell, owner, neigh, rc: float, nlocal: int, new_list: bool = True,
                     form: int = FAST, per_atom: bool = True):
        """ann_pair_compute: x [nall][3] (locals first, then ghosts), species
        [nlocal], owner [nghost], neigh = list of nlocal integer arrays (the
        full list, indices into [0, nall)).  Returns (E, F [nlocal][3],
        eatom [nlocal], virial[6])."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        cell = np.ascontiguousarray(cell, dtype=np.float64)
        own = np.ascontiguousarray(owner, dtype=np.int32)
        nghost = len(x) - nlocal
        rows = [np.ascontiguousarray(r, dtype=np.int32) for r in neigh] if neigh is not None else None
        num = np.array([len(r) for r in rows], np.int32) if rows is not None else None
        ptrs = (P_I * nlocal)(*[r.ctypes.data_as(P_I) for r in rows]) if rows is not None else None
        f = np.zeros((nlocal, 3))
        e = C.c_double(0.0)
        ea = np.zeros(nlocal) if per_atom else None
        vir = np.zeros(6)
        _check(lib().ann_pair_compute(self.h, form, int(new_list), nlocal, nghost, x.ctypes.data_as(P_D),
                                      sp.ctypes.data_as(P_I), cell.ctypes.data_as(P_D), own.ctypes.data_as(P_I),
                                      num.ctypes.data_as(P_I) if num is not None else None, ptrs,
                                      C.c_double(rc), f.ctypes.data_as(P_D), C.byref(e),
                                      ea.ctypes.data_as(P_D) if ea is not None else None, vir.ctypes.data_as(P_D)))
        return e.value, f, ea, vir

    # ------------------------------------------------------------ tuning
    def autotune(self, form: int = FAST) -> list:
        _check(lib().ann_autotune(self.h, form))
        return self.tuning()

    def tuning(self) -> list:
        """[atoms per workgroup G, adjoint, tangent; triplet tiles; single-bin search; persistent queues;
        matrix-core network; fused E+F kernel]"""
        t = (C.c_int * 8)()
        _check(lib().ann_get_tuning(self.h, t, 8))
        return list(t)

    def set_tuning(self, t) -> None:
        v = (C.c_int * len(t))(*[int(x) for x in t])
        _check(lib().ann_set_tuning(self.h, v, len(t)))

    # ------------------------------------------------------------ timing
    def profile(self, enable: bool = True) -> None:
        _check(lib().ann_profile(self.h, int(enable)))

    def profile_only(self, name: str = "") -> None:
        """Time only the kernel family `name` ("" for every family)."""
        _check(lib().ann_profile_only(self.h, name.encode()))

    def profile_get(self, name: str):
        ms, n = C.c_double(), C.c_int()
        _check(lib().ann_profile_get(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value


# ---------------------------------------------------------------- FP64 evaluators
class RadialDesc(C.Structure):
    _fields_ = [("cutoff", C.c_int), ("rc", C.c_double), ("name", C.c_int), ("n", C.c_int),
                ("rs", P_D), ("eta", P_D)]


class AngularDesc(C.Structure):
    _fields_ = [("cutoff", C.c_int), ("rc", C.c_double), ("name", C.c_int), ("n", C.c_int),
                ("eta", P_D), ("zeta", P_D), ("lam", P_I)]


class MlpDesc(C.Structure):
    _fields_ = [("nin", C.c_int), ("nlayer", C.c_int), ("nodes", P_I), ("neuron", C.c_int), ("c", C.c_double),
                ("params", P_D), ("inpw", P_D), ("inpb", P_D), ("outw", C.c_double), ("outb", C.c_double)]


def _f64(x, device=None):
    t = torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64))
    return t.to(device) if device is not None else t


def cutoff_eval(cutoff: int, rc: float, r: torch.Tensor):
    """Cutoff::compute over r (FP64 device tensor): (f, g)."""
    f, g = torch.empty_like(r), torch.empty_like(r)
    _check(lib().ann_cutoff_eval(int(cutoff), C.c_double(rc), r.numel(), _ptr(r), _ptr(f), _ptr(g)))
    return f, g


def radial_eval(cutoff: int, rc: float, name: int, rs, eta, r: torch.Tensor, dEdG: Optional[torch.Tensor] = None):
    """BasisR::symm / symmd / force over distances r: (G, dG/dr, amp)."""
    rs = np.ascontiguousarray(rs, dtype=np.float64)
    eta = np.ascontiguousarray(eta, dtype=np.float64)
    d = RadialDesc(int(cutoff), float(rc), int(name), len(rs), rs.ctypes.data_as(P_D), eta.ctypes.data_as(P_D))
    n = r.numel()
    G = torch.empty((n, len(rs)), dtype=torch.float64, device=r.device)
    dG = torch.empty_like(G)
    amp = torch.empty(n, dtype=torch.float64, device=r.device) if dEdG is not None else None
    _check(lib().ann_radial_eval(C.byref(d), n, _ptr(r), _ptr(G), _ptr(dG), _ptr(dEdG), _ptr(amp)))
    return G, dG, amp


def angular_eval(cutoff: int, rc: float, name: int, eta, zeta, lam, r: torch.Tensor, cos: torch.Tensor,
                 dEdG: Optional[torch.Tensor] = None):
    """BasisA::symm / symmd / force over triplets (r [n][2], cos [n]):
    (G, dG [n][nf][3], phi, eta [n][2])."""
    eta = np.ascontiguousarray(eta, dtype=np.float64)
    zeta = np.ascontiguousarray(zeta, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.int32)
    d = AngularDesc(int(cutoff), float(rc), int(name), len(eta), eta.ctypes.data_as(P_D),
                    zeta.ctypes.data_as(P_D), lam.ctypes.data_as(P_I))
    n = cos.numel()
    G = torch.empty((n, len(eta)), dtype=torch.float64, device=cos.device)
    dG = torch.empty((n, len(eta), 3), dtype=torch.float64, device=cos.device)
    phi = torch.empty(n, dtype=torch.float64, device=cos.device) if dEdG is not None else None
    et = torch.empty((n, 2), dtype=torch.float64, device=cos.device) if dEdG is not None else None
    _check(lib().ann_angular_eval(C.byref(d), n, _ptr(r), _ptr(cos), _ptr(G), _ptr(dG), _ptr(dEdG), _ptr(phi),
                                  _ptr(et)))
    return G, dG, phi, et


class Mlp:
    """One network on the FP64 evaluator (ann_mlp_*): NN::ANN fp / fpbp,
    DODZ::dodi, Cost::grad and D2ODZDI products."""

    def __init__(self, net, device: str = "cuda:0"):
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        nodes = np.ascontiguousarray(net.nodes, dtype=np.int32)
        par = np.ascontiguousarray(net.params, dtype=np.float64)
        iw = np.ascontiguousarray(net.inpw, dtype=np.float64)
        ib = np.ascontiguousarray(net.inpb, dtype=np.float64)
        d = MlpDesc(int(net.nin), len(nodes), nodes.ctypes.data_as(P_I), int(net.neuron), float(net.c),
                    par.ctypes.data_as(P_D), iw.ctypes.data_as(P_D), ib.ctypes.data_as(P_D), float(net.outw),
                    float(net.outb))
        self.h = C.c_void_p()
        _check(lib().ann_mlp_create(C.byref(d), C.byref(self.h)))
        self.nin = int(net.nin)
        self.nparam = int(lib().ann_mlp_size(self.h))

    def __del__(self):
        try:
            lib().ann_mlp_destroy(self.h)
        except Exception:
            pass

    def eval(self, x: torch.Tensor):
        n = x.shape[0]
        out = torch.empty(n, dtype=torch.float64, device=self.device)
        dodi = torch.empty((n, self.nin), dtype=torch.float64, device=self.device)
        _check(lib().ann_mlp_eval(self.h, n, _ptr(x), _ptr(out), _ptr(dodi)))
        return out, dodi

    def grad(self, x: torch.Tensor, dcdo: Optional[torch.Tensor] = None, v: Optional[torch.Tensor] = None,
             per_row: bool = False):
        n = x.shape[0]
        shape = (n, self.nparam) if per_row else (self.nparam,)
        g = torch.empty(shape, dtype=torch.float64, device=self.device)
        _check(lib().ann_mlp_grad(self.h, n, _ptr(x), _ptr(dcdo), _ptr(v), int(per_row), _ptr(g)))
        return g
