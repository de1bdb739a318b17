"""ctypes binding of the FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None
_REF_RNG = None

P_D = C.POINTER(C.c_double)
P_I = C.POINTER(C.c_int)


class OrcModel(C.Structure):
    _fields_ = [("nspecies", C.c_int), ("rc", C.c_double), ("cutoff", C.c_int),
                ("rad_name", P_I), ("rad_n", P_I), ("rad_off", P_I), ("rad_rs", P_D), ("rad_eta", P_D),
                ("ang_name", P_I), ("ang_n", P_I), ("ang_off", P_I), ("ang_eta", P_D),
                ("ang_zeta", P_D), ("ang_lambda", P_I),
                ("nn_nlayer", P_I), ("nn_nodes_off", P_I), ("nn_nodes", P_I), ("nn_neuron", P_I),
                ("nn_c", P_D), ("nn_par_off", P_I), ("params", P_D),
                ("nn_inp_off", P_I), ("inpw", P_D), ("inpb", P_D),
                ("outw", P_D), ("outb", P_D), ("evac", P_D)]


class OrcStruc(C.Structure):
    _fields_ = [("natoms", C.c_int), ("pos", P_D), ("cell", P_D), ("species", P_I)]


def lib():
    global _LIB
    if _LIB is None:
        # ORACLE_LIB selects another build of the same sources (the sanitizer build)
        path = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.orc_fmexp.restype = C.c_double
        L.orc_coscut.restype = C.c_double
        L.orc_sincut.restype = C.c_double
        L.orc_radial_force.restype = C.c_double
        L.orc_nlist_build.restype = C.c_int64
        L.orc_compute.restype = C.c_double
        L.orc_nn_forward.restype = C.c_double
        L.orc_train_loss_grad.restype = C.c_double
        L.orc_train_loss_grad_ex.restype = C.c_double
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own cutoff/special sources compiled into oracle/_ref (or None)."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_cutoff.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        for f in ("ref_cutf", "ref_cutg", "ref_fmexp", "ref_coscut", "ref_sincut"):
            getattr(L, f).restype = C.c_double
        _REF = L
    return _REF


def ref_rng_lib():
    """The reference's random distributions (src/rand/rng.cpp) and std::mt19937
    compiled into oracle/_ref (or None)."""
    global _REF_RNG
    if _REF_RNG is None:
        path = os.path.join(_HERE, "_ref", "libref_rng.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.ref_rng_draw.argtypes = [C.c_int, C.c_double, C.c_uint, C.c_int, C.c_void_p]
        L.ref_mt_raw.argtypes = [C.c_uint, C.c_int, C.c_void_p]
        L.ref_uniform.argtypes = [C.c_uint, C.c_double, C.c_double, C.c_int, C.c_void_p]
        L.ref_taylor_draws.argtypes = [C.c_int, C.c_double, C.c_uint, C.c_int, C.c_void_p]
        _REF_RNG = L
    return _REF_RNG


def _p(a, t):
    return a.ctypes.data_as(t)


class Oracle:
    """Holds an OrcModel built from an ann_amd.model.Model (keeps arrays alive)."""

    def __init__(self, model):
        self.model = model
        self.flat = model.flat()
        f = self.flat
        self.c = OrcModel(model.nspecies, float(model.rc), int(model.cutoff),
                          *[_p(f[k], P_I if f[k].dtype == np.int32 else P_D) for k in (
                              "rad_name", "rad_n", "rad_off", "rad_rs", "rad_eta",
                              "ang_name", "ang_n", "ang_off", "ang_eta", "ang_zeta", "ang_lambda",
                              "nn_nlayer", "nn_nodes_off", "nn_nodes", "nn_neuron", "nn_c",
                              "nn_par_off", "params", "nn_inp_off", "inpw", "inpb", "outw", "outb",
                              "evac")])
        self.ld = model.max_ninput

    # -- structure helpers
    @staticmethod
    def struc(pos, cell, species):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(3, 3)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        s = OrcStruc(len(pos), _p(pos, P_D), _p(cell, P_D), _p(sp, P_I))
        s._keep = (pos, cell, sp)
        return s

    def nlist(self, pos, cell, species, rc: Optional[float] = None, method: int = 1):
        s = self.struc(pos, cell, species)
        return NList(s, self.model.rc if rc is None else rc, method)

    def symm(self, pos, cell, species, nl=None):
        s = self.struc(pos, cell, species)
        nl = nl or NList(s, self.model.rc, 1)
        G = np.zeros((s.natoms, self.ld))
        lib().orc_symm(C.byref(self.c), C.byref(s), nl.h, _p(G, P_D), self.ld)
        return G

    def net(self, species, G):
        G = np.ascontiguousarray(G, dtype=np.float64)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        E = np.zeros(len(sp))
        dE = np.zeros_like(G)
        lib().orc_net(C.byref(self.c), len(sp), _p(sp, P_I), _p(G, P_D), G.shape[1], _p(E, P_D),
                      _p(dE, P_D))
        return E, dE

    def force(self, pos, cell, species, dEdG, mode: int = 0, nl=None):
        s = self.struc(pos, cell, species)
        nl = nl or NList(s, self.model.rc, 1)
        dEdG = np.ascontiguousarray(dEdG, dtype=np.float64)
        F = np.zeros((s.natoms, 3))
        V = np.zeros((3, 3))
        P = np.zeros((nl.npairs, 3))
        lib().orc_force(C.byref(self.c), C.byref(s), nl.h, _p(dEdG, P_D), dEdG.shape[1], mode,
                        _p(F, P_D), _p(V, P_D), _p(P, P_D))
        return F, V, P

    def compute(self, pos, cell, species, mode: int = 0):
        """NNP::energy (with vacuum energies) and exact forces + virial."""
        s = self.struc(pos, cell, species)
        Ea = np.zeros(s.natoms)
        F = np.zeros((s.natoms, 3))
        V = np.zeros((3, 3))
        E = lib().orc_compute(C.byref(self.c), C.byref(s), mode, _p(Ea, P_D), _p(F, P_D), _p(V, P_D))
        return E, Ea, F, V

    def compute_batch(self, batch, nthreads: int = 1):
        off = np.ascontiguousarray(batch.atom_off, dtype=np.int32)
        pos = np.ascontiguousarray(batch.pos, dtype=np.float64)
        cells = np.ascontiguousarray(batch.cells, dtype=np.float64)
        sp = np.ascontiguousarray(batch.species, dtype=np.int32)
        E = np.zeros(batch.nstruc)
        F = np.zeros((batch.ntotal, 3))
        V = np.zeros((batch.nstruc, 3, 3))
        lib().orc_compute_batch(C.byref(self.c), batch.nstruc, _p(off, P_I), _p(pos, P_D),
                                _p(cells, P_D), _p(sp, P_I), nthreads, _p(E, P_D), _p(F, P_D),
                                _p(V, P_D))
        return E, F, V

    def tangent(self, pos, cell, species, u, nl=None):
        s = self.struc(pos, cell, species)
        nl = nl or NList(s, self.model.rc, 1)
        u = np.ascontiguousarray(u, dtype=np.float64)
        v = np.zeros((s.natoms, self.ld))
        lib().orc_symm_tangent(C.byref(self.c), C.byref(s), nl.h, _p(u, P_D), _p(v, P_D), self.ld)
        return v

    def nn_forward(self, species: int, G):
        G = np.ascontiguousarray(G, dtype=np.float64)
        dE = np.zeros(len(G))
        e = lib().orc_nn_forward(C.byref(self.c), int(species), _p(G, P_D), _p(dE, P_D))
        return e, dE

    def nparams(self, species: int) -> int:
        return lib().orc_nparams(C.byref(self.c), int(species))

    def nn_grad_rop(self, species: int, G, v=None, gE: float = 1.0):
        G = np.ascontiguousarray(G, dtype=np.float64)
        g = np.zeros(self.nparams(species))
        if v is None:
            lib().orc_nn_grad_energy(C.byref(self.c), int(species), _p(G, P_D), C.c_double(gE), _p(g, P_D))
        else:
            v = np.ascontiguousarray(v, dtype=np.float64)
            lib().orc_nn_grad_rop(C.byref(self.c), int(species), _p(G, P_D), _p(v, P_D), C.c_double(gE),
                                  _p(g, P_D))
        return g

    def train_loss_grad(self, batch, Eref, Fref, wE=1.0, wF=0.1, loss=1, delta=1.0, norm=1, wEs=None, wFs=None):
        off = np.ascontiguousarray(batch.atom_off, dtype=np.int32)
        pos = np.ascontiguousarray(batch.pos, dtype=np.float64)
        cells = np.ascontiguousarray(batch.cells, dtype=np.float64)
        sp = np.ascontiguousarray(batch.species, dtype=np.int32)
        Eref = np.ascontiguousarray(Eref, dtype=np.float64)
        Fref = np.ascontiguousarray(Fref, dtype=np.float64)
        n = sum(self.nparams(a) for a in range(self.model.nspecies))
        g = np.zeros(n)
        E = np.zeros(batch.nstruc)
        F = np.zeros((batch.ntotal, 3))
        parts = np.zeros(2)
        wEs = None if wEs is None else np.ascontiguousarray(wEs, dtype=np.float64)
        wFs = None if wFs is None else np.ascontiguousarray(wFs, dtype=np.float64)
        L = lib().orc_train_loss_grad_ex(C.byref(self.c), batch.nstruc, _p(off, P_I), _p(pos, P_D),
                                         _p(cells, P_D), _p(sp, P_I), _p(Eref, P_D), _p(Fref, P_D),
                                         C.c_double(wE), C.c_double(wF), int(loss), C.c_double(delta), int(norm),
                                         None if wEs is None else _p(wEs, P_D),
                                         None if wFs is None else _p(wFs, P_D), _p(g, P_D), _p(E, P_D),
                                         _p(F, P_D), _p(parts, P_D))
        return L, g, E, F, parts


class NList:
    """Canonical neighbour list of one structure built by the oracle."""

    def __init__(self, s, rc: float, method: int = 1):
        self._s = s
        h = C.c_void_p()
        self.npairs = int(lib().orc_nlist_build(C.byref(s), C.c_double(rc), method, C.byref(h)))
        self.h = h
        N = s.natoms
        self.off = np.zeros(N + 1, np.int32)
        self.idx = np.zeros(self.npairs, np.int32)
        self.img = np.zeros((self.npairs, 3), np.int32)
        self.r = np.zeros((self.npairs, 3))
        self.d = np.zeros(self.npairs)
        self.rev = np.zeros(self.npairs, np.int32)
        self.minimg = np.zeros(self.npairs, np.int32)
        lib().orc_nlist_get(h, _p(self.off, P_I), _p(self.idx, P_I), _p(self.img, P_I),
                            _p(self.r, P_D), _p(self.d, P_D), _p(self.rev, P_I), _p(self.minimg, P_I))

    def __del__(self):
        try:
            lib().orc_nlist_free(self.h)
        except Exception:
            pass


def adam_step(p, m1, m2, g, gamma, eps, t):
    """One ADAM::step on a copy of p (moments m1, m2 updated in place)."""
    p = np.array(p, dtype=np.float64, copy=True)
    lib().orc_adam_step(len(p), _p(p, P_D), _p(m1, P_D), _p(m2, P_D),
                        _p(np.ascontiguousarray(g, dtype=np.float64), P_D), C.c_double(gamma),
                        C.c_double(eps), t)
    return p
