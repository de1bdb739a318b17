"""Pins the oracle's special functions, cutoffs and basis functions.

* fmexp, coscut, sincut and every cutoff are compared bit for bit with the
  reference's own src/math/special.cpp and src/nnp/cutoff.cpp compiled from
  source into oracle/_ref (skipped where the reference tree is absent);
* cutoffs against their closed forms and int_0^rc fc dr = rc/2 for COS and
  POLY3 (the check of src/nnp/test_nnp_cutoff.cpp:52-58), derivatives by
  central differences (:26-36);
* radial and angular functions against their defining formulas written
  independently with numpy, and every derivative by central differences
  (src/nnp/test_basis_radial.cpp:29-41, src/nnp/test_basis_angular.cpp:49-65,
  152-171).
"""
import ctypes as C
import math

import numpy as np
import pytest

CUTS = [2, 3, 4]  # COS, TANH, POLY3
RC = 6.0


def cut(orc, name, r):
    f, g = C.c_double(), C.c_double()
    orc.lib().orc_cutoff(name, C.c_double(RC), C.c_double(r), C.byref(f), C.byref(g))
    return f.value, g.value


def test_fmexp_and_taylor_cut_match_reference_bitwise(oracle_mod):
    ref = oracle_mod.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    L = oracle_mod.lib()
    xs = np.concatenate([np.linspace(-50, 5, 2001), np.linspace(0, math.pi, 1001)])
    for x in xs:
        assert L.orc_fmexp(C.c_double(x)) == ref.ref_fmexp(C.c_double(x))
    for x in np.linspace(0, math.pi, 2001):
        assert L.orc_coscut(C.c_double(x)) == ref.ref_coscut(C.c_double(x))
        assert L.orc_sincut(C.c_double(x)) == ref.ref_sincut(C.c_double(x))


@pytest.mark.parametrize("name", [1, 2, 3, 4])
def test_cutoffs_match_reference_bitwise(oracle_mod, name):
    ref = oracle_mod.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    for r in np.linspace(0.0, RC * 1.1, 1201):
        f, g = cut(oracle_mod, name, r)
        v, d = C.c_double(), C.c_double()
        ref.ref_cut_compute(name, C.c_double(RC), C.c_double(r), C.byref(v), C.byref(d))
        assert f == v.value and g == d.value, (name, r)


def test_fmexp_accuracy(oracle_mod):
    L = oracle_mod.lib()
    for x in np.linspace(-40, 10, 777):
        assert abs(L.orc_fmexp(C.c_double(x)) / math.exp(x) - 1) < 1e-13


@pytest.mark.parametrize("name", CUTS)
def test_cutoff_closed_form_and_derivative(oracle_mod, name):
    rs = np.linspace(0.01, RC - 0.01, 500)
    for r in rs:
        f, g = cut(oracle_mod, name, r)
        if name == 2:
            assert abs(f - 0.5 * (math.cos(math.pi * r / RC) + 1)) < 1e-11
        elif name == 3:
            assert abs(f - math.tanh(1 - r / RC) ** 3) < 1e-15
        else:
            x = r / RC
            assert abs(f - (2 * x - 3) * x * x - 1) < 1e-15
        h = 1e-6
        fd = (cut(oracle_mod, name, r + h)[0] - cut(oracle_mod, name, r - h)[0]) / (2 * h)
        assert abs(fd - g) < 1e-8
    assert cut(oracle_mod, name, RC)[0] == 0.0 and cut(oracle_mod, name, RC + 1)[0] == 0.0


@pytest.mark.parametrize("name", [2, 4])
def test_cutoff_integral(oracle_mod, name):
    r = np.linspace(0, RC, 200001)
    f = np.array([cut(oracle_mod, name, x)[0] for x in r[::100]])
    integral = np.trapezoid(f, r[::100])
    assert abs(integral - RC / 2) < 1e-6


def radial(orc, name, cutn, rs, eta, r):
    n = len(rs)
    G, dG = np.zeros(n), np.zeros(n)
    orc.lib().orc_radial(name, cutn, C.c_double(RC), n, np.ascontiguousarray(rs, float).ctypes.data_as(orc.P_D),
                         np.ascontiguousarray(eta, float).ctypes.data_as(orc.P_D), C.c_double(r),
                         G.ctypes.data_as(orc.P_D), dG.ctypes.data_as(orc.P_D))
    return G, dG


def test_radial_gaussian_closed_form(oracle_mod):
    rs = np.array([1.0, 2.3, 4.1])
    eta = np.array([0.7, 1.5, 3.0])
    for r in [0.8, 2.0, 3.7, 5.9]:
        G, _ = radial(oracle_mod, 1, 2, rs, eta, r)
        fc = cut(oracle_mod, 2, r)[0]   # the reference's Taylor-series cosine, pinned above
        assert np.allclose(G, np.exp(-(eta * (r - rs)) ** 2) * fc, rtol=1e-12, atol=0)


def test_radial_family_formulas(oracle_mod):
    rs = np.array([1.5, 3.0])
    eta = np.array([0.9, 2.0])
    r = 2.4
    fc = math.tanh(1 - r / RC) ** 3
    x = eta * (r - rs)
    want = {
        2: 1 / np.cosh(x) * fc,
        3: 1 / np.cosh(x) ** 2 * fc,
        4: 0.5 * (np.tanh(-x) + 1) * fc,
        5: 0.5 * np.log1p(np.exp(-2 * x)) * fc,
        6: (-x) * np.exp(2 * math.sqrt(2) * (-x)) / (np.exp(2 * math.sqrt(2) * (-x)) - 1) * fc,
    }
    for name, w in want.items():
        G, _ = radial(oracle_mod, name, 3, rs, eta, r)
        assert np.allclose(G, w, rtol=1e-10), name


@pytest.mark.parametrize("name", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("cutn", CUTS)
def test_radial_derivative_fd(oracle_mod, name, cutn):
    rng = np.random.default_rng(name * 10 + cutn)
    rs = rng.uniform(0.5, 5.0, 4)
    eta = rng.uniform(0.3, 3.0, 4)
    for r in np.linspace(0.3, RC - 0.05, 40):
        _, dG = radial(oracle_mod, name, cutn, rs, eta, r)
        h = 1e-6
        fd = (radial(oracle_mod, name, cutn, rs, eta, r + h)[0] - radial(oracle_mod, name, cutn, rs, eta, r - h)[0]) / (2 * h)
        assert np.allclose(dG, fd, rtol=1e-6, atol=1e-9), (name, r)


def angular(orc, name, cutn, eta, zeta, lam, r0, r1, c):
    n = len(eta)
    outs = [np.zeros(n) for _ in range(4)]
    orc.lib().orc_angular(name, cutn, C.c_double(RC), n, np.ascontiguousarray(eta, float).ctypes.data_as(orc.P_D),
                          np.ascontiguousarray(zeta, float).ctypes.data_as(orc.P_D),
                          np.ascontiguousarray(lam, np.int32).ctypes.data_as(orc.P_I), C.c_double(r0), C.c_double(r1),
                          C.c_double(c), *[o.ctypes.data_as(orc.P_D) for o in outs])
    return outs


def test_angular_gauss_closed_form(oracle_mod):
    eta = np.array([1.5, 4.0, 10.0])
    zeta = np.array([1.0, 4.0, 2.5])
    lam = np.array([1, -1, 1])
    r0, r1, c = 1.1, 2.7, 0.3
    G = angular(oracle_mod, 1, 2, eta, zeta, lam, r0, r1, c)[0]
    f0 = cut(oracle_mod, 2, r0)[0]
    f1 = cut(oracle_mod, 2, r1)[0]
    want = np.abs(0.5 + 0.5 * lam * c) ** zeta * np.exp(-(r0 ** 2 + r1 ** 2) / eta ** 2) * f0 * f1
    assert np.allclose(G, want, rtol=1e-12)
    # no r_jk term: G is independent of the third side for fixed (r0, r1, cos)


@pytest.mark.parametrize("name", [1, 2, 3, 4, 5, 6])
def test_angular_derivatives_fd(oracle_mod, name):
    rng = np.random.default_rng(100 + name)
    for _ in range(20):
        eta = rng.uniform(1.0, 8.0, 3)
        zeta = rng.uniform(1.0, 6.0, 3)
        lam = rng.choice([-1, 1], 3)
        r0, r1 = rng.uniform(0.8, 5.5, 2)
        c = rng.uniform(-0.95, 0.95)
        G, dc, d0, d1 = angular(oracle_mod, name, 2, eta, zeta, lam, r0, r1, c)
        h = 1e-6
        fc = (angular(oracle_mod, name, 2, eta, zeta, lam, r0, r1, c + h)[0] -
              angular(oracle_mod, name, 2, eta, zeta, lam, r0, r1, c - h)[0]) / (2 * h)
        f0 = (angular(oracle_mod, name, 2, eta, zeta, lam, r0 + h, r1, c)[0] -
              angular(oracle_mod, name, 2, eta, zeta, lam, r0 - h, r1, c)[0]) / (2 * h)
        f1 = (angular(oracle_mod, name, 2, eta, zeta, lam, r0, r1 + h, c)[0] -
              angular(oracle_mod, name, 2, eta, zeta, lam, r0, r1 - h, c)[0]) / (2 * h)
        scale = np.abs(G).max() + 1e-12
        assert np.allclose(dc, fc, rtol=1e-5, atol=1e-8 * scale)
        assert np.allclose(d0, f0, rtol=1e-5, atol=1e-8 * scale)
        assert np.allclose(d1, f1, rtol=1e-5, atol=1e-8 * scale)
