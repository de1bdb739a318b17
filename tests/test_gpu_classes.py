"""The FP64 evaluators behind the per-function classes (Cutoff, BasisR,
BasisA, NN::ANN, Cost, DODZ, D2ODZDI) against the FP64 oracle's functions of
the same names, for every cutoff, radial and angular family and neuron.

Bars: max |GPU - oracle| / max |oracle| <= 1e-10 per output (both FP64; the
oracle keeps the reference's fmexp / coscut / sincut, which differ from the
device exp / cos / sin by < 1e-11); network input derivatives also against
central differences of the GPU's own FP64 outputs (<= 1e-7 relative)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)


def _p(a, t=PD):
    return a.ctypes.data_as(t)


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _dev(x):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64)).cuda()


@pytest.mark.parametrize("cut", [1, 2, 3, 4])
def test_cutoff(oracle_mod, cut):
    from ann_amd import gpu
    L = oracle_mod.lib()
    rc = 5.5
    r = np.concatenate([np.linspace(0.05, 1.1 * rc, 2001), [rc]])
    f, g = gpu.cutoff_eval(cut, rc, _dev(r))
    fo, go = np.zeros(len(r)), np.zeros(len(r))
    for i, x in enumerate(r):
        a, b = C.c_double(), C.c_double()
        L.orc_cutoff(cut, C.c_double(rc), C.c_double(x), C.byref(a), C.byref(b))
        fo[i], go[i] = a.value, b.value
    print(f"cutoff {cut}: f {_rel(f.cpu(), fo):.2e} g {_rel(g.cpu(), go) if np.abs(go).max() else 0:.2e}", flush=True)
    assert _rel(f.cpu(), fo) <= 1e-10
    if np.abs(go).max() > 0:
        assert _rel(g.cpu(), go) <= 1e-10
    else:
        assert np.abs(g.cpu().numpy()).max() == 0


@pytest.mark.parametrize("name", [1, 2, 3, 4, 5, 6])
def test_radial(oracle_mod, name):
    from ann_amd import gpu
    L = oracle_mod.lib()
    rng = np.random.default_rng(name)
    cut, rc, n = 1 + name % 4, 6.0, 8
    rs = rng.uniform(0.5, 5.0, n)
    eta = rng.uniform(0.5, 3.0, n)
    r = np.concatenate([rng.uniform(0.3, 1.05 * rc, 1500), [rs[0]]])      # rs[0]: the LOGCOSH2 a == 0 point
    dEdG = rng.standard_normal((len(r), n))
    G, dG, amp = gpu.radial_eval(cut, rc, name, rs, eta, _dev(r), _dev(dEdG))
    Go, dGo, ampo = np.zeros((len(r), n)), np.zeros((len(r), n)), np.zeros(len(r))
    L.orc_radial_force.restype = C.c_double
    for i, x in enumerate(r):
        gi, di = np.zeros(n), np.zeros(n)
        L.orc_radial(name, cut, C.c_double(rc), n, _p(rs), _p(eta), C.c_double(x), _p(gi), _p(di))
        Go[i], dGo[i] = gi, di
        ampo[i] = L.orc_radial_force(name, cut, C.c_double(rc), n, _p(rs), _p(eta), C.c_double(x),
                                     _p(np.ascontiguousarray(dEdG[i])))
    e = (_rel(G.cpu(), Go), _rel(dG.cpu(), dGo), _rel(amp.cpu(), ampo))
    print(f"radial {name} cut {cut}: G {e[0]:.2e} dG {e[1]:.2e} amp {e[2]:.2e}", flush=True)
    assert max(e) <= 1e-10


@pytest.mark.parametrize("name", [1, 2, 3, 4, 5, 6])
def test_angular(oracle_mod, name):
    from ann_amd import gpu
    L = oracle_mod.lib()
    rng = np.random.default_rng(10 + name)
    cut, rc, n = 1 + name % 4, 6.0, 8
    eta = rng.uniform(1.0, 4.0, n)
    zeta = np.array([1, 2, 4, 8, 1.5, 3, 1, 2.5], float)
    lam = np.array([1, -1] * 4, np.int32)
    m = 1500
    r = rng.uniform(0.5, 1.05 * rc, (m, 2))
    cs = rng.uniform(-0.999, 0.999, m)
    dEdG = rng.standard_normal((m, n))
    G, dG, phi, et = gpu.angular_eval(cut, rc, name, eta, zeta, lam, _dev(r), _dev(cs), _dev(dEdG))
    Go, dGo = np.zeros((m, n)), np.zeros((m, n, 3))
    phio, eto = np.zeros(m), np.zeros((m, 2))
    for i in range(m):
        g, dc, d0, d1 = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        L.orc_angular(name, cut, C.c_double(rc), n, _p(eta), _p(zeta), _p(lam, PI), C.c_double(r[i, 0]),
                      C.c_double(r[i, 1]), C.c_double(cs[i]), _p(g), _p(dc), _p(d0), _p(d1))
        Go[i] = g
        dGo[i] = np.stack([dc, d0, d1], 1)
        p, e2 = C.c_double(), np.zeros(2)
        L.orc_angular_force(name, cut, C.c_double(rc), n, _p(eta), _p(zeta), _p(lam, PI), C.c_double(r[i, 0]),
                            C.c_double(r[i, 1]), C.c_double(cs[i]), _p(np.ascontiguousarray(dEdG[i])), C.byref(p),
                            _p(e2))
        phio[i], eto[i] = p.value, e2
    e = (_rel(G.cpu(), Go), _rel(dG.cpu(), dGo), _rel(phi.cpu(), phio), _rel(et.cpu(), eto))
    print(f"angular {name} cut {cut}: G {e[0]:.2e} dG {e[1]:.2e} phi {e[2]:.2e} eta {e[3]:.2e}", flush=True)
    assert max(e) <= 1e-10


NEURONS = ["LINEAR", "SIGMOID", "TANH", "ISRU", "ARCTAN", "RELU", "ELU", "TANHRE", "SQRE", "SWISH", "GELU",
           "MISH", "PFLU", "SERF", "LOGISH", "SOFTPLUS", "SQPLUS", "ATISH"]


def _network(neuron, seed=5):
    from ann_amd import configs
    m = configs.model_family(neuron=neuron, nr=6, hidden=(16, 9), seed=seed)
    rng = np.random.default_rng(seed)
    net = m.nets[0]
    net.params = rng.standard_normal(net.nparams) * 0.4
    net.inpw = rng.uniform(0.5, 2.0, net.nin)
    net.inpb = rng.uniform(-0.3, 0.3, net.nin)
    net.outw, net.outb = 1.7, -0.2
    net.c = 1.3
    return m, net


@pytest.mark.parametrize("neuron", NEURONS)
def test_network(oracle_mod, neuron):
    import torch
    from ann_amd import gpu
    m, net = _network(neuron)
    orc = oracle_mod.Oracle(m)
    L = oracle_mod.lib()
    rng = np.random.default_rng(17)
    nrow = 64
    x = rng.standard_normal((nrow, net.nin))
    dcdo = rng.standard_normal(nrow)
    v = rng.standard_normal((nrow, net.nin))
    mlp = gpu.Mlp(net)
    out, dodi = mlp.eval(_dev(x))
    g_sum = mlp.grad(_dev(x), _dev(dcdo))
    g_row = mlp.grad(_dev(x), _dev(dcdo), per_row=True)
    r_sum = mlp.grad(_dev(x), _dev(dcdo), _dev(v))
    r_row = mlp.grad(_dev(x), _dev(dcdo), _dev(v), per_row=True)
    np_ = orc.nparams(0)
    assert mlp.nparam == np_
    outo, dodio = np.zeros(nrow), np.zeros((nrow, net.nin))
    go, ro = np.zeros((nrow, np_)), np.zeros((nrow, np_))
    for i in range(nrow):
        outo[i], dodio[i] = orc.nn_forward(0, x[i])
        g = np.zeros(np_)
        L.orc_nn_grad_energy(C.byref(orc.c), 0, _p(np.ascontiguousarray(x[i])), C.c_double(dcdo[i]), _p(g))
        go[i] = g
        ro[i] = orc.nn_grad_rop(0, x[i], v[i], gE=float(dcdo[i]))
    e = dict(out=_rel(out.cpu(), outo), dodi=_rel(dodi.cpu(), dodio), grad=_rel(g_row.cpu(), go),
             gsum=_rel(g_sum.cpu(), go.sum(0)), rop=_rel(r_row.cpu(), ro), rsum=_rel(r_sum.cpu(), ro.sum(0)))
    # do/di against central differences of the GPU's FP64 outputs
    h = 1e-5
    k = np.arange(nrow) % net.nin
    xp, xm = x.copy(), x.copy()
    xp[np.arange(nrow), k] += h
    xm[np.arange(nrow), k] -= h
    fd = (mlp.eval(_dev(xp))[0] - mlp.eval(_dev(xm))[0]).cpu().numpy() / (2 * h)
    e["fd"] = _rel(dodi.cpu().numpy()[np.arange(nrow), k], fd)
    print(f"{neuron}: " + " ".join(f"{a} {b:.2e}" for a, b in e.items()), flush=True)
    assert e.pop("fd") <= 1e-7
    assert max(e.values()) <= 1e-10, e
    # the per-row sum of the summed call is the same in row order
    assert torch.equal(g_sum, g_row.sum(0)) or _rel(g_sum.cpu(), g_row.sum(0).cpu()) <= 1e-14


def test_network_errors():
    from ann_amd import gpu
    from ann_amd.gpu import AnnError
    m, net = _network("TANH")
    net.nodes = list(net.nodes[:-1]) + [2]          # two outputs
    with pytest.raises(AnnError) as ei:
        gpu.Mlp(net)
    assert ei.value.code == -2
    m, net = _network("TANH")
    net.neuron = 99
    with pytest.raises(AnnError) as ei:
        gpu.Mlp(net)
    assert ei.value.code == -1
    m, net = _network("TANH")
    net.nodes = [4096] + list(net.nodes[1:])        # node state beyond the LDS budget
    net.params = np.zeros(4096 + 4096 * net.nin + sum(net.nodes[1:]) + 4096 * net.nodes[1] + 9)
    with pytest.raises(AnnError) as ei:
        gpu.Mlp(net)
    assert ei.value.code == -2


@pytest.mark.parametrize("cut,rad,ang,neuron", [(2, 1, 1, "TANH"), (3, 6, 2, "SOFTPLUS"), (4, 3, 5, "GELU")])
def test_cpp_classes(tmp_path, oracle_mod, cut, rad, ang, neuron):
    """The C++ classes one point at a time and batched: single and batched
    results identical, both against the oracle."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    L = oracle_mod.lib()
    rng = np.random.default_rng(cut * 100 + rad * 10 + ang)
    rc, nr, na = 6.0, 5, 6
    rs, reta = rng.uniform(0.5, 5.0, nr), rng.uniform(0.5, 3.0, nr)
    aeta, zeta = rng.uniform(1.0, 4.0, na), np.array([1, 2, 4, 1.5, 3, 8], float)
    lam = np.array([1, -1, 1, -1, 1, -1], np.int32)
    npts, ntrip = 40, 30
    r = rng.uniform(0.3, 1.05 * rc, npts)
    trip = np.stack([rng.uniform(0.5, 1.05 * rc, ntrip), rng.uniform(0.5, 1.05 * rc, ntrip),
                     rng.uniform(-0.99, 0.99, ntrip)], 1)
    dEr, dEa = rng.standard_normal((npts, nr)), rng.standard_normal((ntrip, na))
    m, net = _network(neuron, seed=cut + rad)
    nrow = 12
    x, dcdo, v = rng.standard_normal((nrow, net.nin)), rng.standard_normal(nrow), rng.standard_normal((nrow, net.nin))
    from ann_amd.model import NEURON
    tok = [cut, rc, rad, nr, *rs, *reta, ang, na, *aeta, *zeta, *lam, npts, *r, ntrip, *trip.reshape(-1),
           *dEr.reshape(-1), *dEa.reshape(-1), net.nin, len(net.nodes), *net.nodes, NEURON[neuron], net.c,
           net.outw, net.outb, net.nparams, *net.params, *net.inpw, *net.inpb, nrow, *x.reshape(-1), *dcdo,
           *v.reshape(-1)]
    spec, out = tmp_path / "spec.txt", tmp_path / "out.txt"
    spec.write_text(" ".join(repr(float(t)) if isinstance(t, (float, np.floating)) else str(int(t)) for t in tok))
    print(f"classes cut {cut} rad {rad} ang {ang} {neuron}", flush=True)
    p = subprocess.run([os.path.join(root, "tests", "cpp", "test_classes_gpu"), str(spec), str(out)],
                       capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    got = {}
    for line in out.read_text().splitlines():
        t = line.split()
        got[t[0]] = np.array([float(z) for z in t[2:]])
    for a, b in [("cutf", "cutf_b"), ("cutg", "cutg_b"), ("rsymm", "rsymm_b"), ("rforce", "rforce_b"),
                 ("asymm", "asymm_b"), ("out", "out_b"), ("dodi", "dodi_b")]:
        assert np.array_equal(got[a], got[b]), (a, b)
    assert np.array_equal(got["aforce"].reshape(-1, 3)[:, 0], got["aphi_b"])
    assert np.array_equal(got["aforce"].reshape(-1, 3)[:, 1:], got["aeta_b"].reshape(-1, 2))
    assert got["errors"][0] == 5
    # oracle
    fo, go = [], []
    Rg, Rd, Ramp, Ag, Af, Asf, Asd = [], [], [], [], [], [], []
    for d in r:
        a, b = C.c_double(), C.c_double()
        L.orc_cutoff(cut, C.c_double(rc), C.c_double(d), C.byref(a), C.byref(b))
        fo.append(a.value), go.append(b.value)
        gi, di = np.zeros(nr), np.zeros(nr)
        L.orc_radial(rad, cut, C.c_double(rc), nr, _p(rs), _p(reta), C.c_double(d), _p(gi), _p(di))
        Rg.append(gi), Rd.append(di[0])
    for i, d in enumerate(r):
        Ramp.append(-np.dot(dEr[i], _radial_d(L, rad, cut, rc, rs, reta, d)))
    for t in range(ntrip):
        g, dc, d0, d1 = np.zeros(na), np.zeros(na), np.zeros(na), np.zeros(na)
        L.orc_angular(ang, cut, C.c_double(rc), na, _p(aeta), _p(zeta), _p(lam, PI), C.c_double(trip[t, 0]),
                      C.c_double(trip[t, 1]), C.c_double(trip[t, 2]), _p(g), _p(dc), _p(d0), _p(d1))
        Ag.append(g)
        Af.append([-np.dot(dEa[t], dc), -np.dot(dEa[t], d0), -np.dot(dEa[t], d1)])
        Asf.append(g[0])
        Asd.append([dc[0], d0[0], d1[0]])
    orc = oracle_mod.Oracle(m)
    outo = np.array([orc.nn_forward(0, x[q])[0] for q in range(nrow)])
    dodio = np.concatenate([orc.nn_forward(0, x[q])[1] for q in range(nrow)])
    np_ = orc.nparams(0)
    cost0 = np.zeros(np_)
    L.orc_nn_grad_energy(C.byref(orc.c), 0, _p(np.ascontiguousarray(x[0])), C.c_double(dcdo[0]), _p(cost0))
    d2 = np.stack([orc.nn_grad_rop(0, x[0], np.eye(net.nin)[k], gE=0.0) for k in range(net.nin)], 1)
    costb = sum(orc.nn_grad_rop(0, x[q], None, gE=float(dcdo[q])) for q in range(nrow))
    ropb = sum(orc.nn_grad_rop(0, x[q], v[q], gE=float(dcdo[q])) for q in range(nrow))
    net2 = m.copy()
    net2.nets[0].params = net.params.copy()
    net2.nets[0].params[0] += 0.25
    shift = oracle_mod.Oracle(net2).nn_forward(0, x[0])[0]
    e = dict(cutf=_rel(got["cutf"], fo), cutg=_rel(got["cutg"], go) if np.abs(go).max() else 0.0,
             rsymm=_rel(got["rsymm"], np.concatenate(Rg)), rsymmf=_rel(got["rsymmf"], np.array(Rg)[:, 0]),
             rsymmd=_rel(got["rsymmd"], Rd), rforce=_rel(got["rforce"], Ramp),
             asymm=_rel(got["asymm"], np.concatenate(Ag)), aforce=_rel(got["aforce"], np.ravel(Af)),
             asymmf=_rel(got["asymmf"], Asf), asymmd=_rel(got["asymmd"], np.ravel(Asd)),
             out=_rel(got["out"], outo), dodi=_rel(got["dodi"], dodio), cost0=_rel(got["cost0"], cost0),
             d2=_rel(got["d2odpdi0"], d2.reshape(-1)), costb=_rel(got["cost_b"], costb),
             ropb=_rel(got["rop_b"], ropb), shift=_rel(got["out_shift"], [shift]))
    print(" ".join(f"{a} {b:.1e}" for a, b in e.items()), flush=True)
    assert max(e.values()) <= 1e-10, e


def _radial_d(L, rad, cut, rc, rs, reta, d):
    gi, di = np.zeros(len(rs)), np.zeros(len(rs))
    L.orc_radial(rad, cut, C.c_double(rc), len(rs), _p(rs), _p(reta), C.c_double(d), _p(gi), _p(di))
    return di
