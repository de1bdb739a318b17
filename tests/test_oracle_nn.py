"""Pins the oracle's network: activations, forward, dE/dG, dE/dp and the second
backward, against FP64 torch.autograd (a library, independent of the oracle)
and central differences (the patterns of src/ml/test_nn_grad.cpp)."""
import ctypes as C

import numpy as np
import pytest
import torch

from ann_amd.model import NEURON, Model, init_network

NEURONS = ["TANH", "SOFTPLUS", "SIGMOID", "SWISH", "ARCTAN", "ELU", "SQPLUS", "ISRU", "TANHRE", "SQRE",
           "GELU", "MISH", "PFLU", "SERF", "LOGISH", "ATISH"]


def act(orc, neuron, c, z):
    a, d, d2 = C.c_double(), C.c_double(), C.c_double()
    orc.lib().orc_activation(neuron, C.c_double(c), C.c_double(z), C.byref(a), C.byref(d), C.byref(d2))
    return a.value, d.value, d2.value


def torch_act(name, c, z):
    if name == "TANH":
        return torch.tanh(c * z)
    if name == "SOFTPLUS":
        return torch.nn.functional.softplus(c * z) / c - 0.693147180559945 / c
    if name == "SIGMOID":
        return torch.sigmoid(c * z)
    if name == "SWISH":
        return z * torch.sigmoid(c * z)
    if name == "ARCTAN":
        return 2 / np.pi * torch.atan(c * z * np.pi / 2)
    if name == "ELU":
        return torch.where(z >= 0, z, torch.expm1(c * z))
    if name == "SQPLUS":
        return 0.5 / c * (c * z - 1 + torch.sqrt(1 + (c * z) ** 2))
    # the remaining neurons as the reference defines them (src/ml/nn.cpp:238-780;
    # sqre, serf, logish and atish have no sharpness)
    if name == "ISRU":
        return c * z / torch.sqrt(1 + (c * z) ** 2)
    if name == "TANHRE":
        return torch.where(z >= 0, z, torch.tanh(c * z))
    if name == "SQRE":
        return torch.where(z >= 0, z, z / torch.sqrt(1 + z * z))
    if name == "GELU":
        return 0.5 * z * (1 + torch.erf(0.5 * np.sqrt(np.pi) * c * z))
    if name == "MISH":
        return z * torch.tanh(torch.nn.functional.softplus(c * z))
    if name == "PFLU":
        return 0.5 * z * (1 + c * z / torch.sqrt(1 + (c * z) ** 2))
    if name == "SERF":
        return z * torch.erf(torch.nn.functional.softplus(z))
    if name == "LOGISH":
        return z * torch.log(1 + torch.sigmoid(z)) / 0.693147180559945
    if name == "ATISH":
        return 0.5 * z * (1 + 2 / np.pi * torch.atan(np.pi / 2 * z))
    if name == "LINEAR":
        return z
    raise KeyError(name)


@pytest.mark.parametrize("name", NEURONS)
def test_activation_values_and_derivatives(oracle_mod, name):
    c = 1.3
    for z in np.linspace(-4, 4, 81):
        zt = torch.tensor(z, dtype=torch.float64, requires_grad=True)
        a_t = torch_act(name, c, zt)
        (d_t,) = torch.autograd.grad(a_t, zt, create_graph=True)
        (d2_t,) = torch.autograd.grad(d_t, zt) if d_t.requires_grad else (torch.tensor(0.0),)
        a, d, d2 = act(oracle_mod, NEURON[name], c, z)
        assert abs(a - a_t.item()) < 1e-12
        assert abs(d - d_t.item()) < 1e-12
        if name in ("ELU", "SQRE", "TANHRE") and abs(z) < 1e-12:
            continue                    # a kink at z = 0: the second derivative jumps there
        assert abs(d2 - d2_t.item()) < 1e-10, (z, d2, d2_t.item())


def small_model(neuron, seed=0, nin=6, hidden=(5, 4)):
    rng = np.random.default_rng(seed)
    net = init_network(nin, list(hidden), NEURON[neuron], rng, c=1.1, sigma_b=0.3)
    net.inpw = rng.uniform(0.5, 1.5, nin)
    net.inpb = rng.uniform(-0.2, 0.2, nin)
    net.outw, net.outb = 1.7, -0.3
    # one species, radial basis only (the network is what is tested)
    radial = [[(1, np.linspace(1, 4, nin), np.ones(nin))]]
    angular = [[(1, np.zeros(0), np.zeros(0), np.zeros(0, int))]]
    return Model(["X"], np.array([1.0]), np.array([0.0]), 6.0, 2, radial, angular, [net])


def torch_net(model, G):
    net = model.nets[0]
    x = torch.as_tensor(net.inpw) * (G + torch.as_tensor(net.inpb))
    names = {v: k for k, v in NEURON.items()}
    for l, (W, b) in enumerate(net.layers()):
        z = torch.as_tensor(W) @ x + torch.as_tensor(b)
        x = z if l == len(net.nodes) - 1 else torch_act(names[net.neuron], net.c, z)
    return net.outb + net.outw * x[0]


def params_tensor(model):
    return torch.tensor(model.nets[0].params, dtype=torch.float64, requires_grad=True)


def torch_net_p(model, G, p):
    net = model.nets[0]
    names = {v: k for k, v in NEURON.items()}
    x = torch.as_tensor(net.inpw) * (G + torch.as_tensor(net.inpb))
    cnt, prev, bs = 0, net.nin, []
    for h in net.nodes:
        bs.append(p[cnt:cnt + h])
        cnt += h
    for l, h in enumerate(net.nodes):
        W = p[cnt:cnt + h * prev].reshape(prev, h).T
        cnt += h * prev
        z = W @ x + bs[l]
        x = z if l == len(net.nodes) - 1 else torch_act(names[net.neuron], net.c, z)
        prev = h
    return net.outb + net.outw * x[0]


@pytest.mark.parametrize("neuron", ["TANH", "SOFTPLUS", "SWISH"])
def test_forward_dEdG_and_param_gradient_vs_autograd(oracle_mod, neuron):
    model = small_model(neuron)
    orc = oracle_mod.Oracle(model)
    rng = np.random.default_rng(1)
    for _ in range(5):
        G = rng.uniform(0, 2, 6)
        e, dE = orc.nn_forward(0, G)
        Gt = torch.tensor(G, requires_grad=True)
        p = params_tensor(model)
        Et = torch_net_p(model, Gt, p)
        gG, gp = torch.autograd.grad(Et, (Gt, p))
        assert abs(e - Et.item()) < 1e-12
        assert np.allclose(dE, gG.numpy(), atol=1e-12, rtol=1e-10)
        gE = 0.7
        g = orc.nn_grad_rop(0, G, None, gE)
        assert np.allclose(g, gE * gp.numpy(), atol=1e-12, rtol=1e-10)


@pytest.mark.parametrize("neuron", ["TANH", "SOFTPLUS", "SWISH", "SIGMOID", "MISH", "GELU", "SERF", "LOGISH",
                                    "ATISH", "ISRU", "PFLU", "TANHRE", "SQRE"])
def test_second_backward_vs_double_autograd_and_fd(oracle_mod, neuron):
    """d/dp [gE E + dE/dG . v]: torch double backward (library) and central
    differences of the same scalar w.r.t. the parameters."""
    model = small_model(neuron, seed=3)
    orc = oracle_mod.Oracle(model)
    rng = np.random.default_rng(2)
    G = rng.uniform(0, 2, 6)
    v = rng.normal(size=6)
    gE = -0.4
    g = orc.nn_grad_rop(0, G, v, gE)
    Gt = torch.tensor(G, requires_grad=True)
    p = params_tensor(model)
    Et = torch_net_p(model, Gt, p)
    (dG,) = torch.autograd.grad(Et, Gt, create_graph=True)
    phi = gE * Et + (dG * torch.as_tensor(v)).sum()
    (gp,) = torch.autograd.grad(phi, p)
    assert np.allclose(g, gp.numpy(), atol=1e-11, rtol=1e-9)
    # central differences of phi(p) with the oracle itself
    h = 1e-5
    base = model.nets[0].params.copy()
    for k in rng.choice(len(base), 12, replace=False):
        vals = []
        for s in (1, -1):
            model.nets[0].params = base.copy()
            model.nets[0].params[k] += s * h
            o2 = oracle_mod.Oracle(model)
            e, d = o2.nn_forward(0, G)
            vals.append(gE * e + d @ v)
        model.nets[0].params = base.copy()
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * max(1, abs(g[k])), (k, fd, g[k])


def test_param_packing_order(oracle_mod):
    """Gradient layout = biases layer by layer, then weights column-major
    (Cost::grad, src/ml/nn.cpp:1732-1775): the bias of neuron n of layer 0
    sits at index n, weight (n, m) of layer 0 at sum(nodes) + m*nodes[0] + n."""
    model = small_model("TANH", seed=4)
    orc = oracle_mod.Oracle(model)
    G = np.linspace(0.1, 1.0, 6)
    g = orc.nn_grad_rop(0, G, None, 1.0)
    net = model.nets[0]
    nb = sum(net.nodes)
    # dE/dW0(n,m) = dE/db0(n) * s_m with s = inpw (G + inpb)
    s = net.inpw * (G + net.inpb)
    for n in range(net.nodes[0]):
        for m in range(net.nin):
            assert abs(g[nb + m * net.nodes[0] + n] - g[n] * s[m]) < 1e-14


def test_adam_first_step(oracle_mod):
    """After one step from zero moments the Adam update is gamma*g/(|g|+eps')
    (bias corrections cancel): p1 = p0 - gamma*sign(g) up to eps."""
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -1e-3, 4.0])
    m1, m2 = np.zeros(3), np.zeros(3)
    p1 = oracle_mod.adam_step(p, m1, m2, g, 0.01, 1e-16, 1)
    assert np.allclose(p1, p - 0.01 * np.sign(g), rtol=0, atol=1e-12)
    assert np.allclose(m1, 0.1 * g) and np.allclose(m2, 0.001 * g * g)
