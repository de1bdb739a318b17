"""Error paths of the C ABI (include/ann_gpu.h): invalid descriptors and
arguments, calls out of order and invalid positions come back as status codes
with a message, and the context stays usable.  The reference throws
std::invalid_argument for the same descriptor faults (src/nnp/cutoff.cpp:51-52,
src/nnp/basis_radial.cpp:203, src/ml/nn.cpp:1022-1023)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ARG, UNSUP, CAP, HIP, STATE = -1, -2, -3, -4, -5


def _code(fn):
    from ann_amd.gpu import AnnError
    with pytest.raises(AnnError) as ei:
        fn()
    return ei.value.code, str(ei.value)


def _model_with(change):
    from ann_amd import configs
    m = configs.model_water()
    change(m)
    return m


@pytest.mark.parametrize("what,change,code", [
    ("cutoff name 0", lambda m: setattr(m, "cutoff", 0), ARG),
    ("cutoff name 5", lambda m: setattr(m, "cutoff", 5), ARG),
    ("rc <= 0", lambda m: setattr(m, "rc", 0.0), ARG),
    ("radial name 7", lambda m: m.radial[0].__setitem__(0, (7,) + tuple(m.radial[0][0][1:])), ARG),
    ("angular name 9", lambda m: m.angular[0].__setitem__(0, (9,) + tuple(m.angular[0][0][1:])), ARG),
    ("angular width 0", lambda m: m.angular[1].__setitem__(0, (1, np.zeros(16)) + tuple(m.angular[1][0][2:])), ARG),
    ("neuron 0", lambda m: setattr(m.nets[0], "neuron", 0), UNSUP),
    ("neuron 19 (TEST)", lambda m: setattr(m.nets[1], "neuron", 19), UNSUP),
])
def test_invalid_descriptor(what, change, code):
    from ann_amd.gpu import Engine
    m = _model_with(change)
    print(what, flush=True)
    got, msg = _code(lambda: Engine(m, "cuda:0"))
    assert got == code, (what, got, msg)


def test_invalid_shapes():
    from ann_amd import configs
    from ann_amd.gpu import Engine
    from ann_amd.model import init_network
    m = configs.model_water()
    rng = np.random.default_rng(0)
    m.nets[0] = init_network(80, [129], 3, rng)                 # a layer wider than 128
    assert _code(lambda: Engine(m, "cuda:0"))[0] == UNSUP
    m = configs.model_water()
    m.nets[0] = init_network(80, [8, 8, 8, 8], 3, rng)          # 5 layers
    assert _code(lambda: Engine(m, "cuda:0"))[0] == UNSUP
    m = configs.model_water()
    rs = np.linspace(0.8, 6.0, 33)
    m.radial[0][0] = (1, rs, np.ones(33))                        # 33 functions in one block
    m.nets[0] = init_network(m.ninput(0), [8], 3, rng)
    assert _code(lambda: Engine(m, "cuda:0"))[0] == UNSUP


def test_calls_out_of_order_and_bad_arguments():
    import torch
    from ann_amd import configs
    from ann_amd.gpu import Engine, _check, _ptr, lib
    model = configs.model_water()
    eng = Engine(model, "cuda:0")
    G = torch.empty((192, eng.ldg), dtype=torch.float32, device="cuda:0")
    assert _code(lambda: _check(lib().ann_neighbors(eng.h, C.c_double(0.0))))[0] == STATE
    eng.set_batch(configs.batch_water(1, seed=3))
    assert _code(lambda: _check(lib().ann_symm_fwd(eng.h, 1, _ptr(G))))[0] == STATE
    assert _code(lambda: eng.neighbors(-1.0))[0] == ARG
    assert _code(lambda: eng.neighbors(float("nan")))[0] == ARG
    eng.neighbors()
    assert _code(lambda: _check(lib().ann_symm_fwd(eng.h, 7, _ptr(G))))[0] == ARG      # form
    assert _code(lambda: _check(lib().ann_symm_fwd(eng.h, 1, _ptr(None))))[0] == ARG   # null output
    assert _code(lambda: eng.set_params(np.zeros(3)))[0] == ARG
    assert _code(lambda: eng.set_tuning([3, 1, 1, 0, 0]))[0] == ARG
    assert _code(lambda: eng.set_tuning([1, 1, 1, 0, 0, 2]))[0] == ARG
    assert _code(lambda: eng.set_tuning([1, 1, 1, 0, 0, 0, 3]))[0] == ARG
    assert _code(lambda: eng.set_tuning([1, 1, 1, 0, 0, 0, 0, 2]))[0] == ARG
    assert _code(lambda: eng.set_loss("MSE", -1.0))[0] == ARG
    assert _code(lambda: _check(lib().ann_set_loss(eng.h, 9, C.c_double(1.0), 1)))[0] == ARG
    assert _code(lambda: eng.set_scaling(5, np.ones(80), np.zeros(80)))[0] == ARG
    w = np.ones(80)
    w[3] = np.inf
    assert _code(lambda: eng.set_scaling(0, w, np.zeros(80)))[0] == ARG
    # the context still works
    E, F, _ = eng.energy_forces(1)
    assert np.isfinite(F.cpu().numpy()).all()


def test_nonfinite_positions_are_rejected():
    from ann_amd import configs
    from ann_amd.gpu import Engine
    b = configs.batch_water(2, seed=4)
    pos = b.pos.copy()
    pos[7, 1] = np.nan
    bad = configs.Batch(b.natoms, pos, b.cells, b.species)
    eng = Engine(configs.model_water(), "cuda:0")
    eng.set_batch(bad)
    code, msg = _code(lambda: eng.neighbors())
    assert code == ARG and "finite" in msg
    pos[7, 1] = np.inf
    eng.set_batch(configs.Batch(b.natoms, pos, b.cells, b.species))
    assert _code(lambda: eng.neighbors())[0] == ARG
    eng.set_batch(b)                  # the same context with valid positions
    eng.neighbors()
    assert eng.nlist_size()[0] > 0
