"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle on the
same seeded inputs.

Tolerances (stated, per quantity):
  neighbour list (set, order, image vectors, reverse slots): bit-exact;
  pair geometry: FP32 rounding of the FP64 vector (<= 2 ulp);
  symmetry functions: max |dG| <= 2e-5 * max |G| per structure;
  energies: <= 1e-6 eV/atom;
  forces: relative RMS <= 1e-5 over the batch;
  per-slot adjoint P: relative RMS <= 1e-5.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FORMS = [0, 1, 2]   # plain, fast (moment form where it applies), fast with triplet tiles


def _engine(model, batch):
    from ann_amd.gpu import Engine
    eng = Engine(model, "cuda:0")
    eng.set_batch(batch)
    eng.neighbors()
    return eng


def _cases():
    from ann_amd import configs
    return {
        "ar256": (configs.model_ar256(), configs.struc_ar256()),
        "water4": (configs.model_water(), configs.batch_water(4, seed=21)),
        "uneven6": (configs.model_water(), configs.batch_uneven(6, seed=22, nmax=300)),
        "lips3": (configs.model_lips(), configs.batch_lips(3, seed=23)),
        "random20": (configs.model_lips(), configs.small_random(24, natoms=20, a0=9.0)),
        # search boxes too large to stage in LDS (2500 candidates; 343 bin images)
        "tiny20": (configs.model_lips(), configs.small_random(25, natoms=20, a0=5.9)),
        "tiny1": (configs.model_lips(), configs.single([[0.3, 0.2, 0.1]], np.eye(3) * 2.5, [0])),
        # pairs exactly at the cutoff (simple cubic, 3.0 A spacing, rc = 6.0): the FP64 test decides
        "lattice216": (configs.model_ar256(), configs.single(_lattice(0.0, 0.0), np.eye(3) * 18.0,
                                                              np.zeros(216, np.int32))),
        # the same with 1e-9 A jitter, positions 1000 A away from the cell (large image shifts)
        "lattice_far": (configs.model_ar256(), configs.single(_lattice(1e-9, 1000.0), np.eye(3) * 18.0,
                                                               np.zeros(216, np.int32))),
    }


def _lattice(jitter, shift):
    g = np.array([[i, j, k] for i in range(6) for j in range(6) for k in range(6)], float) * 3.0
    return g + np.random.default_rng(31).normal(0.0, jitter, g.shape) + shift if jitter else g + shift


CASES = ["ar256", "water4", "uneven6", "lips3", "random20", "tiny20", "tiny1"]
NL_CASES = CASES + ["lattice216", "lattice_far"]


@pytest.fixture(scope="module")
def cases():
    return _cases()


def _check_nlist(eng, batch, orc, rc=None):
    nl = eng.nlist()
    off = batch.atom_off
    base = 0
    for k in range(batch.nstruc):
        pos, cell, sp = batch.struc(k)
        ref = orc.nlist(pos, cell, sp, rc)
        a0, a1 = off[k], off[k + 1]
        s0, s1 = nl["off"][a0], nl["off"][a1]
        assert np.array_equal(nl["off"][a0:a1 + 1] - s0, ref.off)
        assert np.array_equal(nl["idx"][s0:s1] - a0, ref.idx)
        assert np.array_equal(nl["img"][s0:s1], ref.img)
        assert np.array_equal(nl["rev"][s0:s1] - s0, ref.rev)
        geo = nl["geo"][s0:s1].astype(np.float64)
        assert np.all(np.abs(geo[:, :3] - ref.r) <= 2 * np.spacing(np.abs(ref.r).astype(np.float32)) + 1e-30)
        assert np.all(np.abs(geo[:, 3] - ref.d) <= 2 * np.spacing(ref.d.astype(np.float32)))
        base += s1 - s0
    assert base == len(nl["idx"])


@pytest.mark.parametrize("name", NL_CASES)
def test_neighbor_list_bit_exact(cases, oracle_mod, name):
    model, batch = cases[name]
    print(f"case {name}", flush=True)
    eng = _engine(model, batch)
    _check_nlist(eng, batch, oracle_mod.Oracle(model))


def test_rebuilds_with_growing_and_shrinking_lists(cases, oracle_mod):
    """A short cutoff, then the model cutoff (the list outgrows its arrays: the
    compaction that ran behind the scan is void and is repeated after growing),
    then the short cutoff again (the compaction behind the scan stands)."""
    model, batch = cases["water4"]
    from ann_amd.gpu import Engine
    eng = Engine(model, "cuda:0")
    eng.set_batch(batch)
    orc = oracle_mod.Oracle(model)
    for rc in (3.0, model.rc, 3.0, model.rc):
        print(f"rc {rc}", flush=True)
        eng.neighbors(rc)
        _check_nlist(eng, batch, orc, rc)


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("name", CASES)
def test_symm_parity(cases, oracle_mod, name, form):
    model, batch = cases[name]
    print(f"case {name} form {form}", flush=True)
    eng = _engine(model, batch)
    G = eng.symm(form).cpu().numpy().astype(np.float64)
    orc = oracle_mod.Oracle(model)
    off = batch.atom_off
    for k in range(batch.nstruc):
        pos, cell, sp = batch.struc(k)
        Gr = orc.symm(pos, cell, sp)
        Gg = G[off[k]:off[k + 1], :Gr.shape[1]]
        err = np.abs(Gg - Gr).max() / max(np.abs(Gr).max(), 1e-30)
        assert err <= 2e-5, (k, err)


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("name", CASES)
def test_energy_force_parity(cases, oracle_mod, name, form):
    model, batch = cases[name]
    print(f"case {name} form {form}", flush=True)
    eng = _engine(model, batch)
    E, F, V = eng.energy_forces(form)
    E, F, V = E.cpu().numpy(), F.cpu().numpy(), V.cpu().numpy()
    orc = oracle_mod.Oracle(model)
    Eo, Fo, Vo = orc.compute_batch(batch, nthreads=4)
    de = np.max(np.abs(E - Eo) / batch.natoms)
    if (Fo ** 2).sum() > 1e-20:
        rf = np.sqrt(((F - Fo) ** 2).sum() / (Fo ** 2).sum())
    else:   # forces vanish by symmetry (one atom and its images): absolute bound in eV/A
        rf = np.abs(F).max()
    rv = np.abs(V - Vo).max() / max(np.abs(Vo).max(), 1e-30)
    print(f"  dE/atom={de:.3e} force relRMS={rf:.3e} virial rel={rv:.3e}", flush=True)
    assert de <= 1e-6
    assert rf <= 1e-5
    assert rv <= 1e-4


@pytest.mark.parametrize("form", FORMS)
def test_slot_adjoint_parity(cases, oracle_mod, form):
    model, batch = cases["water4"]
    eng = _engine(model, batch)
    G = eng.symm(form)
    _, dE = eng.net(G, form)
    P = eng.symm_bwd(dE, form).cpu().numpy()
    orc = oracle_mod.Oracle(model)
    off = batch.atom_off
    nl = eng.nlist()
    dEc = dE.cpu().numpy().astype(np.float64)
    for k in range(batch.nstruc):
        pos, cell, sp = batch.struc(k)
        _, _, Pr = orc.force(pos, cell, sp, dEc[off[k]:off[k + 1], :orc.ld])
        s0, s1 = nl["off"][off[k]], nl["off"][off[k + 1]]
        err = np.sqrt(((P[s0:s1] - Pr) ** 2).sum() / (Pr ** 2).sum())
        assert err <= 1e-5, (k, err)


def test_forms_agree_and_deterministic(cases):
    model, batch = cases["water4"]
    eng = _engine(model, batch)
    E0, F0, _ = eng.energy_forces(1)
    E1, F1, _ = eng.energy_forces(1)
    assert np.array_equal(E0.cpu().numpy(), E1.cpu().numpy())
    assert np.array_equal(F0.cpu().numpy(), F1.cpu().numpy())


def test_capacity_error_is_reported():
    from ann_amd import configs
    from ann_amd.gpu import AnnError, Engine
    model = configs.model_water()
    # 600 atoms inside a sphere of radius < rc: every atom has > 512 neighbours
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 5.0, size=(600, 3))
    b = configs.single(pos, np.eye(3) * 40.0, np.zeros(600, np.int32))
    eng = Engine(model, "cuda:0")
    eng.set_batch(b)
    with pytest.raises(AnnError) as ei:
        eng.neighbors()
    assert ei.value.code == -3


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("name", ["ar256", "water4", "lips3"])
def test_network_parity(cases, oracle_mod, name, form):
    """Atom energies and dE/dG of both network kernels against the FP64
    network on the same (GPU-computed) G."""
    model, batch = cases[name]
    print(f"case {name} form {form}", flush=True)
    eng = _engine(model, batch)
    G = eng.symm(form)
    E, dE = eng.net(G, form)
    E, dE = E.cpu().numpy(), dE.cpu().numpy()
    orc = oracle_mod.Oracle(model)
    Gc = G.cpu().numpy().astype(np.float64)[:, :orc.ld]
    Eo, dEo = orc.net(batch.species, Gc)
    assert np.max(np.abs(E - Eo)) <= 1e-6 * max(1.0, np.abs(Eo).max())
    err = np.sqrt(((dE[:, :orc.ld] - dEo) ** 2).sum() / (dEo ** 2).sum())
    assert err <= 1e-5, err
    assert np.all(dE[:, orc.ld:] == 0)


@pytest.mark.parametrize("tuning", [[2, 2, 2, 0, 0], [4, 4, 4, 0, 1], [1, 1, 1, 1, 0], [2, 1, 1, 0, 0, 1],
                                    [2, 1, 1, 0, 0, 0, 1], [1, 1, 1, 1, 0, 0, 1], [2, 2, 1, 0, 0, 0, 0, 1],
                                    [4, 4, 1, 0, 1, 0, 1, 1]])
@pytest.mark.parametrize("name", ["water4", "lips3"])
def test_tuning_variants_keep_parity(cases, oracle_mod, name, tuning):
    """Every tuning choice (atoms per workgroup, moment/triplet kernels,
    search blocking, work queues, matrix-core against vector-ALU networks,
    the fused E+F kernel) gives the same list bit for bit and E/F within the
    bars."""
    model, batch = cases[name]
    from ann_amd.gpu import Engine
    eng = Engine(model, "cuda:0")
    eng.set_tuning(tuning)
    eng.set_batch(batch)
    eng.neighbors()
    ref = _engine(model, batch).nlist()
    nl = eng.nlist()
    for k in ("off", "idx", "img", "rev"):
        assert np.array_equal(nl[k], ref[k]), k
    E, F, _ = eng.energy_forces(1)
    Eo, Fo, _ = oracle_mod.Oracle(model).compute_batch(batch, nthreads=4)
    assert np.max(np.abs(E.cpu().numpy() - Eo) / batch.natoms) <= 1e-6
    assert np.sqrt(((F.cpu().numpy() - Fo) ** 2).sum() / (Fo ** 2).sum()) <= 1e-5


@pytest.mark.parametrize("name", ["uneven6", "water4"])
def test_persistent_queues_match_grid_dispatch(cases, name):
    """The persistent work-queue dispatch of the moment kernels gives the
    grid dispatch's G, forces and adjoint bit for bit (each atom's wavefront
    does the same work; only the order of atoms over the device differs)."""
    import torch
    from ann_amd.gpu import FAST
    model, batch = cases[name]
    out = []
    for q in (0, 1):
        eng = _engine(model, batch)
        eng.set_tuning([1, 2, 1, 0, 0, q])
        G = eng.symm(FAST)
        E, F, V = eng.energy_forces(FAST)
        out.append((G, E, F, V))
        torch.cuda.synchronize()
    for a, b in zip(*out):
        assert torch.equal(a, b)


def test_autotune_picks_valid_settings(cases, oracle_mod):
    model, batch = cases["water4"]
    eng = _engine(model, batch)
    t = eng.autotune(1)
    assert t[0] in (1, 2, 4) and t[1] in (1, 2, 4) and t[2] in (1, 2, 4) and t[3] in (0, 1) and t[4] in (0, 1)
    assert t[5] in (0, 1) and t[6] in (0, 1) and t[7] in (0, 1)
    E, F, _ = eng.energy_forces(1)
    Eo, Fo, _ = oracle_mod.Oracle(model).compute_batch(batch, nthreads=4)
    assert np.sqrt(((F.cpu().numpy() - Fo) ** 2).sum() / (Fo ** 2).sum()) <= 1e-5


def test_profile_only_times_one_family(cases):
    """Kernel timing restricted to one family records that family only."""
    from ann_amd.gpu import FAST
    model, batch = cases["water4"]
    eng = _engine(model, batch)
    eng.energy_forces(FAST)
    eng.profile(True)
    eng.profile_only("symm_fwd_mom")
    eng.energy_forces(FAST)
    assert eng.profile_get("symm_fwd_mom")[1] == 1
    assert eng.profile_get("net_mfma")[1] == 0
    eng.profile_only("")
    eng.energy_forces(FAST)
    assert eng.profile_get("net_mfma")[1] == 1
    eng.profile(False)
