"""LAMMPS-style pair interface (ann_pair_compute): a periodic domain given as
local atoms plus ghost images with a full neighbour list over both (the data
a LAMMPS pair style receives), against the FP64 oracle on the periodic
structure.  Ghost forces are folded back onto their owners, so the local
forces must equal the periodic forces.

Bars as the E+F parity tests: energy <= 1e-6 eV/atom, forces <= 1e-5
relative RMS, virial <= 1e-4 relative (max component), per-atom energies
<= 1e-6 eV."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SKIN = 0.5


def _domain(pos, cell, rcut):
    """Locals wrapped into the cell, ghosts = every image within rcut of the
    cell, owner map, and the full list of the locals (shuffled rows)."""
    Ri = np.linalg.inv(cell)
    frac = (Ri @ pos.T).T
    frac -= np.floor(frac)
    loc = (cell @ frac.T).T
    # perpendicular widths of the cell: volume / area of the opposite face
    vol = abs(np.linalg.det(cell))
    a = [cell[:, i] for i in range(3)]
    width = [vol / np.linalg.norm(np.cross(a[(i + 1) % 3], a[(i + 2) % 3])) for i in range(3)]
    m = [int(np.ceil(rcut / w)) for w in width]
    w = [rcut / wd for wd in width]
    gx, own = [], []
    for n0 in range(-m[0], m[0] + 1):
        for n1 in range(-m[1], m[1] + 1):
            for n2 in range(-m[2], m[2] + 1):
                if n0 == n1 == n2 == 0:
                    continue
                f = frac + np.array([n0, n1, n2])
                keep = np.all((f > -np.array(w)) & (f < 1 + np.array(w)), axis=1)
                gx.append((cell @ f[keep].T).T)
                own.append(np.nonzero(keep)[0])
    ghosts = np.concatenate(gx)
    owner = np.concatenate(own).astype(np.int32)
    allx = np.concatenate([loc, ghosts])
    rng = np.random.default_rng(3)
    neigh = []
    for i in range(len(loc)):
        d = np.linalg.norm(allx - loc[i], axis=1)
        js = np.nonzero((d < rcut) & (d > 0))[0]
        neigh.append(rng.permutation(js).astype(np.int32))
    return loc, allx, owner, neigh


def _check(orc, E, F, ea, vir, pos, cell, sp, tag):
    Eo, Eao, Fo, Vo = orc.compute(pos, cell, sp)
    N = len(sp)
    de = abs(E - Eo) / N
    rf = float(np.sqrt(((F - Fo) ** 2).sum() / (Fo ** 2).sum()))
    Vl = np.array([Vo[0, 0], Vo[1, 1], Vo[2, 2], Vo[0, 1], Vo[0, 2], Vo[1, 2]])
    rv = float(np.abs(vir - Vl).max() / np.abs(Vo).max())
    evac = np.array([orc.model.evac[s] for s in sp])
    dea = float(np.abs(ea - (Eao + evac)).max())
    print(f"  {tag}: dE/atom {de:.2e} F relRMS {rf:.2e} virial {rv:.2e} eatom {dea:.2e}", flush=True)
    assert de <= 1e-6 and rf <= 1e-5 and rv <= 1e-4 and dea <= 1e-6
    assert abs(ea.sum() - E) <= 1e-9 * max(1.0, abs(E))


@pytest.mark.parametrize("shear", [False, True])
def test_pair_interface_matches_periodic(oracle_mod, shear):
    from ann_amd import configs
    from ann_amd.gpu import Engine
    model = configs.model_water()
    model.evac = model.evac + np.array([0.3, -0.7])      # vacuum energies enter E and eatom
    b = configs.struc_water_md(seed=21, reps=1)
    pos, cell, sp = b.struc(0)
    pos, cell = np.array(pos, float), np.array(cell, float)
    if shear:                               # triclinic: a volume-preserving shear of the whole domain
        S = np.array([[1.0, 0.18, -0.07], [0.0, 1.0, 0.11], [0.0, 0.0, 1.0]])
        pos, cell = (S @ pos.T).T, S @ cell
    rcut = model.rc + SKIN
    loc, allx, owner, neigh = _domain(pos, cell, rcut)
    nlocal = len(loc)
    print(f"pair interface shear={shear}: {nlocal} local, {len(allx) - nlocal} ghost, "
          f"{sum(len(r) for r in neigh)} entries", flush=True)
    eng = Engine(model, "cuda:0")
    E, F, ea, vir = eng.pair_compute(allx, sp, cell, owner, neigh, rcut, nlocal)
    orc = oracle_mod.Oracle(model)
    _check(orc, E, F, ea, vir, loc, cell, sp, "new list")
    # list reuse: locals move (< skin / 2), ghosts with their owners
    rng = np.random.default_rng(4)
    loc2 = loc + rng.uniform(-0.2, 0.2, loc.shape)
    allx2 = np.concatenate([loc2, loc2[owner] + (allx[nlocal:] - loc[owner])])
    E2, F2, ea2, vir2 = eng.pair_compute(allx2, sp, cell, owner, None, rcut, nlocal, new_list=False)
    _check(orc, E2, F2, ea2, vir2, loc2, cell, sp, "reused list")


def test_pair_interface_errors(oracle_mod):
    from ann_amd import configs
    from ann_amd.gpu import AnnError, Engine
    model = configs.model_water()
    b = configs.struc_water_md(seed=22, reps=1)
    pos, cell, sp = b.struc(0)
    rcut = model.rc + SKIN
    loc, allx, owner, neigh = _domain(np.array(pos, float), np.array(cell, float), rcut)
    nlocal = len(loc)
    eng = Engine(model, "cuda:0")
    with pytest.raises(AnnError) as ei:              # reuse before any list
        eng.pair_compute(allx, sp, cell, owner, None, rcut, nlocal, new_list=False)
    assert ei.value.code == -5
    g = next(int(j) for r in neigh for j in r if j >= nlocal) - nlocal
    bad = owner.copy()
    bad[g] = (bad[g] + 1) % nlocal                   # a listed ghost that is not an image of its owner
    with pytest.raises(AnnError) as ei:
        eng.pair_compute(allx, sp, cell, bad, neigh, rcut, nlocal)
    assert ei.value.code == -1 and "lattice image" in str(ei.value)
    half = [r[r < i] for i, r in enumerate(neigh)]   # a half list is not symmetric
    with pytest.raises(AnnError) as ei:
        eng.pair_compute(allx, sp, cell, owner, half, rcut, nlocal)
    assert ei.value.code == -1
    E, F, ea, vir = eng.pair_compute(allx, sp, cell, owner, neigh, rcut, nlocal)
    assert np.isfinite(E)
