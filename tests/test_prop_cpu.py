"""CPU checks of the propagation oracle (W3XYP3/W3QCK3, W3KTP3/W3QCK1/W3QCK2) and of the case
generator's maps.  No reference vectors exist for these routines; the oracle is pinned by the
reference's folded constants, the maps' defining properties, and what the mathematics fixes:
the sweeps at CFL 0 (identity) and CFL 1 (a shift by one cell), conservation of the
flux-form sweeps and of the garden-sprinkler averaging, identity of W3KTP3 with both
velocities off, and the sub-step count's closed form."""
import numpy as np
import pytest

from ww3gpu import prop_case as pc

import oracle as O


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def test_constants_match_the_reference_folds():
    # prop_params.hpp:32-35 states the float bits of each folded constant
    want = [0x40B46E2E, 0x3C8EFA35, 0x4AC247EB, 0x47D9038E]
    got = [int(np.float32(O.prop_const(i)).view(np.uint32)) for i in range(4)]
    assert got == want


@pytest.mark.parametrize("global_", [False, True])
def test_w3map3_maps(global_):
    c = pc.make_xyp3(nx=24, ny=18, seed=3, land=0.3, global_=global_)
    st = c["st"]
    nx, ny = st.shape
    nsea = c["nsea"]
    nmx0, nmx1, nmx2, nmy0, nmy1, nmy2, nact, ncent = c["imaps"]
    flat = st.reshape(-1)
    x2 = c["mapx2"][:nmx2]
    assert len(set(x2.tolist())) == nmx2
    for k, v in enumerate(x2):
        ix, iy = (v - 1) // ny + 1, (v - 1) % ny + 1
        ix2 = ix % nx + 1
        a, b = st[ix - 1, iy - 1] == 1, st[ix2 - 1, iy - 1] == 1
        assert 2 <= iy <= ny - 1
        assert (a and b) if k < nmx0 else ((a and not b) if k < nmx1 else (b and not a))
    y2 = c["mapy2"][:nmy2]
    assert len(set(y2.tolist())) == nmy2
    for k, v in enumerate(y2):
        ix, iy = (v - 1) // ny + 1, (v - 1) % ny + 1
        a, b = st[ix - 1, iy - 1] == 1, st[ix - 1, iy] == 1
        assert (a and b) if k < nmy0 else ((a and not b) if k < nmy1 else (b and not a))
    act = c["mapaxy"][:nact]
    assert np.all(flat[act - 1] == 1)
    assert nact == int(np.sum(st[:, 1:ny - 1] == 1))
    cxy = c["mapcxy"]
    assert sorted(cxy.tolist()) == list(range(1, nsea + 1))
    mapsf = c["mapsf"]
    for isea in cxy[:ncent]:
        ix, iy = mapsf[isea - 1], mapsf[nsea + isea - 1]
        assert st[ix - 1, iy - 1] == 1
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                assert st[(ix - 1 + dx) % nx, iy - 1 + dy] >= 1
    # the central list ascends in ISEA, the rest descends
    assert np.all(np.diff(cxy[:ncent]) > 0) and np.all(np.diff(cxy[ncent:]) < 0)


def _periodic_x(nx, ny, cfl_face, q_cells):
    """A field with every cell active and central for an x sweep with closure."""
    q = np.zeros(ny * (nx + 3), np.float32)
    q[ny:ny + nx * ny] = q_cells.reshape(-1)
    cfll = np.zeros(ny * (nx + 1), np.float32)
    cfll[:nx * ny] = cfl_face.reshape(-1)
    trans = np.ones(3 * nx * ny, np.float32)
    allc = np.arange(1, nx * ny + 1, dtype=np.int32)
    return q, cfll, trans, allc


def test_qck3_cfl0_identity_and_cfl1_shift():
    nx, ny = 12, 5
    rng = np.random.default_rng(0)
    cells = rng.uniform(0.5, 1.5, (nx, ny)).astype(np.float32)
    q, cfll, trans, allc = _periodic_x(nx, ny, np.zeros((nx, ny)), cells)
    O.qck3(nx, ny, nx, ny, cfll, trans, q, True, ny, allc, allc.size, allc, allc.size, allc.size,
           allc.size)
    assert np.array_equal(_bits(q[ny:ny + nx * ny]), _bits(cells.reshape(-1)))
    q, cfll, trans, allc = _periodic_x(nx, ny, np.ones((nx, ny)), cells)
    O.qck3(nx, ny, nx, ny, cfll, trans, q, True, ny, allc, allc.size, allc, allc.size, allc.size,
           allc.size)
    got = q[ny:ny + nx * ny].reshape(nx, ny)
    np.testing.assert_allclose(got, np.roll(cells, 1, axis=0), rtol=3e-6)


def test_qck3_and_qck1_conserve_under_closure():
    nx, ny = 16, 6
    rng = np.random.default_rng(1)
    cells = (1.0 + 0.5 * np.sin(np.linspace(0, 2 * np.pi, nx))[:, None] *
             np.ones(ny)[None, :] + 0.1 * rng.random((nx, ny))).astype(np.float32)
    face = rng.uniform(-0.8, 0.8, (nx, ny)).astype(np.float32)
    q, cfll, trans, allc = _periodic_x(nx, ny, face, cells)
    O.qck3(nx, ny, nx, ny, cfll, trans, q, True, ny, allc, allc.size, allc, allc.size, allc.size,
           allc.size)
    s0, s1 = cells.astype(np.float64).sum(), q[ny:ny + nx * ny].astype(np.float64).sum()
    assert abs(s1 - s0) < 1e-5 * s0
    assert not np.array_equal(q[ny:ny + nx * ny], cells.reshape(-1))
    q, cfll, _, allc = _periodic_x(nx, ny, face, cells)
    O.qck1(nx, ny, nx, ny, cfll, q, True, ny, allc, allc.size, allc, allc.size, allc.size, allc.size)
    s2 = q[ny:ny + nx * ny].astype(np.float64).sum()
    assert abs(s2 - s0) < 1e-5 * s0


def test_ktp3_without_velocities_is_identity():
    c = pc.make_ktp3(nsea=40, seed=4, flcth=False, flck=False)
    va, cth, ck = O.ktp3_batch(c)
    assert np.array_equal(_bits(va), _bits(c["va"]))
    assert np.all(cth == 0) and np.all(ck == 0)


def test_ktp3_refraction_conserves_each_wavenumber():
    c = pc.make_ktp3(nsea=40, nk=5, nth=12, seed=5, flcth=True, flck=False, flcur=True, ntimes=3)
    va, cth, _ = O.ktp3_batch(c)
    nspec = c["nspec"]
    a0 = c["va"].reshape(-1, nspec)[c["jsea"]].reshape(-1, 5, 12).astype(np.float64)
    a1 = va.reshape(-1, nspec)[c["jsea"]].reshape(-1, 5, 12).astype(np.float64)
    np.testing.assert_allclose(a1.sum(axis=2), a0.sum(axis=2), rtol=2e-5)
    assert not np.array_equal(a0, a1)
    assert np.all(cth[c["jsea"] - 1] > 0)


def test_xyp3_averaging_conserves_without_propagation():
    c = pc.make_xyp3(nx=30, ny=20, seed=6, land=0.2, global_=True, obstructions=0.0)
    c["ipar"][7] = 0  # FLCX
    c["ipar"][8] = 0  # FLCY
    va = O.xyp3_batch(c)
    nspec, nsea = c["nspec"], c["nsea"]
    mapsf = c["mapsf"]
    st = c["mapsta"][mapsf[2 * nsea:] - 1]
    act = np.nonzero(st >= 1)[0] + 1
    cg = c["cg"].reshape(nsea + 1, -1)
    clats = c["clats"]
    ik = np.arange(nspec) // c["ipar"][4] + 1
    w = (clats[act][:, None] / cg[act][:, ik]).astype(np.float64)
    s0 = (c["va"].reshape(nsea + 1, nspec)[act] * w).sum(axis=0)
    s1 = (va.reshape(nsea + 1, nspec)[act] * w).sum(axis=0)
    np.testing.assert_allclose(s1, s0, rtol=2e-5)
    assert not np.array_equal(va, c["va"])


def test_xyp3_ntloc_closed_form():
    c = pc.make_xyp3(nx=10, ny=8, seed=7)
    nt = O.xyp3_ntloc(c)
    dtg, dtcfl = c["dtg"], float(c["rpar"][0])
    # ITH = 1 of IK = 1 moves at CG0 along x: NTLOC = 1 + INT(DTG/DTCFL)
    assert nt[0] == 1 + int(np.float32(dtg) / np.float32(dtcfl))
    # slower (higher) frequencies need no more sub-steps
    nth = c["ipar"][4]
    per_k = nt.reshape(-1, nth).max(axis=1)
    assert np.all(np.diff(per_k) <= 0)
