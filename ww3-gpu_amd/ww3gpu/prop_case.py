"""Seeded, self-contained PR3/UQ propagation cases in the reference ABI's layouts: a grid with
land, its W3MAP3/W3MAPT maps, group velocities, currents, boundary input and a spectrum field,
for ww_xyp3_batch and ww_ktp3_batch.

The maps are what W3MAP3 (kokkos/src/ww_kokkos/prop_map3.hpp:77-270) and W3MAPT
(prop_mapt.hpp:59-66) compute from MAPSTA and the transparencies; the library takes them as
inputs, as the reference does, so a case builds them the same way the model would.  Sea points
are numbered row by row (IY outer, IX inner) as W3GRID numbers them; MAPSF(ISEA,3) =
IY + (IX-1)*NY.  Rows 1 and NY never hold active points (W3GRID makes their sea points
excluded, MAPSTA = -1); a regional grid's first and last columns are active boundary points
(MAPSTA = 2) when boundary input is on, excluded otherwise.
"""
from __future__ import annotations

import numpy as np

from . import srce_case as sc
from . import synth

f32 = np.float32

XYP3_R = ["DTCFL", "DTH", "XFR", "WDCG", "WDTH"]
XYP3_I = ["NX", "NY", "NSEA", "NK", "NTH", "NSPEC", "ICLOSE", "FLCX", "FLCY", "FLAGLL", "FLCUR"]
KTP3_R = ["DMIN", "CTMAX", "FACHFA"]
KTP3_I = ["NK", "NTH", "NSPEC", "NK2", "FLCTH", "FLCK", "FLCUR"]
ICLOSE_NONE, ICLOSE_SMPL = -1, 2


def _smooth(rng, nx, ny, passes=6):
    f = rng.standard_normal((nx, ny))
    for _ in range(passes):
        f = 0.25 * (np.roll(f, 1, 0) + np.roll(f, -1, 0) + np.roll(f, 1, 1) + np.roll(f, -1, 1))
    return (f - f.mean()) / (f.std() + 1e-12)


def mapsta_grid(nx: int, ny: int, land: float, global_: bool, boundary: bool, rng) -> np.ndarray:
    """MAPSTA(IY,IX) as an (NX, NY) array (index [ix-1, iy-1]): 1 sea, 0 land, 2 boundary,
    -1 excluded."""
    f = _smooth(rng, nx, ny)
    thr = np.quantile(f, land) if land > 0 else -np.inf
    st = np.where(f > thr, 1, 0).astype(np.int32)
    st[:, 0][st[:, 0] != 0] = -1
    st[:, ny - 1][st[:, ny - 1] != 0] = -1
    if not global_:
        for ix in (0, nx - 1):
            col = st[ix]
            col[col != 0] = 2 if boundary else -1
    return st


def sea_points(st: np.ndarray):
    """MAPSF(NSEA,3), Fortran column order flattened, and the sea cells' (ix, iy)."""
    nx, ny = st.shape
    iy, ix = np.nonzero(st.T != 0)  # IY outer, IX inner
    ix1, iy1 = ix + 1, iy + 1
    ixy = iy1 + (ix1 - 1) * ny
    mapsf = np.concatenate([ix1, iy1, ixy]).astype(np.int32)
    return mapsf, ix1, iy1, ixy


def w3map3(st: np.ndarray, mapsf: np.ndarray):
    """MAPX2, MAPY2, MAPAXY, MAPCXY and the counts NMX0..NMY2, NACT, NCENT (prop_map3.hpp:
    77-233), each list in the Fortran's loop order (IX outer, IY inner)."""
    nx, ny = st.shape
    nsea = mapsf.size // 3
    ixy0 = (np.arange(nx)[:, None] * ny)
    iyv = np.arange(1, ny + 1)[None, :]
    s1 = st == 1
    ix2 = (np.arange(nx) + 1) % nx  # IX2 = 1 + MOD(IX, NX)
    inner = np.zeros((nx, ny), bool)
    inner[:, 1:ny - 1] = True
    a, b = s1, s1[ix2]
    secs_x = [a & b & inner, a & ~b & inner, ~a & b & inner]
    lst, cnt = [], []
    for m in secs_x:
        lst.append((ixy0 + iyv)[m])
        cnt.append(sum(x.size for x in lst))
    mapx2 = np.concatenate(lst).astype(np.int32)
    nmx = cnt
    up = np.zeros((nx, ny), bool)
    up[:, :ny - 1] = True
    a = s1
    b = np.zeros_like(s1)
    b[:, :ny - 1] = s1[:, 1:]
    secs_y = [a & b & up, a & ~b & up, ~a & b & up]
    lst, cnt = [], []
    for m in secs_y:
        lst.append((ixy0 + iyv)[m])
        cnt.append(sum(x.size for x in lst))
    mapy2 = np.concatenate(lst).astype(np.int32)
    nmy = cnt
    mapaxy = (ixy0 + iyv)[s1 & inner].astype(np.int32)
    # MAPCXY: central points from the front in ISEA order, the others from the back
    ix = mapsf[:nsea] - 1
    iy = mapsf[nsea:2 * nsea] - 1
    flat = st.reshape(-1)  # (ix, iy) -> ix*ny + iy, the column-major address of MAPSTA(IY,IX)

    def sta(iy_, ix_):
        return flat[np.clip(ix_ * ny + iy_, 0, nx * ny - 1)]

    ix0 = np.where(ix == 0, nx - 1, ix - 1)
    ix2p = np.where(ix == nx - 1, 0, ix + 1)
    stc = st[ix, iy]
    excl = (stc == 2) | (stc < 0)
    cen = np.ones(nsea, bool)
    for dy in (-1, 0, 1):
        for xx in (ix0, ix, ix2p):
            if dy == 0 and xx is ix:
                continue
            cen &= sta(iy + dy, xx) >= 1
    cen &= ~excl
    isea = np.arange(1, nsea + 1, dtype=np.int32)
    mapcxy = np.concatenate([isea[cen], isea[~cen][::-1]]).astype(np.int32)
    counts = np.array(nmx + nmy + [mapaxy.size, int(cen.sum())], np.int32)
    pad = lambda v: np.concatenate([v, np.zeros(nx * ny - v.size, np.int32)]).astype(np.int32)
    return pad(mapx2), pad(mapy2), pad(mapaxy), mapcxy, counts


def spectral_maps(nk: int, nth: int):
    """MAPTH2 ((NK+2)*NTH) and MAPWN2 (NSPEC+NTH) of W3MAP3 section 5 (prop_map3.hpp:238-270)."""
    nk2 = nk + 2
    mapth2 = np.zeros(nk2 * nth, np.int32)
    for ik in range(1, nk + 1):
        for ith in range(1, nth + 1):
            mapth2[ith + (ik - 1) * nth - 1] = (ik + 1) + (ith - 1) * nk2
    mapwn2 = []
    for ik in range(1, nk):
        for ith in range(1, nth + 1):
            mapwn2.append((ik + 1) + (ith - 1) * nk2)
    for ith in range(1, nth + 1):
        mapwn2.append(nk + 1 + (ith - 1) * nk2)
    for ith in range(1, nth + 1):
        mapwn2.append(1 + (ith - 1) * nk2)
    return mapth2, np.array(mapwn2, np.int32)


def w3mapt(atrnx, atrny, mapsf, nxy):
    """MAPTRN at the sea cells (prop_mapt.hpp:59-66): the averaging is off where the least
    transparent side is below 0.95."""
    nsea = mapsf.size // 3
    ixy = mapsf[2 * nsea:] - 1
    ax = atrnx.reshape(3, nxy)
    ay = atrny.reshape(3, nxy)

    def mn(a, b):
        return np.where(a < b, a, b)

    m = mn(mn(ay[2, ixy], ay[0, ixy]), mn(ax[2, ixy], ax[0, ixy]))
    out = np.zeros(nxy, np.int32)
    out[ixy] = (m < f32(0.95)).astype(np.int32)
    return out


def make_xyp3(nx=40, ny=30, nk=6, nth=8, seed=0, land=0.25, global_=False, flagll=True,
              flcur=False, boundary=False, dtg=900.0, cfl=0.8, obstructions=0.08, on=None,
              dsec_day=3600.0, dup_boundary=True):
    """A ww_xyp3_batch case; every array in the ABI's layout (Fortran order, flattened)."""
    rng = np.random.default_rng(seed)
    g = sc.grid_arrays(nk, nth)
    nspec = nk * nth
    st = mapsta_grid(nx, ny, land, global_, boundary, rng)
    mapsf, ix, iy, ixy = sea_points(st)
    nsea = ix.size
    nxy = nx * ny
    # geometry: 0.25 degree lon-lat (or 25 km Cartesian cells) with a mild curvilinear twist
    dlon = 0.25 if flagll else 25000.0
    lat = np.linspace(-40.0, 40.0, ny) if flagll else np.zeros(ny)
    clat_cell = np.cos(np.radians(lat)).astype(np.float32)
    tw = 0.05 * _smooth(rng, nx, ny, 3)
    dpdx = (np.ones((nx, ny)) / dlon * (1 + 0.1 * tw)).astype(np.float32)
    dqdy = (np.ones((nx, ny)) / dlon * (1 - 0.1 * tw)).astype(np.float32)
    dpdy = (0.02 * tw / dlon).astype(np.float32)
    dqdx = (-0.02 * tw / dlon).astype(np.float32)
    gsqrt = (dlon * dlon * (1 + 0.05 * tw)).astype(np.float32)
    fortran = lambda a: np.ascontiguousarray(a.astype(np.float32).reshape(-1))  # (nx, ny) -> (iy fastest)
    clats = np.ones(nsea + 1, np.float32)
    clats[1:] = clat_cell[iy - 1]
    # depths and the group velocity CG(0:NK+1, 0:NSEA)
    depth = np.exp(rng.uniform(np.log(10.0), np.log(5000.0), nsea))
    sig = g["sig0"].astype(np.float64)
    k = synth.wavenumber(sig, depth)
    cgv = synth.group_velocity(sig, k, depth)
    cg = np.ones((nsea + 1, nk + 2), np.float32)
    cg[1:] = cgv.astype(np.float32)
    cx = np.zeros(nsea + 1, np.float32)
    cy = np.zeros(nsea + 1, np.float32)
    if flcur:
        cx[1:] = (0.4 * _smooth(rng, nx, ny)[ix - 1, iy - 1]).astype(np.float32)
        cy[1:] = (0.4 * _smooth(rng, nx, ny)[ix - 1, iy - 1]).astype(np.float32)
    # transparencies: 1 except a few obstructed cells
    atrnx = np.ones((3, nxy), np.float32)
    atrny = np.ones((3, nxy), np.float32)
    obs = rng.random(nxy) < obstructions
    atrnx[:, obs] = rng.uniform(0.3, 1.0, (3, obs.sum())).astype(np.float32)
    obs2 = rng.random(nxy) < obstructions
    atrny[:, obs2] = rng.uniform(0.3, 1.0, (3, obs2.sum())).astype(np.float32)
    mapx2, mapy2, mapaxy, mapcxy, counts = w3map3(st, mapsf)
    maptrn = w3mapt(atrnx.reshape(-1), atrny.reshape(-1), mapsf, nxy)
    mapsta_f = np.ascontiguousarray(st.reshape(-1)).astype(np.int32)  # (ix-1)*ny + iy-1
    # the spectrum field VA(NSPEC, 0:NSEA): smooth in space, positive, a few zeros
    base = np.exp(0.5 * _smooth(rng, nx, ny))[ix - 1, iy - 1]
    spec = rng.gamma(2.0, 1.0, (nspec,)) * 0.01
    va = np.zeros((nsea + 1, nspec), np.float32)
    va[1:] = (base[:, None] * spec[None, :] * (1 + 0.2 * rng.standard_normal((nsea, nspec))))
    va[1:][rng.random((nsea, nspec)) < 0.02] = 0.0
    va = np.maximum(va, 0).astype(np.float32)
    # CFL time step: the fastest component crosses `cfl` cells per DTCFL
    cg0 = f32(f32(0.575) * sc.GRAV) / g["sig0"][1]
    dx_m = dlon * 111000.0 if flagll else dlon
    dtcfl = f32(cfl * dx_m * float(np.min(clat_cell)) / float(cg0) * 0.5)
    rpar = np.array([dtcfl, g["dth"], g["xfr"], 1.0, 1.5], np.float32)
    ipar = np.array([nx, ny, nsea, nk, nth, nspec, ICLOSE_SMPL if global_ else ICLOSE_NONE, 1, 1,
                     1 if flagll else 0, 1 if flcur else 0], np.int32)
    # boundary input at the MAPSTA = 2 points (one listed twice: the last write wins)
    bpts = np.nonzero(st[ix - 1, iy - 1] == 2)[0] + 1
    if boundary and bpts.size:
        isbpi = bpts.astype(np.int32)
        if dup_boundary and isbpi.size > 2:
            isbpi = np.concatenate([isbpi, isbpi[1:2]]).astype(np.int32)
        nbi = isbpi.size
        bbpi0 = (rng.gamma(2.0, 1.0, (nbi + 1, nspec)) * 0.01).astype(np.float32)
        bbpin = (rng.gamma(2.0, 1.0, (nbi + 1, nspec)) * 0.01).astype(np.float32)
        tscal = np.array([dsec_day, 1800.0, 3600.0], np.float32)
        flbpi = 1
    else:
        isbpi = np.zeros(1, np.int32)
        nbi = 0
        bbpi0 = np.zeros((1, nspec), np.float32)
        bbpin = np.zeros((1, nspec), np.float32)
        tscal = np.array([dsec_day, 0.0, 0.0], np.float32)
        flbpi = 0
    onv = np.ones(nspec, np.int32) if on is None else np.asarray(on, np.int32)
    return {
        "rpar": rpar, "ipar": ipar, "imaps": counts, "tscal": tscal, "dtg": float(dtg),
        "mapsf": mapsf, "clats": clats, "sig": g["sig0"].astype(np.float32),
        "ecos": g["ecos"], "esin": g["esin"], "dpdx": fortran(dpdx), "dpdy": fortran(dpdy),
        "dqdx": fortran(dqdx), "dqdy": fortran(dqdy), "gsqrt": fortran(gsqrt),
        "atrnx": atrnx.reshape(-1).copy(), "atrny": atrny.reshape(-1).copy(),
        "mapx2": mapx2, "mapy2": mapy2, "mapaxy": mapaxy, "mapcxy": mapcxy, "maptrn": maptrn,
        "mapsta": mapsta_f, "cg": cg.reshape(-1).copy(), "cx": cx, "cy": cy,
        "flbpi": flbpi, "nbi": nbi, "isbpi": isbpi, "bbpi0": bbpi0.reshape(-1).copy(),
        "bbpin": bbpin.reshape(-1).copy(), "on": onv, "ncol": nsea + 1,
        "va": va.reshape(-1).copy(), "nsea": nsea, "nspec": nspec, "st": st,
    }


XYP3_ARGS = ["rpar", "ipar", "imaps", "tscal", "dtg", "mapsf", "clats", "sig", "ecos", "esin",
             "dpdx", "dpdy", "dqdx", "dqdy", "gsqrt", "atrnx", "atrny", "mapx2", "mapy2",
             "mapaxy", "mapcxy", "maptrn", "mapsta", "cg", "cx", "cy", "flbpi", "nbi", "isbpi",
             "bbpi0", "bbpin", "on", "ncol", "va"]


def make_ktp3(nsea=200, nk=6, nth=8, seed=0, flcth=True, flck=True, flcur=True, itstep=1,
              ntimes=2, dtg=900.0, npts=None, ctmax=0.7):
    """A ww_ktp3_batch case (prop3_shim.cpp:1013-1040's argument table)."""
    rng = np.random.default_rng(seed)
    g = sc.grid_arrays(nk, nth)
    nspec = nk * nth
    nk2 = nk + 2
    th = np.concatenate([g["th"]] * nk + [g["th"]])
    ec = g["ecos"].astype(np.float32)
    es = g["esin"].astype(np.float32)
    ec2 = (ec * ec).astype(np.float32)
    esc = (ec * es).astype(np.float32)
    es2 = (es * es).astype(np.float32)
    mapwn = np.concatenate([np.repeat(np.arange(1, nk + 1), nth), np.full(nth, nk + 1)]).astype(np.int32)
    mapth2, mapwn2 = spectral_maps(nk, nth)
    depth = np.exp(rng.uniform(np.log(3.0), np.log(500.0), nsea))
    sig = g["sig0"].astype(np.float64)
    k = synth.wavenumber(sig, depth)
    cgv = synth.group_velocity(sig, k, depth)
    cg = np.ones((nsea + 1, nk2), np.float32)
    wn = np.ones((nsea + 1, nk2), np.float32)
    cg[1:] = cgv.astype(np.float32)
    wn[1:] = k.astype(np.float32)
    npts = nsea if npts is None else npts
    isea = rng.permutation(np.arange(1, nsea + 1))[:npts].astype(np.int32)
    jsea = isea.copy()
    ps = np.zeros((npts, 10), np.float32)
    ps[:, 0] = rng.uniform(-2e-7, 2e-7, npts)          # CTHG0S
    ps[:, 1] = np.maximum(depth[isea - 1], 1.0)         # MAX(DMIN, DW)
    ps[:, 2] = rng.normal(0, 2e-3, npts)                # DDDX
    ps[:, 3] = rng.normal(0, 2e-3, npts)                # DDDY
    ps[:, 4] = rng.normal(0, 0.3, npts)                 # CX
    ps[:, 5] = rng.normal(0, 0.3, npts)                 # CY
    ps[:, 6:10] = rng.normal(0, 2e-5, (npts, 4))        # DCXDX, DCXDY, DCYDX, DCYDY
    spec = rng.gamma(2.0, 1.0, (nsea + 1, nspec)) * 0.01
    spec[0] = 0
    va = spec.astype(np.float32)
    ntloc = 1
    facth = f32(dtg / (float(g["dth"]) * ntloc))
    fack = f32(dtg / ntloc)
    fachfa = f32(float(g["xfr"]) ** (-5.0 - 2.0))
    rpar = np.array([1.0, ctmax, fachfa], np.float32)
    ipar = np.array([nk, nth, nspec, nk2, int(flcth), int(flck), int(flcur)], np.int32)
    return {
        "rpar": rpar, "ipar": ipar, "sig": g["sig0"].astype(np.float32),
        "dsip": g["dsip0"].astype(np.float32), "ecos": ec, "esin": es, "ec2": ec2, "esc": esc,
        "es2": es2, "mapwn": mapwn, "mapth2": mapth2, "mapwn2": mapwn2, "itstep": int(itstep),
        "ntimes": int(ntimes), "facth": float(facth), "fack": float(fack), "nsea": nsea,
        "cg": cg.reshape(-1).copy(), "wn": wn.reshape(-1).copy(), "npts": npts, "jsea": jsea,
        "isea": isea, "pscal": ps.reshape(-1).copy(), "ncol": nsea + 1, "va": va.reshape(-1).copy(),
        "cflthmax": np.zeros(nsea, np.float32), "cflkmax": np.zeros(nsea, np.float32),
        "nspec": nspec,
    }


KTP3_ARGS = ["rpar", "ipar", "sig", "dsip", "ecos", "esin", "ec2", "esc", "es2", "mapwn",
             "mapth2", "mapwn2", "itstep", "ntimes", "facth", "fack", "nsea", "cg", "wn", "npts",
             "jsea", "isea", "pscal", "ncol", "va", "cflthmax", "cflkmax"]
