"""ST4 / W3SRCE on the CPU: the library's INSIN4 tables against the oracle and against what
the geometry fixes, the oracle's W3SRCE on special cases, and the C ABI's error paths
(no device call is made here)."""
import numpy as np
import pytest

import oracle as O
from ww3gpu import capi, srce_case as sc
from ww3gpu.srce import insin4_tables


def _case(nk=32, nth=36, npts=8, seed=0, **kw):
    return sc.make_case(npts, nk=nk, nth=nth, seed=seed, **kw)


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.int32)


@pytest.mark.parametrize("nk,nth", [(32, 36), (25, 24), (10, 12)])
def test_insin4_tables_equal_oracle(nk, nth):
    c = _case(nk, nth)
    t = insin4_tables(c)
    o = O.insin4_tables(c)
    assert t["nsat"] == o["nsat"]
    assert np.array_equal(_bits(t["satw"]), _bits(o["satw"]))
    assert np.array_equal(t["sati"], o["sati"])
    assert np.array_equal(_bits(t["cumulw"]), _bits(o["cumulw"]))


def test_insin4_tables_geometry():
    """What the geometry fixes: SATINDICES walks ITH-SDSNTH..ITH+SDSNTH with wrap, the
    centre weight is cos(0)**p = 1 and the weights are symmetric and equal cos(j*DTH)**p; CUMULW is the
    phase-speed-difference kernel |c - c2| (law of cosines) * DSIP / (c2/2) * DTH, zero where
    IK2 > IK - DIKCUMUL, and no entry is negative."""
    c = _case(32, 36)
    t = insin4_tables(c)
    g = c["grid"]
    r = c["st4_rpar"]
    nth, nk, nspec = g["nth"], g["nk"], g["nspec"]
    dth = float(g["dth"])
    sdsnth = min(int(round(float(r[sc.ST4_R["SSDSDTH"]]) / (dth * 180.0 / np.pi))), nth // 2 - 1)
    nsat = t["nsat"]
    for ith in range(1, nth + 1):
        idx = t["sati"][ith - 1]
        w = t["satw"][ith - 1]
        for row in range(2 * sdsnth + 1):
            ii = ith - sdsnth + row
            assert idx[row] == (ii - 1) % nth + 1
        assert w[sdsnth] == 1.0
        # symmetric up to the rounding of TH(ITH) - TH(J) in float32
        assert np.allclose(w[:sdsnth][::-1], w[sdsnth + 1:2 * sdsnth + 1], rtol=4e-6, atol=1e-6)
        assert np.all(idx[2 * sdsnth + 1:] == 1) and np.all(w[2 * sdsnth + 1:] == 0.0)
        expect = np.cos(np.arange(-sdsnth, sdsnth + 1) * dth) ** float(r[sc.ST4_R["SSDSCOS"]])
        assert np.allclose(w[:2 * sdsnth + 1], expect, rtol=2e-6, atol=2e-7)
    assert nsat == 2 * (nth // 2 - 1) + 1
    cw = t["cumulw"]  # [IS-1][IS2-1]
    dik = int(np.floor(float(r[sc.ST4_R["SSDSBRF1"]]) / (float(g["xfr"]) - 1.0) + 0.5))
    sig = g["sig0"].astype(np.float64)
    dsip = g["dsip0"].astype(np.float64)
    rng = np.random.default_rng(3)
    for _ in range(400):
        ik, ik2 = rng.integers(1, nk + 1, 2)
        ith, ith2 = rng.integers(1, nth + 1, 2)
        v = cw[ith - 1 + (ik - 1) * nth, ith2 - 1 + (ik2 - 1) * nth]
        if ik2 > ik - dik:
            assert v == 0.0
            continue
        cc, c2 = 9.806 / sig[ik], 9.806 / sig[ik2]
        ang = (ith2 - ith) * dth
        ref = np.sqrt(cc * cc + c2 * c2 - 2 * cc * c2 * np.cos(ang)) * dsip[ik2] / (0.5 * c2) * dth
        assert v >= 0.0 and np.isclose(v, ref, rtol=2e-5, atol=1e-9)
    assert cw.shape == (nspec, nspec)


def test_insin4_tables_isotropic_and_no_cumulative():
    c = _case(10, 12)
    c["st4_rpar"][sc.ST4_R["SSDSDTH"]] = 180.0
    c["st4_rpar"][sc.ST4_R["SSDSC"] + 2] = 0.5
    t = insin4_tables(c)
    assert np.all(t["satw"] == 1.0) and np.all(t["sati"] == 1)
    assert not np.any(t["cumulw"])


def test_insin4_tables_rejects_bad_packets():
    L = capi.lib()
    c = _case(10, 12)
    g = c["grid"]
    r, i = c["st4_rpar"], c["st4_ipar"]
    out = np.zeros(1 << 16, np.float32)
    p = capi.ptr
    assert L.ww3gpu_insin4_tables(p(r), r.size - 1, p(i), i.size, p(g["sig0"]), p(g["dsip0"]), p(g["ecos"]),
                                  p(g["th"]), p(out), p(out), p(out)) == capi.Status.BAD_SHAPE
    assert L.ww3gpu_insin4_tables(p(r), r.size, p(i), i.size, None, p(g["dsip0"]), p(g["ecos"]),
                                  p(g["th"]), p(out), p(out), p(out)) == capi.Status.BAD_SHAPE


def test_srce_calls_without_runtime_report_not_initialised():
    L = capi.lib()
    c = _case(10, 12)
    r, i = c["st4_rpar"], c["st4_ipar"]
    p = capi.ptr
    # this container has no device, so the runtime never becomes ready
    assert L.ww_st4_init(5, p(r), r.size, p(i), i.size) == capi.Status.NOT_INITIALISED
    assert L.ww_st4_commit(5) == capi.Status.NOT_INITIALISED
    one = np.zeros(4096, np.float32)
    L.ww_srce_batch(5, 1, 0, 0, 0, 300.0, 1, *([p(one)] * 9))
    assert L.ww_srce_last_error() == capi.Status.NOT_INITIALISED
    assert L.ww3gpu_srce_substeps(5) == -1
    assert L.ww_st4_enabled(0) == 0


def test_packet_layouts_match_the_header(tmp_path):
    """The Python packet maps (srce_case) and the C header's enums give every slot the same
    index: a C program that includes the header prints the enum values."""
    import os
    import subprocess
    inc = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "include")
    names = ([("WW_ST4_R_" + k, v) for k, v in sc.ST4_R.items()] +
             [("WW_ST4_I_" + k, v) for k, v in sc.ST4_I.items()] +
             [("WW_SRCE_R_" + k, v) for k, v in sc.SRCE_R.items()] +
             [("WW_SRCE_I_" + k, v) for k, v in sc.SRCE_I.items()] +
             [("WW_SRCE_IN_" + k, i) for i, k in enumerate(sc.PIN)] +
             [("WW_SRCE_SC_" + k, i) for i, k in enumerate(sc.PSC)] +
             [("WW_ST4_NRPAR", sc.ST4_NRPAR), ("WW_ST4_NIPAR", sc.ST4_NIPAR),
              ("WW_SRCE_NRPAR", sc.SRCE_NRPAR), ("WW_SRCE_NIPAR", sc.SRCE_NIPAR),
              ("WW_SRCE_NPIN", len(sc.PIN)), ("WW_SRCE_NPSC", len(sc.PSC))])
    src = '#include <stdio.h>\n#include "ww3gpu_c.h"\nint main(void){\n'
    src += "".join(f'printf("%d\\n", (int){n});\n' for n, _ in names) + "return 0;}\n"
    (tmp_path / "p.c").write_text(src)
    exe = tmp_path / "p"
    subprocess.run(["gcc", "-I", inc, str(tmp_path / "p.c"), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert got == [v for _, v in names]
    assert sc.ST4_NRPAR == 84 and sc.SRCE_NRPAR == 29


# ------------------------------------------------------------------------------------
# The oracle's W3SRCE on cases whose answer the equations fix
# ------------------------------------------------------------------------------------
def _oracle(c):
    return O.srce_batch(c, O.insin4_tables(c), nthreads=4)


def test_oracle_zero_spectrum_calm_wind_stays_zero():
    """No energy and no wind: every source term is zero, so is the spectrum after the step,
    and no NaN appears (the U10 = 0 path takes FHIGH = +Inf on the first sub-step)."""
    c = _case(25, 24, npts=4)
    c["spec"][:] = 0.0
    c["pin"][:, sc.PIN.index("U10ABS")] = 0.0
    c["psc"][:] = 0.0
    out = _oracle(c)
    assert np.all(out["spec"] == 0.0)
    assert np.all(np.isfinite(out["psc"][:, :sc.PSC.index("FPI")]))


def test_oracle_it0_restores_the_spectrum():
    """IT = 0 (the first call of a run): W3SRCE computes the fluxes but hands back the
    spectrum it was given, clipped at 0."""
    c = _case(25, 24, npts=6)
    c["flags"]["it"] = 0
    out = _oracle(c)
    assert np.array_equal(_bits(out["spec"]), _bits(np.maximum(c["spec"], 0.0)))


def test_oracle_ice_without_dispersion_has_no_ice_flux():
    """Ice fraction > 0 with INFLAGS2(4) and no ice dispersion: ATT = 1, so the ice block
    changes no value and its fluxes are zero."""
    c = _case(25, 24, npts=6)
    c["flags"]["inflags24"] = 1
    c["pin"][:, sc.PIN.index("ICE")] = 0.3
    out = _oracle(c)
    assert np.all(out["psc"][:, sc.PSC.index("PHICE")] == 0.0)
    assert np.all(out["tauice"] == 0.0)


def test_oracle_energy_budget_closes():
    """PHIOC = (PHIAW + E_before - E_after - PHIBBL) * DWAT*GRAV / DTG: the energy flux to the
    ocean is what the wind put in minus what the spectrum gained, minus the bottom loss.
    Checked against the energies of the spectra before and after, in float64, so a dropped
    term or a wrong Jacobian (DDEN/CG) in the finish block fails it."""
    c = _case(32, 36, npts=12, seed=5)
    out = _oracle(c)
    g = c["grid"]
    dden = g["dden"].astype(np.float64)
    nk, nth = g["nk"], g["nth"]
    P = sc.PSC.index
    dg = 1000.0 * 9.806
    for p in range(c["npts"]):
        cg = c["cg1"][p].astype(np.float64)
        e0 = (c["spec"][p].reshape(nk, nth).astype(np.float64).sum(1) * dden / cg).sum()
        e1 = (out["spec"][p].reshape(nk, nth).astype(np.float64).sum(1) * dden / cg).sum()
        phiaw = out["psc"][p, P("PHIAW")] * 300.0 / dg
        phibbl = out["psc"][p, P("PHIBBL")] * 300.0 / dg
        ref = (phiaw + (e0 - e1) - phibbl) * dg / 300.0
        got = float(out["psc"][p, P("PHIOC")])
        assert np.isclose(got, ref, rtol=2e-3, atol=1e-4 * max(1.0, abs(ref))), (p, got, ref)


def test_group_lanes_control():
    L = capi.lib()
    for bad in (-1, 8, 48, 128):
        assert L.ww3gpu_srce_group_lanes(bad) == capi.Status.BAD_SHAPE
    assert L.ww3gpu_srce_group_lanes(0) == capi.Status.OK


def test_width_thresholds_control():
    """ww3gpu_srce_width_thresholds: 0, 0 restores the defaults; negative values or
    wide_below > half_below are BAD_SHAPE (host-only; no device needed)."""
    L = capi.lib()
    assert L.ww3gpu_srce_width_thresholds(-1, 0) == capi.Status.BAD_SHAPE
    assert L.ww3gpu_srce_width_thresholds(100, 200) == capi.Status.BAD_SHAPE
    assert L.ww3gpu_srce_width_thresholds(5000, 2500) == capi.Status.OK
    assert L.ww3gpu_srce_width_thresholds(0, 0) == capi.Status.OK
