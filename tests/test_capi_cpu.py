"""The C ABI library on the CPU: it loads, exports every declared symbol, and its host
INSNL1 reproduces the reference's golden tables.  No device call is made here."""
import os
import re

import numpy as np
import pytest

import oracle as O
from ww3gpu import capi, synth
from ww3gpu.snl1 import insnl1_host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "ww3gpu_c.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ww_\w+|ww3gpu_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = capi.load()
    names = declared_functions()
    assert "ww_snl1" in names and "ww_kokkos_init" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python declarations cover the header exactly
    assert sorted(capi.SIGNATURES) == names


def test_version():
    assert capi.lib().ww3gpu_version().startswith(b"ww3gpu")


def test_insnl1_host_equals_fixture(fixture_snl1):
    f = fixture_snl1
    c = f.cfg
    t = insnl1_host(f.nk, f.nth, c["xfr"], c["dth"], c["lam"], f.sig)
    assert t["structured"]
    assert list(t["dims"][:5]) == list(f.dims5)
    assert np.array_equal(t["scal"].view(np.uint32), f.scal.view(np.uint32))
    assert np.array_equal(t["ipim"], f.ipim)
    assert np.array_equal(t["ic"], f.ic)
    # AF11 follows -O3 WW3 (x^5*x^6); the fixture's -O0 build used x^3*x^8.
    o = O.insnl1(f.nk, f.nth, c["xfr"], c["dth"], c["lam"], f.sig, chain=0)
    assert np.array_equal(t["af11"].view(np.uint32), o.af11.view(np.uint32))


@pytest.mark.parametrize("nk,nth", [(32, 24), (32, 36), (25, 24), (36, 24), (10, 12), (30, 72), (6, 4)])
def test_insnl1_host_equals_oracle(nk, nth):
    cfg, sig = synth.dia_config(nk, nth)
    t = insnl1_host(nk, nth, cfg["xfr"], cfg["dth"], cfg["lam"], sig)
    o = O.insnl1(nk, nth, cfg["xfr"], cfg["dth"], cfg["lam"], sig, chain=0)
    assert t["structured"]
    assert np.array_equal(t["dims"], o.dims)
    assert np.array_equal(t["scal"].view(np.uint32), o.scal.view(np.uint32))
    assert np.array_equal(t["ipim"], o.ipim) and np.array_equal(t["ic"], o.ic)
    assert np.array_equal(t["af11"].view(np.uint32), o.af11.view(np.uint32))


def test_calls_before_init_report_not_initialised():
    L = capi.lib()
    a = np.zeros(4, np.float32)
    L.ww_snl1(987, 1, a.ctypes.data, a.ctypes.data, a.ctypes.data, a.ctypes.data, a.ctypes.data)
    assert L.ww_snl1_last_error() == capi.Status.NOT_INITIALISED
    assert L.ww3gpu_snl1_tables(987, None, None, None, None, None) == capi.Status.NOT_INITIALISED
    assert L.ww3gpu_snl1_lds_bytes(987) == -1


def test_insnl1_host_rejects_bad_shapes():
    L = capi.lib()
    sig = np.ones(4, np.float32)
    assert L.ww3gpu_insnl1(0, 24, 1.1, 0.26, 0.25, sig.ctypes.data, None, None, None, None, None,
                           None) == capi.Status.BAD_SHAPE
    assert L.ww3gpu_insnl1(4, 24, 1.1, 0.26, 0.25, None, None, None, None, None, None,
                           None) == capi.Status.BAD_SHAPE


def test_flag_resolution(monkeypatch):
    L = capi.lib()
    monkeypatch.setenv("WW3GPU_TEST_FLAG_ON", "1")
    monkeypatch.setenv("WW3GPU_TEST_FLAG_OFF", "yes")
    assert L.ww_kokkos_flag(b"WW3GPU_TEST_FLAG_ON") == 1
    assert L.ww_kokkos_flag(b"WW3GPU_TEST_FLAG_OFF") == 0
    assert L.ww_kokkos_flag(None) == 0


_FLAG_PROBE = """
import ctypes, sys
L = ctypes.CDLL(sys.argv[1])
names = sys.argv[2].split(",")
vals = [L.ww_kokkos_flag(n.encode()) for n in names]
slot = L.ww_kokkos_fallback_slot(b"W3SRCE")
L.ww_kokkos_fallback_add(slot)
L.ww_kokkos_fallback_add(slot)
extra = [L.ww_kokkos_fallback_slot(b"W3SRCE"), L.ww_kokkos_fallback_slot(b"X" * 25)]
print(vals, L.ww_kokkos_preset_status(), L.ww_kokkos_strict(), slot, extra)
"""


def _flags_in(env, names):
    import subprocess
    import sys
    e = {k: v for k, v in os.environ.items() if not k.startswith("WW_KOKKOS")}
    e.update(env)
    r = subprocess.run([sys.executable, "-c", _FLAG_PROBE, capi.LIB_PATH, ",".join(names)],
                       env=e, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout.strip(), r.stderr


def test_flag_resolution_rules():
    """kokkos/src/fortran_iface/part_flags.cpp:97-135: a set flag decides alone; unset, a preset
    naming the part gives 1, opt-in parts ignore the master, WW_KOKKOS_ALL=1 gives 2."""
    names = ["WW_KOKKOS_SRCE", "WW_KOKKOS_SNL1", "WW_KOKKOS_XYP3", "WW_KOKKOS_SIN4"]
    out, err = _flags_in({"WW_KOKKOS_ALL": "1", "WW_KOKKOS_SIN4": "0"}, names)
    assert out == "[2, 2, 0, 0] 0 0 0 [0, -1]"
    assert "W3SRCE=2" in err
    out, _ = _flags_in({"WW_KOKKOS_PRESET": "resident", "WW_KOKKOS_STRICT": "1"}, names)
    assert out == "[1, 0, 1, 0] 1 1 0 [0, -1]"
    out, _ = _flags_in({"WW_KOKKOS_PRESET": "resident", "WW_KOKKOS_ALL": "1", "WW_KOKKOS_XYP3": "yes"}, names)
    assert out == "[1, 2, 0, 2] 1 0 0 [0, -1]"
    out, _ = _flags_in({"WW_KOKKOS_PRESET": "nosuch"}, names)
    assert out == "[0, 0, 0, 0] -1 0 0 [0, -1]"


def declared_arity():
    src = open(os.path.join(ROOT, "include", "ww3gpu_c.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for name, params in re.findall(r"\b(ww_\w+|ww3gpu_\w+)\s*\(([^)]*)\)\s*;", src):
        params = params.strip()
        out[name] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_python_signatures_match_header_arity():
    arity = declared_arity()
    for name, (_, args) in capi.SIGNATURES.items():
        assert len(args) == arity[name], f"{name}: {len(args)} ctypes args, header has {arity[name]}"
