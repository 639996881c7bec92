"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

from rift_amd import kp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(rift_\w+)\s*\(", text, re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_headers_declare_entry_points():
    names = declared_functions()
    assert "rift_create" in names and "rift_step" in names and "rift_geometry" in names
    assert len(names) >= 35


def test_library_exports_every_declared_symbol():
    L = kp.load()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    bound = {name for name, _, _ in kp.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_error_reporting_without_gpu_work():
    L = kp.load()
    # a null descriptor is rejected before any device call
    h = ctypes.c_void_p()
    rc = L.rift_create(None, ctypes.byref(h))
    assert rc == -1
    assert b"null" in L.rift_last_error()


def test_geometry_matches_oracle_restatement():
    """The product's C++ InitSimulation geometry equals the oracle's."""
    import numpy as np
    import oracle_lib as O
    from rift_amd import synth
    dem = 30.0 * synth.fractal_dem(32, 0.6, 9)
    h0 = np.maximum(20.0 - dem, 0.0)
    K = 5.0 * np.exp(0.8 * synth.normal(3, dem.size).reshape(dem.shape))
    n = 0.02 + 0.04 * synth.uniform(4, dem.size).reshape(dem.shape)
    geo = kp.geometry(dem, h0=h0, n=n, K=K)
    o = O.Oracle(dem, 10.0, h0=h0, n=n, K=K, psi=0.11, dtheta=0.3)
    bx, by = o.faces()
    np.testing.assert_array_equal(geo["BX"], bx)
    np.testing.assert_array_equal(geo["BY"], by)
    np.testing.assert_array_equal(geo["n"], o.field("n"))
    np.testing.assert_array_equal(geo["K"], o.field("K"))
    # h0 enters as w = BC + h0 on i, j >= 2 (InitGrid_k)
    w = o.field("w")
    bc = o.field("bc")
    np.testing.assert_allclose(o.interior(w - bc), o.interior(geo["h0"]), atol=1e-12)


def test_kappa_matches_oracle():
    import oracle_lib as O
    for dx in (1.0, 5.0, 10.0, 0.3):
        assert kp.kappa(dx, dx) == O.lib().orc_kappa(dx, dx)


def test_binding_constants_match_the_header():
    """Every flag, option and count rift_kp.h defines has the same value in
    the Python binding (a flag bit or option number out of step would select
    another feature without any error)."""
    text = open(os.path.join(ROOT, "include", "rift_kp.h")).read()
    defs = {}
    for m in re.finditer(r"^#define RIFT_(\w+)\s+\((?:1u\s*<<\s*(\d+))\)", text, re.M):
        defs[m.group(1)] = 1 << int(m.group(2))
    for m in re.finditer(r"^#define RIFT_(OPT_\w+|NVARIANTS)\s+(\d+)", text, re.M):
        defs[m.group(1)] = int(m.group(2))
    assert {"TILE_ACCEL", "MASS_BUDGET", "OPT_DT_KERNEL", "NVARIANTS"} <= set(defs)
    for name, value in defs.items():
        assert getattr(kp, name) == value, name
