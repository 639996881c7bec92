"""Host C++ readers against the reference's own host sources.

oracle/_ref/libref_host.so is compiled from the reference's source.cc and
config.h in place (oracle/Makefile, target `ref`); these tests compare the
product's CSV series reader and keyword parser (include/rift_host.h) with it
on the same files, bit for bit, and the oracle's series interpolation with
Source::InterpolateRate (source.cc:65-73) replayed the way UpdateSource
calls it (simulator.cu:500-538).  Without the reference tree (and so without
the built shim) they are skipped; the file-format round trips at the end need
only the product library.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
from rift_amd import kp

REF_LIB = os.path.join(O.ORACLE_DIR, "_ref", "libref_host.so")


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF_LIB):
        import subprocess
        subprocess.run(["make", "-s", "-C", O.ORACLE_DIR, "ref"], check=False)
    if not os.path.exists(REF_LIB):
        pytest.skip("reference host sources not available: oracle/_ref not built")
    L = C.CDLL(REF_LIB)
    L.ref_source_read.argtypes = [C.c_char_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int]
    L.ref_source_read.restype = C.c_int
    L.ref_source_replay.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.c_int, C.c_float,
                                    C.POINTER(C.c_float)]
    L.ref_config_query.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_double)]
    L.ref_config_string.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
    return L


def _ref_series(ref, path):
    t = np.empty(256, np.float32)
    r = np.empty(256, np.float32)
    n = ref.ref_source_read(path.encode(), t.ctypes.data_as(C.POINTER(C.c_float)),
                            r.ctypes.data_as(C.POINTER(C.c_float)), 256)
    return t[:n].copy(), r[:n].copy()


SERIES = {
    "zero_tail": "# hyetograph, mm/hr\n0,0\n0.25,12.5\n0.5,40\n1.0,0\n",
    "nonzero_tail": "0,10\n0.1,20\n0.3,35.5\n0.7,3\n",
    "spaces_crlf": "# q in cfs\r\n0 , 100\r\n 0.5, 250.25\r\n1.5 ,0\r\n",
    "blank_lines": "0,1\n\n0.2,2\n\n0.4,0\n",
    "one_point": "0.5,7\n",
    "fractions": "0.0333333,1e-3\n0.0666667,2.5e2\n0.1,0.000001\n",
}


@pytest.mark.parametrize("name", sorted(SERIES))
def test_series_reader_matches_reference(ref, tmp_path, name):
    p = tmp_path / f"{name}.csv"
    p.write_bytes(SERIES[name].encode())
    want_t, want_r = _ref_series(ref, str(p))
    got_t, got_r = kp.read_series(str(p))
    assert got_t.tobytes() == want_t.tobytes()
    assert got_r.tobytes() == want_r.tobytes()


@pytest.mark.parametrize("which,unit", [(O.SER_HYETO, 3600.0 * 1000.0), (O.SER_DRAIN, 3600.0 * 1000.0),
                                        (O.SER_HYDRO, 35.3147)])
def test_oracle_series_interpolation_matches_reference(ref, tmp_path, which, unit):
    """Rates at times before, inside, on the knots of and after the series,
    including the stale value re-divided every step past the end (the
    reference keeps interpolated_rate and divides it again)."""
    text = "0.05,10\n0.25,30\n0.5,5\n"   # nonzero tail: a zero point is appended
    p = tmp_path / "s.csv"
    p.write_text(text)
    hours = np.array([0.05, 0.25, 0.5], np.float32)
    vals = np.array([10, 30, 5], np.float32)
    ts = np.concatenate([[0.0, 100.0, 180.0], np.linspace(180.0, 3600.0, 57), [900.0, 1800.0],
                         np.linspace(2200.0, 2600.0, 9)])
    want = np.empty(len(ts), np.float32)
    ref.ref_source_replay(str(p).encode(), ts.ctypes.data_as(C.POINTER(C.c_double)), len(ts),
                          C.c_float(unit), want.ctypes.data_as(C.POINTER(C.c_float)))
    flag = {O.SER_HYETO: O.RAIN_AVG, O.SER_DRAIN: O.DRAIN_AVG, O.SER_HYDRO: O.HYDROGRAPH}[which]
    orc = O.Oracle(np.zeros((8, 8)), 1.0, flags=O.TILE_ACCEL | flag, n_const=0.03)
    orc.set_series_hours(which, hours, vals)
    if which == O.SER_HYDRO:
        orc.set_sources(np.array([5 * orc.nx + 5], np.int32))
    got = np.empty(len(ts), np.float32)
    for k, t in enumerate(ts):
        orc.set_time(float(t), 0.1)
        orc.phase("update_source")
        got[k] = orc.rates()[{O.SER_HYETO: 0, O.SER_DRAIN: 1, O.SER_HYDRO: 2}[which]]
    assert got.tobytes() == want.tobytes()


CONFIG = """; RIFT keyword file
DEM = dem.asc
t0 = 0
tf\t= 7200.5
n = 0.035
n_file = n_map.asc extra words
small = 3e-2
neg = -1
   indented_key   =   12.75   ; trailing comment
tabbed\t\t=\tvalue_with_tab
euler_integration = 1
h_print = 0
source_rect_X = (-94.5,-94.25)
source_pt = (1.5e3,2.5)
output = ./out dir
dotted = 1.2.3
empty_after_comment = 5;=
"""

KEYS = ["DEM", "t0", "tf", "n", "n_file", "small", "neg", "indented_key", "tabbed",
        "euler_integration", "h_print", "source_rect_X", "source_pt", "output", "dotted",
        "empty_after_comment", "missing", "tf\t", ";"]


def test_config_matches_reference(ref, tmp_path):
    p = tmp_path / "run.cfg"
    p.write_text(CONFIG)
    cfg = kp.Config(str(p))
    out = (C.c_double * 2)()
    path = str(p).encode()
    for key in KEYS:
        ref.ref_config_query(path, key.encode(), 0, out)
        assert cfg.has(key) == bool(out[0]), key
        if not cfg.has(key):
            with pytest.raises(kp.RiftError):
                cfg.double(key)
            continue
        buf = C.create_string_buffer(256)
        ref.ref_config_string(path, key.encode(), buf, 256)
        assert cfg.string(key) == buf.value.decode(), key
        ref.ref_config_query(path, key.encode(), 1, out)
        assert cfg.is_number(key) == bool(out[0]), key
    for key in ["t0", "tf", "n", "small", "neg", "indented_key", "euler_integration", "dotted"]:
        ref.ref_config_query(path, key.encode(), 2, out)
        assert cfg.double(key) == out[0], key
    for key in ["euler_integration", "h_print", "t0"]:
        ref.ref_config_query(path, key.encode(), 3, out)
        assert cfg.bool(key) == bool(out[0]), key
    for key in ["source_rect_X", "source_pt"]:
        ref.ref_config_query(path, key.encode(), 4, out)
        assert cfg.complex(key) == complex(out[0], out[1]), key


def test_config_rejects_what_the_reference_rejects(tmp_path):
    """The reference prints and exits on these (config.h:126-150); the
    product reports RIFT_E_ARG with the reason."""
    cases = {
        "no_separator": ("DEM dem.asc\n", "separator"),
        "leading_equals": ("= 5\n", "Bad format"),
        "no_value": ("tf =   \n", "Bad format"),
        "duplicate": ("tf = 1\ntf = 2\n", "unique"),
    }
    for name, (text, msg) in cases.items():
        p = tmp_path / f"{name}.cfg"
        p.write_text(text)
        with pytest.raises(kp.RiftError, match=msg):
            kp.Config(str(p))
    with pytest.raises(kp.RiftError, match="couldn't be found"):
        kp.Config(str(tmp_path / "absent.cfg"))
    p = tmp_path / "typed.cfg"
    p.write_text("flag = true\nx = abc\n")
    cfg = kp.Config(str(p))
    with pytest.raises(kp.RiftError, match="not a valid"):
        cfg.bool("flag")
    with pytest.raises(kp.RiftError, match="not a valid"):
        cfg.double("x")


# --- ESRI-ASCII rasters (io.cc:260-361) -----------------------------------

def test_raster_orientation_and_header(tmp_path):
    """Row 0 of the array is the south row: the first data line of the file
    is the north row."""
    p = tmp_path / "r.asc"
    p.write_text("ncols 3\nnrows 2\nxllcorner 100.5\nyllcorner -20\ncellsize 0.25\n"
                 "NODATA_value -9999\n1 2 3\n4 5 6\n")
    data, hdr = kp.read_raster(str(p))
    np.testing.assert_array_equal(data, [[4, 5, 6], [1, 2, 3]])
    assert hdr == {"xllcorner": 100.5, "yllcorner": -20.0, "cellsize": 0.25, "nodata": -9999.0}


def test_raster_nodata_rules(tmp_path):
    p = tmp_path / "nd.asc"
    p.write_text("ncols 2\nnrows 2\nxllcorner 0\nyllcorner 0\ncellsize 1\nNODATA_value -1\n"
                 "1 -1\n3 4\n")
    with pytest.raises(kp.RiftError, match="NODATA"):
        kp.read_raster(str(p))
    data, _ = kp.read_raster(str(p), allow_nodata=True)
    np.testing.assert_array_equal(data, [[3, 4], [1, -9999.0]])
    q = tmp_path / "short.asc"
    q.write_text("ncols 2\nnrows 2\nxllcorner 0\nyllcorner 0\ncellsize 1\nNODATA_value -1\n1 2\n3\n")
    with pytest.raises(kp.RiftError, match="read error"):
        kp.read_raster(str(q))


def test_write_grid_format_and_round_trip(tmp_path):
    """writeGrid writes the interior (2 ghosts off each side), 15 significant
    digits, north row first, each value followed by a space."""
    rng = np.random.default_rng(5)
    nx, ny = 9, 7
    f = rng.standard_normal((ny, nx)) * 1e3
    p = tmp_path / "out.asc"
    kp.write_grid(str(p), f, 10.25, 20.5, 0.5)
    lines = p.read_text().split("\n")
    assert lines[0] == "ncols         5" and lines[1] == "nrows         3"
    assert lines[2] == "xllcorner     10.25" and lines[3] == "yllcorner     20.5"
    assert lines[4] == "cellsize      0.5" and lines[5] == "NODATA_value  -9999"
    assert lines[6].endswith(" ") and len(lines[6].split()) == 5
    assert lines[6].split()[0] == "%.15g" % f[ny - 3, 2]
    data, hdr = kp.read_raster(str(p))
    np.testing.assert_allclose(data, f[2:ny - 2, 2:nx - 2], rtol=1e-14)


def test_simulation_open_reports_input_errors_before_device_work(tmp_path, monkeypatch):
    """ReadUserParams and the file readers fail before any device call: a
    missing raster and a bad keyword file are reported with their names."""
    monkeypatch.chdir(tmp_path)
    (tmp_path / "a.cfg").write_text("DEM = nowhere.asc\nt0 = 0\ntf = 10\n")
    with pytest.raises(kp.RiftError, match="nowhere.asc"):
        kp.Simulation(str(tmp_path / "a.cfg"))
    (tmp_path / "b.cfg").write_text("DEM = x.asc\nt0 0\n")
    with pytest.raises(kp.RiftError, match="separator"):
        kp.Simulation(str(tmp_path / "b.cfg"))
