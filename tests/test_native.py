"""The native libraries on the CPU: they build, load and export every entry
point their headers declare, and the host layer's deterministic pieces (grid,
RK coefficients, initial field, NPY format, diagnostics) agree with the oracle
and with numpy.  No GPU call is made here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def declared(header, pattern=r"^\s*(?:extern\s+)?[\w\s\*]+?\b(\w+)\s*\(", prefix=None):
    names = []
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    for line in text.splitlines():
        if line.strip().startswith("#") or "typedef" in line:
            continue
        m = re.match(pattern, line)
        if m and (prefix is None or m.group(1).startswith(prefix)):
            names.append(m.group(1))
    return names


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if l.strip()}


def test_hip_library_exports_every_c_abi_symbol(sns):
    names = declared("sns_hip.h", prefix="snsg_")
    assert len(names) >= 25
    syms = exported(sns.HIP_LIB)
    missing = [n for n in names if n not in syms]
    assert not missing, missing
    L = sns.hip_lib()
    for n in names:
        assert hasattr(L, n)


@pytest.mark.parametrize("header", ["param.h", "fluid.h", "temperature.h", "parallel.h",
                                    "fileio.h", "save.h", "logging.h", "common.h",
                                    "statistics.h"])
def test_host_library_exports_reference_call_surface(sns, header):
    names = declared(header, pattern=r"^\s*extern\s+[\w\s\*]+?\b(\w+)\s*\(")
    assert names
    syms = exported(sns.HOST_LIB)
    missing = [n for n in names if n not in syms]
    assert not missing, missing


def test_driver_binary_present(sns):
    assert os.access(sns.DRIVER, os.X_OK)


@pytest.mark.parametrize("itot,jtot,stretch", [(32, 64, 3.0), (61, 30, 2.0), (128, 256, 64.0)])
def test_host_grid_matches_oracle_bitwise(sns, itot, jtot, stretch):
    g = sns.grid(itot, jtot, 2.0, stretch)
    o = po.Oracle(itot, jtot, stretch=stretch)
    for n in ["xf", "xc", "dxf", "dxc", "yf", "yc"]:
        assert np.array_equal(g[n].view(np.uint64), getattr(o, n).view(np.uint64)), n
    assert g["dy"] == o.dy


def test_host_rk_coefficients_match_oracle(sns):
    a, b, g = sns.rk_coefs()
    oa, ob, og = po.Oracle(8, 8).rk()
    assert list(a) == oa and list(b) == ob and list(g) == og


@pytest.mark.parametrize("linear", [False, True])
def test_host_initial_temperature_matches_oracle(sns, linear):
    o = po.Oracle(40, 24, stretch=3.0)
    o.init_temperature(11, 1e-3, linear=linear)
    t = sns.initial_temperature(40, 24, o.xc, 11, 1e-3, linear)
    assert np.array_equal(t[1:-1, 1:-1], o.f["temp"][1:-1, 1:-1])


def test_npy_files_round_trip_with_numpy(sns, tmp_path):
    L = sns.host_lib()
    L.fileio_w_1d_serial.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t, C.c_size_t,
                                     C.c_void_p]
    L.fileio_w_0d_serial.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t, C.c_void_p]
    L.fileio_r_0d_serial.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t, C.c_void_p]
    d = str(tmp_path).encode()
    x = np.linspace(0, 1, 37)
    L.fileio_w_1d_serial(d, b"xf", b"'float64'", 8, x.size, x.ctypes.data)
    y = np.load(os.path.join(str(tmp_path), "xf.npy"))
    assert np.array_equal(x, y)
    raw = open(os.path.join(str(tmp_path), "xf.npy"), "rb").read()
    hlen = raw[8] | (raw[9] << 8)
    assert (10 + hlen) % 64 == 0 and raw[10 + hlen - 1:10 + hlen] == b"\n"
    step = C.c_int(1234)
    L.fileio_w_0d_serial(d, b"step", b"'int32'", 4, C.addressof(step))
    assert int(np.load(os.path.join(str(tmp_path), "step.npy"))) == 1234
    np.save(os.path.join(str(tmp_path), "time.npy"), np.float64(3.25))
    t = C.c_double(0.0)
    assert L.fileio_r_0d_serial(d, b"time", 8, C.addressof(t)) == 0
    assert t.value == 3.25


def test_host_diagnostics_match_oracle(sns):
    o = po.Oracle(24, 32, stretch=3.0, Ra=1e7, Pr=2.0)
    rng = np.random.default_rng(4)
    for n in ["ux", "uy", "temp"]:
        o.f[n][...] = rng.uniform(-1, 1, o.f[n].shape)
    o.apply_boundaries()
    L = sns.host_lib()
    dp = C.POINTER(C.c_double)
    p = lambda a: a.ctypes.data_as(dp)
    out = np.zeros(11)
    L.sns_logging_compute_raw.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                          C.c_double, dp, dp, C.c_double, dp, dp, dp, dp]
    L.sns_logging_compute_raw(24, 32, 1.0, 2.0, 1e7, 2.0, p(o.dxf), p(o.dxc), o.dy,
                              p(o.f["ux"]), p(o.f["uy"]), p(o.f["temp"]), p(out))
    expect = np.concatenate([o.divergence(), o.momentum(), o.energy(), o.nusselt()])
    assert np.array_equal(out, expect)
