"""MI355X (gfx950) implementation of the SimpleNavierStokesSolver time-step hot path.

Native parts (built in-tree by ``simplenavierstokessolver_amd.build``):

* ``lib/libsns_hip.so`` -- hand-written HIP kernels and the C-ABI declared in
  ``include/sns_hip.h``;
* ``lib/libsns.so`` -- the C host layer with the reference's call surface
  (``param_*``, ``fluid_*``, ``temperature_*``, ``fileio_*``);
* ``lib/sns_run`` -- the driver with the reference's time loop.

This module is a ctypes front end used by the tests and the benchmark.  It
never falls back to a CPU path: if the native library is missing, loading
fails with an error.
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_PKG, "lib")
HIP_LIB = os.path.join(LIB_DIR, "libsns_hip.so")
HOST_LIB = os.path.join(LIB_DIR, "libsns.so")
DRIVER = os.path.join(LIB_DIR, "sns_run")

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class SnsgConfig(C.Structure):
    _fields_ = [("itot", C.c_int), ("jtot", C.c_int), ("ly", C.c_double),
                ("Ra", C.c_double), ("Pr", C.c_double),
                ("with_temperature", C.c_int), ("with_thermal_forcing", C.c_int),
                ("implicitx", C.c_int), ("implicity", C.c_int),
                ("alpha", C.c_double * 3), ("beta", C.c_double * 3), ("gamma", C.c_double * 3)]


class SnsgState(C.Structure):
    _fields_ = [(n, _vp) for n in ["ux", "uy", "p", "psi", "temp", "temp_alt",
                                   "srcuxa", "srcuya", "srctempa", "ux_star", "uy_star",
                                   "p_alt"]] + [
        ("ld_ux", C.c_long), ("ld_c", C.c_long)]


class SnsgJob(C.Structure):
    _fields_ = [("cfg", SnsgConfig), ("dxf", _dp), ("dxc", _dp), ("seed", C.c_uint64),
                ("amplitude", C.c_double), ("nsteps", C.c_int), ("keep", C.c_int)]


class SnsgServerStats(C.Structure):
    _fields_ = [(n, C.c_long) for n in ("submitted", "admitted", "completed", "active",
                                        "queued", "active_peak", "steps")] + [
        ("cell_steps", C.c_double), ("busy_s", C.c_double), ("p50_ms", C.c_double),
        ("p99_ms", C.c_double), ("max_ms", C.c_double), ("mem_used", C.c_size_t),
        ("mem_peak", C.c_size_t), ("lanes", C.c_int)]


class SnsgLaunch(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("fft_n1", "ysolve_chunk", "ysolve_cols",
                                       "ysolve_recompute", "split_corr")]


LAUNCH_FIELDS = tuple(n for n, _ in SnsgLaunch._fields_)
# launch-shape table written by `python -m simplenavierstokessolver_amd.autotune`
TUNING_FILE = os.path.join(_PKG, "tuning", "launch_gfx950.json")


# field ids of snsg_state_upload / snsg_state_download (include/sns_hip.h)
STATE_FIELDS = {"ux": 0, "uy": 1, "p": 2, "psi": 3, "temp": 4, "ux_star": 5, "uy_star": 6,
                "srcuxa": 7, "srcuya": 8, "srctempa": 9}


class GpuError(RuntimeError):
    pass


_hip = None
_host = None


def _load_hip():
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise GpuError("native library %s is missing: run simplenavierstokessolver_amd.build"
                           % HIP_LIB)
        L = C.CDLL(HIP_LIB)
        L.snsg_last_error.restype = C.c_char_p
        L.snsg_create.argtypes = [C.c_int, C.POINTER(SnsgConfig), _dp, _dp, C.POINTER(_vp)]
        L.snsg_destroy.argtypes = [_vp]
        L.snsg_set_treatment.argtypes = [_vp, C.c_int, C.c_int]
        L.snsg_set_dt.argtypes = [_vp, C.c_double]
        L.snsg_synchronize.argtypes = [_vp]
        L.snsg_device_count.argtypes = [C.POINTER(C.c_int)]
        L.snsg_malloc.argtypes = [_vp, C.c_size_t, C.POINTER(_vp)]
        L.snsg_copy.argtypes = [_vp, _vp, _vp, C.c_size_t]
        L.snsg_copy_shaped.argtypes = [_vp, _vp, _vp, C.c_size_t, C.c_int, C.c_int]
        L.snsg_free.argtypes = [_vp, _vp]
        L.snsg_upload.argtypes = [_vp, _vp, _vp, C.c_size_t]
        L.snsg_download.argtypes = [_vp, _vp, _vp, C.c_size_t]
        for n in ["snsg_fluid_update_boundaries_ux", "snsg_fluid_update_boundaries_uy",
                  "snsg_fluid_update_boundaries_p", "snsg_temperature_update_boundaries_temp"]:
            getattr(L, n).argtypes = [_vp, _vp]
        L.snsg_temperature_compute_force.argtypes = [_vp, _vp, _vp]
        L.snsg_temperature_update_temp.argtypes = [_vp, C.c_int] + [_vp] * 6
        L.snsg_fluid_update_velocity.argtypes = [_vp, C.c_int] + [_vp] * 10
        L.snsg_fluid_compute_potential.argtypes = [_vp, C.c_int, _vp, _vp, _vp]
        L.snsg_fluid_correct_velocity.argtypes = [_vp, C.c_int, _vp, _vp, _vp]
        L.snsg_fluid_update_pressure.argtypes = [_vp, C.c_int, _vp, _vp]
        L.snsg_param_decide_dt.argtypes = [_vp, _vp, _vp, _dp]
        L.snsg_transpose.argtypes = [_vp, C.c_int, C.c_int, C.c_size_t, _vp, _vp]
        L.snsg_state_alloc.argtypes = [_vp, C.POINTER(SnsgState)]
        L.snsg_state_upload.argtypes = [_vp, C.POINTER(SnsgState), C.c_int, _vp]
        L.snsg_state_download.argtypes = [_vp, C.POINTER(SnsgState), C.c_int, _vp]
        L.snsg_state_decide_dt.argtypes = [_vp, C.POINTER(SnsgState), _dp]
        L.snsg_diagnostics.argtypes = [_vp, _vp, _vp, _vp, _dp]
        L.snsg_state_diagnostics.argtypes = [_vp, C.POINTER(SnsgState), _dp]
        L.snsg_diagnostics_noalloc.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.snsg_statistics_collect.argtypes = [_vp] * 10
        L.snsg_state_init_temperature.argtypes = [_vp, C.POINTER(SnsgState), C.c_uint64,
                                                  C.c_double]
        L.snsg_instance_bytes.argtypes = [C.POINTER(SnsgConfig), C.POINTER(C.c_size_t)]
        L.snsg_memory.argtypes = [_vp, C.POINTER(C.c_size_t)]
        L.snsg_uniform_grid.argtypes = [_vp, C.POINTER(C.c_int)]
        L.snsg_launch_check.argtypes = [C.c_int, C.c_int, C.POINTER(SnsgLaunch)]
        L.snsg_set_launch.argtypes = [_vp, C.POINTER(SnsgLaunch)]
        L.snsg_get_launch.argtypes = [_vp, C.POINTER(SnsgLaunch)]
        L.snsg_tuning_add.argtypes = [C.c_int, C.c_int, C.POINTER(SnsgLaunch)]
        L.snsg_tuning_clear.argtypes = []
        L.snsg_server_create.argtypes = [C.c_int, C.c_size_t, C.c_int, C.POINTER(_vp)]
        L.snsg_server_submit.argtypes = [_vp, C.POINTER(SnsgJob), C.POINTER(C.c_int)]
        L.snsg_server_wait.argtypes = [_vp, C.c_int, C.c_double]
        L.snsg_server_drain.argtypes = [_vp, C.c_double]
        L.snsg_server_status.argtypes = [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.snsg_server_download.argtypes = [_vp, C.c_int, C.c_int, _vp]
        L.snsg_server_job_time.argtypes = [_vp, C.c_int, _dp, _dp]
        L.snsg_server_release.argtypes = [_vp, C.c_int]
        L.snsg_server_stats_get.argtypes = [_vp, C.POINTER(SnsgServerStats)]
        L.snsg_server_reset_stats.argtypes = [_vp]
        L.snsg_server_destroy.argtypes = [_vp]
        L.snsg_batch_create.argtypes = [C.c_int, C.POINTER(_vp), C.POINTER(C.POINTER(SnsgState)),
                                        C.c_int, C.POINTER(_vp)]
        L.snsg_batch_advance.argtypes = [_vp, C.c_int]
        L.snsg_batch_synchronize.argtypes = [_vp]
        L.snsg_batch_classes.argtypes = [_vp, C.POINTER(C.c_int)]
        L.snsg_batch_destroy.argtypes = [_vp]
        L.snsg_state_statistics_collect.argtypes = [_vp, C.POINTER(SnsgState)] + [_vp] * 6
        L.snsg_state_free.argtypes = [_vp, C.POINTER(SnsgState)]
        L.snsg_integrate.argtypes = [_vp, C.POINTER(SnsgState)]
        L.snsg_advance.argtypes = [_vp, C.POINTER(SnsgState), C.c_int]
        L.snsg_get_time.argtypes = [_vp, _dp, _dp]
        L.snsg_set_time.argtypes = [_vp, C.c_double]
        L.snsg_check_division.argtypes = [_vp, C.c_int] + [_vp] * 5
        L.snsg_timing_enable.argtypes = [_vp, C.c_char_p]
        L.snsg_timing_disable.argtypes = [_vp]
        L.snsg_timing_read.argtypes = [_vp, C.POINTER(C.c_int), _dp]
        _hip = L
    return _hip


def _load_host():
    global _host
    if _host is None:
        _load_hip()
        if not os.path.exists(HOST_LIB):
            raise GpuError("native library %s is missing: run simplenavierstokessolver_amd.build"
                           % HOST_LIB)
        L = C.CDLL(HOST_LIB)
        L.sns_set_coordinate.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                         _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.sns_rk_coefs.argtypes = [_dp, _dp, _dp]
        L.sns_initial_temperature.argtypes = [C.c_int, C.c_int, _dp, C.c_uint64, C.c_double,
                                              C.c_int, _dp]
        _host = L
    return _host


def hip_lib():
    return _load_hip()


def host_lib():
    return _load_host()


def _check(rc, what):
    if rc != 0:
        raise GpuError("%s failed (%d): %s" % (what, rc, _load_hip().snsg_last_error().decode()))


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def device_count():
    n = C.c_int(0)
    _check(_load_hip().snsg_device_count(C.byref(n)), "snsg_device_count")
    return n.value


def grid(itot, jtot, ly=2.0, stretch=3.0, lx=1.0):
    """Coordinates from the host layer (src/param/init.c:124-195 restated)."""
    L = _load_host()
    g = {"xf": np.zeros(itot + 1), "xc": np.zeros(itot + 2), "dxf": np.zeros(itot),
         "dxc": np.zeros(itot + 1), "yf": np.zeros(jtot + 1), "yc": np.zeros(jtot)}
    dy = C.c_double(0.0)
    L.sns_set_coordinate(itot, jtot, lx, ly, stretch, _p(g["xf"]), _p(g["xc"]), _p(g["dxf"]),
                         _p(g["dxc"]), _p(g["yf"]), _p(g["yc"]), C.byref(dy))
    g["dy"] = dy.value
    return g


def tanh_grid(itot, jtot, ly=2.0, stretch=3.0, lx=1.0):
    """Wall-refined x coordinates from a hyperbolic tangent, an option with
    no counterpart in the reference (whose stretched grid is the clipped
    Chebyshev one, src/param/init.c:144-165):
        XF(i) = lx/2 (1 + tanh(s (2 (i-1)/itot - 1)) / tanh(s)),  i = 1..itot+1,
    s = stretch, with XF(1) = 0 and XF(itot+1) = lx exactly; the cell centres,
    spacings and y coordinates follow from XF as in the reference
    (src/param/init.c:170-187).  The instance takes them through
    Instance(..., coords=tanh_grid(...))."""
    g = grid(itot, jtot, ly, float(itot), lx)  # y as the reference; x replaced
    xi = np.arange(itot + 1) / itot
    xf = 0.5 * lx * (1.0 + np.tanh(stretch * (2.0 * xi - 1.0)) / np.tanh(stretch))
    xf[0], xf[itot] = 0.0, lx
    xc = np.empty(itot + 2)
    xc[0], xc[itot + 1] = 0.0, lx
    xc[1:itot + 1] = 0.5 * (xf[:-1] + xf[1:])
    g["xf"], g["xc"] = xf, xc
    g["dxf"] = xf[1:] - xf[:-1]
    g["dxc"] = xc[1:] - xc[:-1]
    return g


def rk_coefs():
    a, b, c = np.zeros(3), np.zeros(3), np.zeros(3)
    _load_host().sns_rk_coefs(_p(a), _p(b), _p(c))
    return a, b, c


def initial_temperature(itot, jtot, xc, seed, amplitude, linear=False):
    """Interior of the seeded initial temperature in the (jtot+2, itot+2) layout."""
    t = np.zeros((jtot + 2, itot + 2))
    _load_host().sns_initial_temperature(itot, jtot, _p(np.ascontiguousarray(xc)), seed,
                                         amplitude, int(linear), _p(t))
    return t


FIELD_SHAPES = {
    "ux": lambda i, j: (j + 2, i + 1), "uy": lambda i, j: (j + 2, i + 2),
    "p": lambda i, j: (j + 2, i + 2), "psi": lambda i, j: (j + 2, i + 2),
    "temp": lambda i, j: (j + 2, i + 2), "tempforcex": lambda i, j: (j, i - 1),
    "srcuxa": lambda i, j: (j, i - 1), "srcuxb": lambda i, j: (j, i - 1),
    "srcuxg": lambda i, j: (j, i - 1), "srcuya": lambda i, j: (j, i),
    "srcuyb": lambda i, j: (j, i), "srcuyg": lambda i, j: (j, i),
    "srctempa": lambda i, j: (j, i), "srctempb": lambda i, j: (j, i),
    "srctempg": lambda i, j: (j, i),
    "ux1": lambda i, j: (j + 2, i + 1), "ux2": lambda i, j: (j + 2, i + 1),
    "uy1": lambda i, j: (j + 2, i + 2), "uy2": lambda i, j: (j + 2, i + 2),
    "temp1": lambda i, j: (j + 2, i + 2), "temp2": lambda i, j: (j + 2, i + 2),
}


def make_config(itot, jtot, ly=2.0, Ra=1e8, Pr=1.0, implicitx=0, implicity=0,
                with_temperature=True, with_thermal_forcing=True):
    """snsg_config of an instance (RK coefficients of src/param/init.c:197-213)."""
    alpha, beta, gamma = rk_coefs()
    cfg = SnsgConfig()
    cfg.itot, cfg.jtot, cfg.ly, cfg.Ra, cfg.Pr = itot, jtot, ly, Ra, Pr
    cfg.with_temperature = int(with_temperature)
    cfg.with_thermal_forcing = int(with_thermal_forcing)
    cfg.implicitx, cfg.implicity = int(implicitx), int(implicity)
    for k in range(3):
        cfg.alpha[k], cfg.beta[k], cfg.gamma[k] = alpha[k], beta[k], gamma[k]
    return cfg


def instance_bytes(cfg):
    """Device bytes of an instance of this configuration (snsg_instance_bytes)."""
    n = C.c_size_t(0)
    _check(_load_hip().snsg_instance_bytes(C.byref(cfg), C.byref(n)), "snsg_instance_bytes")
    return n.value


def _launch(shapes):
    bad = set(shapes) - set(LAUNCH_FIELDS)
    if bad:
        raise ValueError("unknown launch shapes: %s" % ", ".join(sorted(bad)))
    return SnsgLaunch(**{k: int(shapes.get(k, 0)) for k in LAUNCH_FIELDS})


def launch_check(itot, jtot, **shapes):
    """Raises GpuError unless the kernels take these launch shapes for an
    itot x jtot grid (snsg_launch_check; host-only, needs no GPU)."""
    _check(_load_hip().snsg_launch_check(int(itot), int(jtot), C.byref(_launch(shapes))),
           "snsg_launch_check")


def tuning_add(itot, jtot, **shapes):
    """Table entry of (itot, jtot): instances created afterwards take these
    launch shapes (snsg_tuning_add)."""
    _check(_load_hip().snsg_tuning_add(int(itot), int(jtot), C.byref(_launch(shapes))),
           "snsg_tuning_add")


def tuning_clear():
    _check(_load_hip().snsg_tuning_clear(), "snsg_tuning_clear")


def read_tuning(path=TUNING_FILE):
    """Entries of a launch-shape table file: [(itot, jtot, shapes), ...]."""
    import json
    with open(path) as f:
        doc = json.load(f)
    return [(int(e["itot"]), int(e["jtot"]), {k: int(e.get(k, 0)) for k in LAUNCH_FIELDS})
            for e in doc["entries"]]


def load_tuning(path=TUNING_FILE):
    """Registers every entry of a launch-shape table file; returns their
    number."""
    entries = read_tuning(path)
    for itot, jtot, shapes in entries:
        tuning_add(itot, jtot, **shapes)
    return len(entries)


def _state_shape(name):
    """Reference-layout shape name of a state field."""
    return {"ux_star": "ux", "uy_star": "uy"}.get(name, name)


class Instance:
    """One simulation on one GPU: the reference's arrays as device buffers and
    the reference-named operations on them (through include/sns_hip.h)."""

    def __init__(self, itot, jtot, ly=2.0, stretch=3.0, Ra=1e8, Pr=1.0, implicitx=0,
                 implicity=0, with_temperature=True, with_thermal_forcing=True, device=0,
                 coords=None):
        """coords: x coordinates other than the reference's (a dict of grid()'s
        keys, e.g. tanh_grid()); stretch is then unused."""
        self.L = _load_hip()
        self.itot, self.jtot = itot, jtot
        self.g = grid(itot, jtot, ly, stretch) if coords is None else coords
        self.alpha, self.beta, self.gamma = rk_coefs()
        cfg = make_config(itot, jtot, ly, Ra, Pr, implicitx, implicity, with_temperature,
                          with_thermal_forcing)
        self.cfg = cfg
        self.Ra, self.Pr, self.ly = Ra, Pr, ly
        self.ctx = _vp()
        _check(self.L.snsg_create(device, C.byref(cfg), _p(self.g["dxf"]), _p(self.g["dxc"]),
                                  C.byref(self.ctx)), "snsg_create")
        self.buf = {}
        self.state = None

    # ---- lifetime
    def close(self):
        if self.ctx:
            if self.state is not None:
                _check(self.L.snsg_state_free(self.ctx, C.byref(self.state)), "snsg_state_free")
                self.state = None
            for v in self.buf.values():
                self.L.snsg_free(self.ctx, v)
            self.buf = {}
            _check(self.L.snsg_destroy(self.ctx), "snsg_destroy")
            self.ctx = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- memory
    def shape(self, name):
        return FIELD_SHAPES[name](self.itot, self.jtot)

    def alloc(self, *names):
        for n in names or [k for k in FIELD_SHAPES if k not in self.STAT_NAMES]:
            if n not in self.buf:
                nbytes = 8 * int(np.prod(self.shape(n)))
                ptr = _vp()
                _check(self.L.snsg_malloc(self.ctx, nbytes, C.byref(ptr)), "snsg_malloc")
                self.buf[n] = ptr
        return self

    def upload(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.shape(name), (name, a.shape, self.shape(name))
        self.alloc(name)
        _check(self.L.snsg_upload(self.ctx, self.buf[name], a.ctypes.data, a.nbytes), "upload")

    def download(self, name):
        a = np.zeros(self.shape(name))
        _check(self.L.snsg_download(self.ctx, a.ctypes.data, self.buf[name], a.nbytes), "download")
        return a

    def sync(self):
        _check(self.L.snsg_synchronize(self.ctx), "snsg_synchronize")

    def malloc(self, nbytes):
        """Raw zero-initialised device buffer (freed by free_raw or close)."""
        ptr = _vp()
        _check(self.L.snsg_malloc(self.ctx, int(nbytes), C.byref(ptr)), "snsg_malloc")
        self.buf["_raw%d" % ptr.value] = ptr
        return ptr

    def free_raw(self, ptr):
        self.buf.pop("_raw%d" % ptr.value, None)
        _check(self.L.snsg_free(self.ctx, ptr), "snsg_free")

    def copy(self, dst, src, nbytes, loads=0, nontemporal=False):
        """Device-to-device copy on the instance's stream: snsg_copy, or the
        tiled snsg_copy_shaped when loads (per thread: 1, 2, 4, 8) is set."""
        if loads:
            _check(self.L.snsg_copy_shaped(self.ctx, dst, src, int(nbytes), int(loads),
                                           int(bool(nontemporal))), "snsg_copy_shaped")
        else:
            _check(self.L.snsg_copy(self.ctx, dst, src, int(nbytes)), "snsg_copy")

    # ---- parameters
    def set_dt(self, dt):
        self.dt = float(dt)
        _check(self.L.snsg_set_dt(self.ctx, self.dt), "snsg_set_dt")

    def set_treatment(self, implicitx, implicity):
        _check(self.L.snsg_set_treatment(self.ctx, int(implicitx), int(implicity)),
               "snsg_set_treatment")

    def set_launch(self, **shapes):
        """Launch shapes of this instance (all of them; 0 or absent: the
        rule), before its first fused step (snsg_set_launch)."""
        _check(self.L.snsg_set_launch(self.ctx, C.byref(_launch(shapes))), "snsg_set_launch")

    def launch(self):
        """The launch shapes in effect, rules resolved (snsg_get_launch)."""
        l = SnsgLaunch()
        _check(self.L.snsg_get_launch(self.ctx, C.byref(l)), "snsg_get_launch")
        return {k: getattr(l, k) for k in LAUNCH_FIELDS}

    # ---- reference functions (host-layer conventions: B <- A by pointer exchange)
    def update_boundaries(self, name):
        fn = {"ux": "snsg_fluid_update_boundaries_ux", "uy": "snsg_fluid_update_boundaries_uy",
              "p": "snsg_fluid_update_boundaries_p", "psi": "snsg_fluid_update_boundaries_p",
              "temp": "snsg_temperature_update_boundaries_temp"}[name]
        _check(getattr(self.L, fn)(self.ctx, self.buf[name]), fn)

    def compute_force(self):
        _check(self.L.snsg_temperature_compute_force(self.ctx, self.buf["temp"],
                                                     self.buf["tempforcex"]), "compute_force")

    def _swap(self, a, b):
        self.buf[a], self.buf[b] = self.buf[b], self.buf[a]

    def update_temp(self, rkstep):
        self._swap("srctempa", "srctempb")
        b = self.buf
        _check(self.L.snsg_temperature_update_temp(self.ctx, rkstep, b["ux"], b["uy"], b["temp"],
                                                   b["srctempa"], b["srctempb"], b["srctempg"]),
               "update_temp")

    def update_velocity(self, rkstep):
        self._swap("srcuxa", "srcuxb")
        self._swap("srcuya", "srcuyb")
        b = self.buf
        _check(self.L.snsg_fluid_update_velocity(self.ctx, rkstep, b["ux"], b["uy"], b["p"],
                                                 b["tempforcex"], b["srcuxa"], b["srcuxb"],
                                                 b["srcuxg"], b["srcuya"], b["srcuyb"],
                                                 b["srcuyg"]), "update_velocity")

    def compute_potential(self, rkstep):
        b = self.buf
        _check(self.L.snsg_fluid_compute_potential(self.ctx, rkstep, b["ux"], b["uy"], b["psi"]),
               "compute_potential")

    def correct_velocity(self, rkstep):
        b = self.buf
        _check(self.L.snsg_fluid_correct_velocity(self.ctx, rkstep, b["psi"], b["ux"], b["uy"]),
               "correct_velocity")

    def update_pressure(self, rkstep):
        b = self.buf
        _check(self.L.snsg_fluid_update_pressure(self.ctx, rkstep, b["psi"], b["p"]),
               "update_pressure")

    def dt_adv_raw(self, ux="ux", uy="uy"):
        v = C.c_double(0.0)
        _check(self.L.snsg_param_decide_dt(self.ctx, self.buf[ux], self.buf[uy], C.byref(v)),
               "decide_dt")
        return v.value

    def transpose(self, g_isize, g_jsize, elemsize, send_ptr, recv_ptr):
        _check(self.L.snsg_transpose(self.ctx, g_isize, g_jsize, elemsize, send_ptr, recv_ptr),
               "transpose")

    def integrate_reference_order(self):
        """One RK3 step through the reference-named operations (src/main.c:79-101)."""
        for rk in range(3):
            for n in ("ux", "uy", "p", "temp"):
                self.update_boundaries(n)
            if self.cfg.with_temperature:
                self.compute_force()
                self.update_temp(rk)
            self.update_velocity(rk)
            self.compute_potential(rk)
            self.correct_velocity(rk)
            self.update_pressure(rk)

    # ---- fused step on instance-owned state
    def state_alloc(self):
        if self.state is None:
            self.state = SnsgState()
            _check(self.L.snsg_state_alloc(self.ctx, C.byref(self.state)), "snsg_state_alloc")
        return self.state

    def state_upload(self, name, arr):
        """Copy a reference-layout array into the padded state field."""
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.shape(_state_shape(name)), (name, a.shape)
        _check(self.L.snsg_state_upload(self.ctx, C.byref(self.state), STATE_FIELDS[name],
                                        a.ctypes.data), "snsg_state_upload")

    def state_init_temperature(self, seed, amplitude):
        """Seeded initial temperature on the device (snsg_state_init_temperature)."""
        _check(self.L.snsg_state_init_temperature(self.ctx, C.byref(self.state), int(seed),
                                                  float(amplitude)),
               "snsg_state_init_temperature")

    def uniform_grid(self):
        u = C.c_int(0)
        _check(self.L.snsg_uniform_grid(self.ctx, C.byref(u)), "snsg_uniform_grid")
        return bool(u.value)

    def memory(self):
        n = C.c_size_t(0)
        _check(self.L.snsg_memory(self.ctx, C.byref(n)), "snsg_memory")
        return n.value

    def state_download(self, name, shape_name=None):
        """A state field in the reference layout."""
        a = np.zeros(self.shape(shape_name or _state_shape(name)))
        _check(self.L.snsg_state_download(self.ctx, C.byref(self.state), STATE_FIELDS[name],
                                          a.ctypes.data), "snsg_state_download")
        return a

    def integrate(self):
        _check(self.L.snsg_integrate(self.ctx, C.byref(self.state)), "snsg_integrate")

    def advance(self, nsteps):
        """nsteps time-loop iterations with the time step decided on the device."""
        _check(self.L.snsg_advance(self.ctx, C.byref(self.state), int(nsteps)), "snsg_advance")

    def get_time(self):
        t, dt = C.c_double(0.0), C.c_double(0.0)
        _check(self.L.snsg_get_time(self.ctx, C.byref(t), C.byref(dt)), "snsg_get_time")
        return t.value, dt.value

    def set_time(self, t):
        _check(self.L.snsg_set_time(self.ctx, float(t)), "snsg_set_time")

    DIAG_NAMES = ("divmax", "divsum", "mom_x", "mom_y", "energy_ux", "energy_uy", "energy_t",
                  "nu_wall", "nu_inje", "nu_eps_k", "nu_eps_h")

    def diagnostics(self, state=False):
        """The monitors of src/logging.c evaluated on the device, as a dict."""
        out = np.zeros(11)
        if state:
            _check(self.L.snsg_state_diagnostics(self.ctx, C.byref(self.state), _p(out)),
                   "snsg_state_diagnostics")
        else:
            b = self.buf
            _check(self.L.snsg_diagnostics(self.ctx, b["ux"], b["uy"], b.get("temp"), _p(out)),
                   "snsg_diagnostics")
        return dict(zip(self.DIAG_NAMES, out))

    STAT_NAMES = ("ux1", "ux2", "uy1", "uy2", "temp1", "temp2")

    def statistics_collect(self, state=False):
        """One sample into the accumulators ux1 .. temp2 (src/statistics/collect.c),
        from the instance's ux, uy, temp buffers or from the state."""
        self.alloc(*self.STAT_NAMES)
        acc = [self.buf[n] for n in self.STAT_NAMES]
        if state:
            _check(self.L.snsg_state_statistics_collect(self.ctx, C.byref(self.state), *acc),
                   "snsg_state_statistics_collect")
        else:
            b = self.buf
            _check(self.L.snsg_statistics_collect(self.ctx, b["ux"], b["uy"], b.get("temp"), *acc),
                   "snsg_statistics_collect")

    def state_dt_adv_raw(self):
        v = C.c_double(0.0)
        _check(self.L.snsg_state_decide_dt(self.ctx, C.byref(self.state), C.byref(v)),
               "snsg_state_decide_dt")
        return v.value

    # ---- timing
    def timing_enable(self, name_filter=None):
        _check(self.L.snsg_timing_enable(self.ctx, name_filter.encode() if name_filter else None),
               "timing_enable")

    def timing_read(self):
        n, ms = C.c_int(0), C.c_double(0.0)
        _check(self.L.snsg_timing_read(self.ctx, C.byref(n), C.byref(ms)), "timing_read")
        return n.value, ms.value

    def timing_disable(self):
        _check(self.L.snsg_timing_disable(self.ctx), "timing_disable")


def decide_dt(g, Ra, Pr, dt_adv_raw, implicitx, implicity):
    """Host part of src/param/decide_dt.c:11-101 with a forced treatment:
    the advective bound (device reduction) times 0.7, the diffusive bounds,
    and the reference's selection rule fed a cost table that picks the
    forced treatment.  Evaluated in C by the host layer's formula restated
    here with the same operation order."""
    import math
    dt_adv = dt_adv_raw * 0.7
    diffusivity = max(math.sqrt(Pr) / math.sqrt(Ra), 1. / math.sqrt(Pr) / math.sqrt(Ra))
    dx = 1.0
    for v in g["dxc"][1:-1]:
        dx = min(dx, v)
    dt_dif_x = 0.25 / diffusivity * dx ** 2.0 * 0.9
    dt_dif_y = 0.25 / diffusivity * g["dy"] ** 2.0 * 0.9
    if implicitx and implicity:
        return dt_adv
    if implicitx:
        return min(dt_adv, dt_dif_y)
    if implicity:
        return min(dt_adv, dt_dif_x)
    return min(dt_adv, min(dt_dif_x, dt_dif_y))


class Batch:
    """Instances advanced together (snsg_batch_*): each kernel launched once
    per shape class for all its members, one stream and one replayed HIP
    graph per class.  The instances must have their states allocated and
    must outlive the batch."""

    def __init__(self, instances, eager=False):
        self.L = _load_hip()
        self.instances = list(instances)
        n = len(self.instances)
        ctxs = (_vp * n)(*[i.ctx for i in self.instances])
        sts = (C.POINTER(SnsgState) * n)(*[C.pointer(i.state) for i in self.instances])
        self.b = _vp()
        _check(self.L.snsg_batch_create(n, ctxs, sts, 1 if eager else 0, C.byref(self.b)),
               "snsg_batch_create")

    def advance(self, nsteps):
        _check(self.L.snsg_batch_advance(self.b, int(nsteps)), "snsg_batch_advance")

    def sync(self):
        _check(self.L.snsg_batch_synchronize(self.b), "snsg_batch_synchronize")

    def classes(self):
        n = C.c_int(0)
        _check(self.L.snsg_batch_classes(self.b, C.byref(n)), "snsg_batch_classes")
        return n.value

    def close(self):
        if self.b:
            self.L.snsg_batch_destroy(self.b)
            self.b = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving engine (snsg_server_*): jobs submitted at any time are
    admitted against a device-memory budget and advanced continuously, one
    lane per shape class."""

    def __init__(self, mem_budget, max_active=0, device=0):
        self.L = _load_hip()
        self.s = _vp()
        _check(self.L.snsg_server_create(int(device), int(mem_budget), int(max_active),
                                         C.byref(self.s)), "snsg_server_create")
        self.sizes = {}

    def submit(self, itot, jtot, nsteps, Ra=1e8, Pr=1.0, implicitx=0, implicity=0, seed=1,
               amplitude=1e-3, ly=2.0, stretch=None, keep=False, with_temperature=True):
        g = grid(itot, jtot, ly, float(itot) if stretch is None else stretch)
        job = SnsgJob()
        job.cfg = make_config(itot, jtot, ly, Ra, Pr, implicitx, implicity, with_temperature)
        job.dxf, job.dxc = _p(g["dxf"]), _p(g["dxc"])
        job.seed, job.amplitude = int(seed), float(amplitude)
        job.nsteps, job.keep = int(nsteps), int(keep)
        jid = C.c_int(0)
        _check(self.L.snsg_server_submit(self.s, C.byref(job), C.byref(jid)), "snsg_server_submit")
        self.sizes[jid.value] = (itot, jtot)
        return jid.value

    def wait(self, jid, timeout=600.0):
        _check(self.L.snsg_server_wait(self.s, int(jid), float(timeout)), "snsg_server_wait")

    def drain(self, timeout=600.0):
        _check(self.L.snsg_server_drain(self.s, float(timeout)), "snsg_server_drain")

    def status(self, jid):
        st, n = C.c_int(0), C.c_int(0)
        _check(self.L.snsg_server_status(self.s, int(jid), C.byref(st), C.byref(n)),
               "snsg_server_status")
        return st.value, n.value

    def download(self, jid, name):
        itot, jtot = self.sizes[jid]
        a = np.zeros(FIELD_SHAPES[_state_shape(name)](itot, jtot))
        _check(self.L.snsg_server_download(self.s, int(jid), STATE_FIELDS[name], a.ctypes.data),
               "snsg_server_download")
        return a

    def job_time(self, jid):
        t, dt = C.c_double(0.0), C.c_double(0.0)
        _check(self.L.snsg_server_job_time(self.s, int(jid), C.byref(t), C.byref(dt)),
               "snsg_server_job_time")
        return t.value, dt.value

    def release(self, jid):
        _check(self.L.snsg_server_release(self.s, int(jid)), "snsg_server_release")

    def stats(self):
        st = SnsgServerStats()
        _check(self.L.snsg_server_stats_get(self.s, C.byref(st)), "snsg_server_stats_get")
        return {n: getattr(st, n) for n, _ in SnsgServerStats._fields_}

    def reset_stats(self):
        _check(self.L.snsg_server_reset_stats(self.s), "snsg_server_reset_stats")

    def close(self):
        if self.s:
            self.L.snsg_server_destroy(self.s)
            self.s = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
