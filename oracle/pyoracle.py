"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product package never
imports this module.

Arrays use the reference layouts (include/arrays/*.h in the reference):
ux (jtot+2, itot+1), uy/p/psi/temp (jtot+2, itot+2) with one y-halo row on
each side; source terms and tempforcex have no halos.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)


class OrcParam(C.Structure):
    _fields_ = [
        ("itot", C.c_int), ("jtot", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double),
        ("stretch", C.c_double),
        ("Ra", C.c_double), ("Pr", C.c_double),
        ("with_temperature", C.c_int), ("with_thermal_forcing", C.c_int),
        ("xf", _dp), ("xc", _dp), ("dxf", _dp), ("dxc", _dp), ("yf", _dp), ("yc", _dp),
        ("dy", C.c_double),
        ("alpha", C.c_double * 3), ("beta", C.c_double * 3), ("gamma", C.c_double * 3),
        ("time", C.c_double), ("dt", C.c_double),
        ("step", C.c_int), ("implicitx", C.c_int), ("implicity", C.c_int),
        ("expimp_wtimes", (C.c_double * 2) * 2),
    ]


_STATE_FIELDS = ["ux", "uy", "p", "psi",
                 "srcuxa", "srcuxb", "srcuxg", "srcuya", "srcuyb", "srcuyg",
                 "temp", "tempforcex", "srctempa", "srctempb", "srctempg"]


class OrcState(C.Structure):
    _fields_ = [(n, _dp) for n in _STATE_FIELDS]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P, S = C.POINTER(OrcParam), C.POINTER(OrcState)
        for name in ["orc_set_coordinate", "orc_set_rk_coefs"]:
            getattr(L, name).argtypes = [P]
        L.orc_param_decide_dt.argtypes = [P, S]
        for name in ["orc_fluid_update_boundaries_ux", "orc_fluid_update_boundaries_uy",
                     "orc_fluid_update_boundaries_p", "orc_temperature_update_boundaries_temp"]:
            getattr(L, name).argtypes = [P, _dp]
        L.orc_temperature_compute_force.argtypes = [P, S]
        for name in ["orc_temperature_update_temp", "orc_fluid_update_velocity",
                     "orc_fluid_compute_potential", "orc_fluid_correct_velocity",
                     "orc_fluid_update_pressure"]:
            getattr(L, name).argtypes = [P, C.c_int, S]
        L.orc_integrate.argtypes = [P, S]
        L.orc_run.argtypes = [P, S, C.c_int]
        for name in ["orc_dgtsv_b", "orc_dgtsv_p"]:
            getattr(L, name).argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.orc_zgtsv_b.argtypes = [C.c_int, _dp, _dp, _dp, C.c_void_p]
        L.orc_rfft.argtypes = [C.c_int, _dp, C.c_void_p]
        L.orc_irfft.argtypes = [C.c_int, C.c_void_p, _dp]
        L.orc_parallel_get_size.argtypes = [C.c_int] * 3
        L.orc_parallel_get_offset.argtypes = [C.c_int] * 3
        L.orc_parallel_transpose_sim.argtypes = [C.c_int, C.c_int, C.c_int, C.c_size_t,
                                                 C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]
        for name in ["orc_check_divergence", "orc_check_momentum", "orc_check_energy",
                     "orc_check_nusselt"]:
            getattr(L, name).argtypes = [P, S, _dp]
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_init_temperature.argtypes = [P, C.c_uint64, C.c_double, C.c_int, _dp]
        _lib = L
    return _lib


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def forced_wtimes(implicitx, implicity):
    """expimp_wtimes table that makes the reference's selection rule
    (src/param/decide_dt.c:74-97) pick the given treatment."""
    big = np.finfo(np.float64).max
    w = [[big, big], [big, big]]
    w[implicity][implicitx] = 1.0
    return w


class Oracle:
    """One single-rank simulation held in numpy arrays."""

    def __init__(self, itot, jtot, ly=2.0, stretch=3.0, Ra=1e8, Pr=1.0,
                 with_temperature=True, with_thermal_forcing=True, lx=1.0):
        self.L = lib()
        self.itot, self.jtot = itot, jtot
        self.xf = np.zeros(itot + 1)
        self.xc = np.zeros(itot + 2)
        self.dxf = np.zeros(itot)
        self.dxc = np.zeros(itot + 1)
        self.yf = np.zeros(jtot + 1)
        self.yc = np.zeros(jtot)
        pr = OrcParam()
        pr.itot, pr.jtot = itot, jtot
        pr.lx, pr.ly, pr.stretch, pr.Ra, pr.Pr = lx, ly, stretch, Ra, Pr
        pr.with_temperature = int(with_temperature)
        pr.with_thermal_forcing = int(with_thermal_forcing)
        for n in ["xf", "xc", "dxf", "dxc", "yf", "yc"]:
            setattr(pr, n, _ptr(getattr(self, n)))
        self.param = pr
        self.L.orc_set_coordinate(C.byref(pr))
        self.L.orc_set_rk_coefs(C.byref(pr))
        self.set_treatment(0, 0)
        shapes = {
            "ux": (jtot + 2, itot + 1), "uy": (jtot + 2, itot + 2),
            "p": (jtot + 2, itot + 2), "psi": (jtot + 2, itot + 2),
            "srcuxa": (jtot, itot - 1), "srcuxb": (jtot, itot - 1), "srcuxg": (jtot, itot - 1),
            "srcuya": (jtot, itot), "srcuyb": (jtot, itot), "srcuyg": (jtot, itot),
            "temp": (jtot + 2, itot + 2), "tempforcex": (jtot, itot - 1),
            "srctempa": (jtot, itot), "srctempb": (jtot, itot), "srctempg": (jtot, itot),
        }
        self.f = {n: np.zeros(shapes[n]) for n in _STATE_FIELDS}
        st = OrcState()
        for n in _STATE_FIELDS:
            setattr(st, n, _ptr(self.f[n]))
        self.state = st

    # -- configuration
    @property
    def dy(self):
        return self.param.dy

    @property
    def dt(self):
        return self.param.dt

    @dt.setter
    def dt(self, v):
        self.param.dt = float(v)

    def set_treatment(self, implicitx, implicity):
        self.param.implicitx = int(implicitx)
        self.param.implicity = int(implicity)
        w = forced_wtimes(implicitx, implicity)
        for a in range(2):
            for b in range(2):
                self.param.expimp_wtimes[a][b] = w[a][b]

    def rk(self):
        return (list(self.param.alpha), list(self.param.beta), list(self.param.gamma))

    # -- state
    def init_temperature(self, seed, amplitude, linear=False):
        self.L.orc_init_temperature(C.byref(self.param), seed, amplitude, int(linear),
                                    _ptr(self.f["temp"]))

    def apply_boundaries(self):
        P = C.byref(self.param)
        self.L.orc_fluid_update_boundaries_ux(P, _ptr(self.f["ux"]))
        self.L.orc_fluid_update_boundaries_uy(P, _ptr(self.f["uy"]))
        self.L.orc_fluid_update_boundaries_p(P, _ptr(self.f["p"]))
        self.L.orc_temperature_update_boundaries_temp(P, _ptr(self.f["temp"]))

    def copy_fields(self):
        return {n: a.copy() for n, a in self.f.items()}

    def load_fields(self, fields):
        for n, a in fields.items():
            self.f[n][...] = a

    # -- reference functions
    def bc(self, name, arr):
        fn = {"ux": "orc_fluid_update_boundaries_ux", "uy": "orc_fluid_update_boundaries_uy",
              "p": "orc_fluid_update_boundaries_p",
              "temp": "orc_temperature_update_boundaries_temp"}[name]
        getattr(self.L, fn)(C.byref(self.param), _ptr(arr))

    def compute_force(self):
        self.L.orc_temperature_compute_force(C.byref(self.param), C.byref(self.state))

    def update_temp(self, rkstep):
        self.L.orc_temperature_update_temp(C.byref(self.param), rkstep, C.byref(self.state))

    def update_velocity(self, rkstep):
        self.L.orc_fluid_update_velocity(C.byref(self.param), rkstep, C.byref(self.state))

    def compute_potential(self, rkstep):
        self.L.orc_fluid_compute_potential(C.byref(self.param), rkstep, C.byref(self.state))

    def correct_velocity(self, rkstep):
        self.L.orc_fluid_correct_velocity(C.byref(self.param), rkstep, C.byref(self.state))

    def update_pressure(self, rkstep):
        self.L.orc_fluid_update_pressure(C.byref(self.param), rkstep, C.byref(self.state))

    def decide_dt(self):
        self.L.orc_param_decide_dt(C.byref(self.param), C.byref(self.state))
        return self.param.dt

    def integrate(self):
        self.L.orc_integrate(C.byref(self.param), C.byref(self.state))

    def run(self, nsteps):
        self.L.orc_run(C.byref(self.param), C.byref(self.state), int(nsteps))

    # -- diagnostics
    def _diag(self, fn, n):
        out = np.zeros(n)
        getattr(self.L, fn)(C.byref(self.param), C.byref(self.state), _ptr(out))
        return out

    def divergence(self):
        return self._diag("orc_check_divergence", 2)

    def momentum(self):
        return self._diag("orc_check_momentum", 2)

    def energy(self):
        return self._diag("orc_check_energy", 3)

    def nusselt(self):
        return self._diag("orc_check_nusselt", 4)


# -- free functions --------------------------------------------------------

def dgtsv_b(l, c, u, q):
    l, c, u, q = (np.array(a, dtype=np.float64) for a in (l, c, u, q))
    lib().orc_dgtsv_b(len(q), _ptr(l), _ptr(c), _ptr(u), _ptr(q))
    return q


def zgtsv_b(l, c, u, q):
    l, c, u = (np.array(a, dtype=np.float64) for a in (l, c, u))
    q = np.array(q, dtype=np.complex128)
    lib().orc_zgtsv_b(len(q), _ptr(l), _ptr(c), _ptr(u), q.ctypes.data)
    return q


def dgtsv_p(l, c, u, q):
    l, c, u, q = (np.array(a, dtype=np.float64) for a in (l, c, u, q))
    lib().orc_dgtsv_p(len(q), _ptr(l), _ptr(c), _ptr(u), _ptr(q))
    return q


def rfft(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(x)
    out = np.zeros(n // 2 + 1, dtype=np.complex128)
    lib().orc_rfft(n, _ptr(x), out.ctypes.data)
    return out


def irfft(y, n):
    y = np.ascontiguousarray(y, dtype=np.complex128)
    out = np.zeros(n)
    lib().orc_irfft(n, y.ctypes.data, _ptr(out))
    return out


def get_size(num, size, rank):
    return lib().orc_parallel_get_size(num, size, rank)


def get_offset(num, size, rank):
    return lib().orc_parallel_get_offset(num, size, rank)


def transpose_sim(nranks, g_isize, g_jsize, sendbufs, recvbufs):
    """Simulated MPI_Alltoallw transpose over nranks slabs (numpy buffers)."""
    es = sendbufs[0].dtype.itemsize
    S = (C.c_void_p * nranks)(*[b.ctypes.data for b in sendbufs])
    R = (C.c_void_p * nranks)(*[b.ctypes.data for b in recvbufs])
    lib().orc_parallel_transpose_sim(nranks, g_isize, g_jsize, es, S, R)


VARIANT_DIRECT_DFT = 1      # Poisson transforms as a direct DFT
VARIANT_REVERSED_SOLVE = 2  # tri-diagonal systems eliminated last row first
VARIANT_PIN_GAUGE = 4       # J = 0 Poisson system pinned at psi_hat(itot) = 0


def set_variant(bits):
    """Alternative operation orders for the same arithmetic (0: the
    reference's), for rounding-floor runs."""
    lib().orc_set_variant(int(bits))


def num_threads():
    """Threads the oracle's OpenMP loops use (OMP_NUM_THREADS)."""
    return int(lib().orc_num_threads())


def splitmix64(x):
    return lib().orc_splitmix64(x)


# -- y-slab MPI restatement (oracle/sns_oracle_mpi) --------------------------

MPI_EXE = os.path.join(_HERE, "sns_oracle_mpi")
MPIEXEC_CANDIDATES = ["/opt/conda/bin/mpiexec"]


def mpiexec():
    """An MPI launcher that can run the MPI oracle, or None."""
    for c in MPIEXEC_CANDIDATES:
        if os.path.exists(c):
            return c
    return None


def mpi_run(ranks, itot, jtot, Ra, Pr, implicitx, implicity, seed, amplitude, nsteps,
            stretch=None, pin_gauge=False, out=None, timeout=None):
    """Run the MPI oracle on `ranks` local ranks; returns its JSON summary
    and, with `out` (a file path), the interior rows of ux, uy, temp, p."""
    import json
    exe, launcher = MPI_EXE, mpiexec()
    if not os.path.exists(exe) or launcher is None:
        raise FileNotFoundError("MPI oracle or mpiexec not available")
    cmd = [launcher, "-launcher", "fork", "-n", str(int(ranks)), exe, str(itot), str(jtot),
           repr(float(Ra)), repr(float(Pr)), str(int(implicitx)), str(int(implicity)),
           str(int(seed)), repr(float(amplitude)), str(int(nsteps)),
           repr(float(itot if stretch is None else stretch)), str(int(bool(pin_gauge)))]
    if out:
        cmd.append(out)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run(cmd, capture_output=True, text=True, check=True, timeout=timeout, env=env)
    summary = json.loads(r.stdout.strip().splitlines()[-1])
    fields = None
    if out:
        d = np.fromfile(out)
        fields, off = {}, 0
        for n, w in (("ux", itot + 1), ("uy", itot + 2), ("temp", itot + 2), ("p", itot + 2)):
            fields[n] = d[off:off + jtot * w].reshape(jtot, w)
            off += jtot * w
    return summary, fields
