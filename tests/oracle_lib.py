"""ctypes access to the oracle (oracle/liboracle_mhd.so) and to the reference's
own C functions compiled by oracle/Makefile (oracle/_ref/libmhd_ref.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only to check or to time the CPU
path beside the GPU one.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle_mhd.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libmhd_ref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_I = ctypes.c_int


def _bind(lib, name, restype, argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


class Oracle:
    """The C restatement, FP32 (reference evaluation order) and FP64."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        self.lib = lib = ctypes.CDLL(path)
        for sfx, p, R in (("_f32", _f32p, ctypes.c_float), ("_f64", _f64p, ctypes.c_double)):
            _bind(lib, "mhd_oracle_fluidx" + sfx, None, [p, p, _I, _I, _I, R])
            _bind(lib, "mhd_oracle_advectbyzx" + sfx, None, [p, p, _I, _I, _I, R])
            _bind(lib, "mhd_oracle_transpose12" + sfx, None, [p, p, p, p, _I, _I, _I])
            _bind(lib, "mhd_oracle_transpose13" + sfx, None, [p, p, p, p, _I, _I, _I])
            _bind(lib, "mhd_oracle_calcfl" + sfx, R, [p, p, _I, _I, _I, _I])
            _bind(lib, "mhd_oracle_step" + sfx, R, [p, p, _I, _I, _I, ctypes.POINTER(R), R, _I])
            _bind(lib, "mhd_oracle_run" + sfx, None, [p, p, _I, _I, _I, _I, R, _I, ctypes.c_void_p])
        _bind(lib, "mhd_oracle_read_ref_order", None, [_f32p, _f32p, _f32p, _I, _I, _I])
        _bind(lib, "mhd_oracle_monitors_f32", None, [_f32p, _f32p, _I, _I, _I, _f64p])
        _bind(lib, "mhd_oracle_threads", _I, [])

    @staticmethod
    def _sfx(a):
        return "_f64" if a.dtype == np.float64 else "_f32"

    def fluidx(self, u, b, nx, ny, nz, dt):
        getattr(self.lib, "mhd_oracle_fluidx" + self._sfx(u))(u, b, nx, ny, nz, dt)

    def advectbyzx(self, u, b, nx, ny, nz, dt):
        getattr(self.lib, "mhd_oracle_advectbyzx" + self._sfx(u))(u, b, nx, ny, nz, dt)

    def transpose12(self, u, b, nx, ny, nz):
        uo, bo = np.empty_like(u), np.empty_like(b)
        getattr(self.lib, "mhd_oracle_transpose12" + self._sfx(u))(uo, bo, u, b, nx, ny, nz)
        return uo, bo

    def transpose13(self, u, b, nx, ny, nz):
        uo, bo = np.empty_like(u), np.empty_like(b)
        getattr(self.lib, "mhd_oracle_transpose13" + self._sfx(u))(uo, bo, u, b, nx, ny, nz)
        return uo, bo

    def calcfl(self, u, b, nx, ny, nz, avg_b=False):
        return getattr(self.lib, "mhd_oracle_calcfl" + self._sfx(u))(u, b, nx, ny, nz, int(avg_b))

    def threads(self):
        """Threads of the OpenMP pencil loops."""
        return self.lib.mhd_oracle_threads()

    def run(self, u, b, nx, ny, nz, nsteps, tf=float("inf"), avg_b=False):
        """nsteps reference loop iterations in place; returns the dt sequence."""
        dts = np.zeros(max(nsteps, 1), dtype=u.dtype)
        getattr(self.lib, "mhd_oracle_run" + self._sfx(u))(
            u, b, nx, ny, nz, nsteps, tf, int(avg_b), dts.ctypes.data_as(ctypes.c_void_p))
        return dts[:nsteps]

    def read_ref_order(self, rec, n):
        u = np.empty(5 * n ** 3, np.float32)
        b = np.empty(3 * n ** 3, np.float32)
        self.lib.mhd_oracle_read_ref_order(np.ascontiguousarray(rec, np.float32), u, b, n, n, n)
        return u, b

    def monitors(self, u, b, nx, ny, nz):
        out = np.zeros(6, np.float64)
        self.lib.mhd_oracle_monitors_f32(u, b, nx, ny, nz, out)
        return out


class Reference:
    """The reference's own C functions (Heterogeneous_MHD/C), FP32 only."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise RuntimeError(f"reference build missing: {path}")
        self.lib = lib = ctypes.CDLL(path)
        F = ctypes.c_float
        _bind(lib, "ref_fluidx", None, [_f32p, _f32p, _I, _I, _I, F])
        _bind(lib, "ref_advectbyzx", None, [_f32p, _f32p, _I, _I, _I, F])
        _bind(lib, "ref_transpose12", None, [_f32p, _f32p, _I, _I, _I])
        _bind(lib, "ref_transpose13", None, [_f32p, _f32p, _I, _I, _I])
        _bind(lib, "ref_calcfl", F, [_f32p, _f32p, _I, _I, _I])
        _bind(lib, "ref_run", None, [_f32p, _f32p, _I, _I, _I, _I, F, ctypes.c_void_p])

    def fluidx(self, u, b, nx, ny, nz, dt):
        self.lib.ref_fluidx(u, b, nx, ny, nz, dt)

    def advectbyzx(self, u, b, nx, ny, nz, dt):
        self.lib.ref_advectbyzx(u, b, nx, ny, nz, dt)

    def transpose12(self, u, b, nx, ny, nz):
        self.lib.ref_transpose12(u, b, nx, ny, nz)

    def transpose13(self, u, b, nx, ny, nz):
        self.lib.ref_transpose13(u, b, nx, ny, nz)

    def calcfl(self, u, b, nx, ny, nz):
        return self.lib.ref_calcfl(u, b, nx, ny, nz)

    def run(self, u, b, nx, ny, nz, nsteps, tf=float("inf")):
        dts = np.zeros(max(nsteps, 1), dtype=np.float32)
        self.lib.ref_run(u, b, nx, ny, nz, nsteps, tf, dts.ctypes.data_as(ctypes.c_void_p))
        return dts[:nsteps]


def reference_available():
    return os.path.exists(REF_SO)
