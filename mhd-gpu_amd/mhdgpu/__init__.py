"""Python host mirror of the MI355X relaxing-TVD MHD library (libmhdgpu.so).

The functions keep the reference's operator interface for the timestep
(HIP/hip_subroutine.h:4-12, Heterogeneous_MHD/C/spu_function.h:1-15):
``fluidx``, ``advectbyzx``, ``transpose12``, ``transpose13``, ``cfl`` and
``init``/``main``, over AoS ``u`` (5 per cell) and ``b`` (3 per cell) arrays,
with the same argument meaning.  Errors raise ``MHDError`` carrying the
library's code and message (the reference prints and continues or exits).

Everything runs through the C ABI of include/mhd_api.h with ctypes; device
memory is held in ``DeviceArray`` objects.  There is no CPU fallback: if the
library is missing, importing works but the first call raises.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MHDGPU_LIB", os.path.join(os.path.dirname(_HERE), "libmhdgpu.so"))

MHD_OK = 0
ERRORS = {-1: "MHD_ERR_ARG", -2: "MHD_ERR_SHAPE", -3: "MHD_ERR_HIP", -4: "MHD_ERR_NOMEM", -5: "MHD_ERR_STATE",
          -6: "MHD_ERR_IO", -7: "MHD_ERR_FULL", -8: "MHD_ERR_NONFINITE"}

# every C entry point of include/mhd_api.h: name -> (restype, argtypes)
_P = ctypes.c_void_p
_I = ctypes.c_int
_IP = ctypes.POINTER(ctypes.c_int)
_FP = ctypes.POINTER(ctypes.c_float)
_DP = ctypes.POINTER(ctypes.c_double)
_SZ = ctypes.c_size_t
_SZP = ctypes.POINTER(ctypes.c_size_t)
SIGNATURES = {
    "mhd_last_error": (ctypes.c_char_p, []),
    "mhd_version": (ctypes.c_char_p, []),
    "mhd_dev_alloc": (_I, [ctypes.POINTER(_P), _SZ]),
    "mhd_dev_free": (_I, [_P]),
    "mhd_memcpy_h2d": (_I, [_P, _P, _SZ]),
    "mhd_memcpy_d2h": (_I, [_P, _P, _SZ]),
    "mhd_memcpy_d2d": (_I, [_P, _P, _SZ]),
    "mhd_device_sync": (_I, []),
    "mhd_dev_copy_kernel": (_I, [_P, _P, _SZ]),
    "mhd_mem_info": (_I, [_SZP, _SZP]),
    "mhd_selfcheck": (_I, [ctypes.POINTER(ctypes.c_double)]),
    "mhd_selfcheck_division": (_I, [ctypes.c_int, ctypes.POINTER(ctypes.c_ulonglong)]),
    "mhd_selfcheck_limiter": (_I, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "mhd_selfcheck_sqrt": (_I, [ctypes.POINTER(ctypes.c_ulonglong)]),
    "mhd_fluidx": (_I, [_P, _P, _IP, _IP, _IP, _P]),
    "mhd_advectbyzx": (_I, [_P, _P, _IP, _IP, _IP, _P]),
    "mhd_advectbyzx_noalloc": (_I, [_P, _P, _IP, _IP, _IP, _P]),
    "mhd_transpose12": (_I, [_P, _P, _P, _P, _IP, _IP, _IP]),
    "mhd_transpose13": (_I, [_P, _P, _P, _P, _IP, _IP, _IP]),
    "mhd_transpose12_": (_I, [_P, _P, _P, _P, _IP, _IP, _IP, _FP]),
    "mhd_transpose13_": (_I, [_P, _P, _P, _P, _IP, _IP, _IP, _FP]),
    "mhd_cfl": (_I, [_P, _P, _IP, _IP, _IP, _I, _P, _FP]),
    "mhd_init": (_I, [ctypes.c_char_p, _I, _P, _P, _IP, _IP, _IP]),
    "mhd_write_result": (_I, [ctypes.c_char_p, ctypes.c_char_p, _P, _P, _IP, _IP, _IP]),
    "mhd_main": (_I, [_P, _P, _IP, _IP, _IP, _I, ctypes.c_float, _FP, _I, _IP]),
    "mhd_ctx_create": (_I, [ctypes.POINTER(_P), _I, _I, _I, _I, _I]),
    "mhd_ctx_destroy": (_I, [_P]),
    "mhd_ctx_set_state": (_I, [_P, _P, _P]),
    "mhd_ctx_get_state": (_I, [_P, _P, _P]),
    "mhd_ctx_reset_clock": (_I, [_P, ctypes.c_double]),
    "mhd_ctx_set_advect_form": (_I, [_P, _I]),
    "mhd_ctx_step": (_I, [_P, _I]),
    "mhd_ctx_step_timed": (_I, [_P, _I, _FP]),
    "mhd_ctx_sync": (_I, [_P]),
    "mhd_ctx_dts": (_I, [_P, _DP, _I, _IP]),
    "mhd_ctx_time": (_I, [_P, _DP, _DP]),
    "mhd_ctx_monitors": (_I, [_P, _DP]),
    "mhd_ctx_monitor_history": (_I, [_P, _DP, _I, _IP]),
    "mhd_ctx_device_state": (_I, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    "mhd_ctx_layout": (_I, [_P, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long),
                            ctypes.POINTER(ctypes.c_long)]),
    "mhd_ctx_profile_step": (_I, [_P, _I, _FP]),
    "mhd_batch_create": (_I, [ctypes.POINTER(_P), _I, _IP, _I]),
    "mhd_batch_destroy": (_I, [_P]),
    "mhd_batch_set_state": (_I, [_P, _I, _P, _P]),
    "mhd_batch_get_state": (_I, [_P, _I, _P, _P]),
    "mhd_batch_step": (_I, [_P, _I]),
    "mhd_batch_step_timed": (_I, [_P, _I, _FP]),
    "mhd_batch_sync": (_I, [_P]),
    "mhd_batch_dts": (_I, [_P, _I, _DP, _I, _IP]),
    "mhd_batch_set_schedule": (_I, [_P, _I, _I]),
    "mhd_batch_autotune": (_I, [_P, _I, _IP, _IP, _FP]),
    "mhd_server_create": (_I, [ctypes.POINTER(_P), _I, _SZ, _I, ctypes.c_longlong]),
    "mhd_server_submit": (_I, [_P, _I, _I, _I, _I, _P, _P, _P, _P, ctypes.POINTER(ctypes.c_longlong)]),
    "mhd_server_poll": (_I, [_P, ctypes.POINTER(ctypes.c_longlong), _DP, _DP, _IP, _I, _IP]),
    "mhd_server_pending": (_I, [_P, _IP, _IP]),
    "mhd_server_stats": (_I, [_P, _DP]),
    "mhd_server_timing": (_I, [_P, _DP]),
    "mhd_server_destroy": (_I, [_P]),
}


class MHDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    """The loaded libmhdgpu.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MHDError(-3, f"{LIB_PATH} not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != MHD_OK:
        raise MHDError(rc, lib().mhd_last_error().decode(errors="replace"))
    return rc


def _iref(v):
    return ctypes.byref(ctypes.c_int(int(v)))


class DeviceArray:
    """A device allocation made by the library, with host copies in and out."""

    def __init__(self, nbytes, dtype=np.float32):
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        p = ctypes.c_void_p()
        check(lib().mhd_dev_alloc(ctypes.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes, a.dtype)
        check(lib().mhd_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes))
        return d

    def to_host(self, count=None):
        n = self.nbytes // self.dtype.itemsize if count is None else count
        out = np.empty(n, self.dtype)
        check(lib().mhd_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def copy_from(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        check(lib().mhd_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes))

    def free(self):
        if self.ptr:
            lib().mhd_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---- step-level operators (reference interface; device arrays, FP32) -------

def fluidx(u, b, nx, ny, nz, dt):
    """One relaxing-TVD sweep along i (Heterogeneous_MHD/C/fluidx.c:19); dt is a device scalar."""
    check(lib().mhd_fluidx(u.ptr, b.ptr, _iref(nx), _iref(ny), _iref(nz), dt.ptr))


def advectbyzx(u, b, nx, ny, nz, dt):
    """CT advection of b2, b3 along i (Heterogeneous_MHD/C/advectbyzx.c:14)."""
    check(lib().mhd_advectbyzx(u.ptr, b.ptr, _iref(nx), _iref(ny), _iref(nz), dt.ptr))


def advectbyzx_noalloc(u, b, nx, ny, nz, dt):
    """advectbyzx without device allocation (fixed per-device scratch, at most 128^3 cells)."""
    check(lib().mhd_advectbyzx_noalloc(u.ptr, b.ptr, _iref(nx), _iref(ny), _iref(nz), dt.ptr))


def transpose12(ut, bt, u, b, nx, ny, nz):
    """(i,j,k)->(j,i,k); afterwards u, b hold the (ny,nx,nz) frame (transpose12.c:7)."""
    check(lib().mhd_transpose12(ut.ptr, bt.ptr, u.ptr, b.ptr, _iref(nx), _iref(ny), _iref(nz)))


def transpose13(ut, bt, u, b, nx, ny, nz):
    """(i,j,k)->(k,j,i); afterwards u, b hold the (nz,ny,nx) frame (transpose13.c:9)."""
    check(lib().mhd_transpose13(ut.ptr, bt.ptr, u.ptr, b.ptr, _iref(nx), _iref(ny), _iref(nz)))


def transpose_host(which, u, b, nx, ny, nz):
    """The Fortran-callable host-array transposes (cuda_transpose12_,
    HIP/cu/cuda_transpose12.cu:69): returns (ut, bt); u, b are unchanged."""
    u = np.ascontiguousarray(u, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    ut, bt = np.empty_like(u), np.empty_like(b)
    fn = lib().mhd_transpose12_ if which == 12 else lib().mhd_transpose13_
    check(fn(ut.ctypes.data, bt.ctypes.data, u.ctypes.data, b.ctypes.data, _iref(nx), _iref(ny), _iref(nz),
             ctypes.byref(ctypes.c_float(0.0))))
    return ut, bt


def cfl(u, b, nx, ny, nz, d_cfl, avg_b=False):
    """1/max fast speed (calcfl.c:13); written to the device scalar d_cfl, returned as float."""
    h = ctypes.c_float()
    check(lib().mhd_cfl(u.ptr, b.ptr, _iref(nx), _iref(ny), _iref(nz), int(avg_b), d_cfl.ptr, ctypes.byref(h)))
    return h.value


def init(path, n, order=0):
    """Read an initial-state file (8 float32 per cell) into host (u, b) (init.c:11)."""
    nc = n ** 3
    u = np.empty(5 * nc, np.float32)
    b = np.empty(3 * nc, np.float32)
    check(lib().mhd_init(str(path).encode(), int(order), u.ctypes.data, b.ctypes.data, _iref(n), _iref(n),
                         _iref(n)))
    return u, b


def write_result(path_u, path_b, u, b, nx, ny, nz):
    u = np.ascontiguousarray(u, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    check(lib().mhd_write_result(str(path_u).encode(), str(path_b).encode(), u.ctypes.data, b.ctypes.data,
                                 _iref(nx), _iref(ny), _iref(nz)))


def main(u, b, nx, ny, nz, nsteps, tf=math.inf):
    """Program-level loop (cuda_main, HIP/cu/cuda_main.cu:14) on host arrays, in place.
    Returns the dt sequence."""
    cap = max(int(nsteps), 1) if nsteps > 0 else 1 << 16
    dts = np.zeros(cap, np.float32)
    done = ctypes.c_int()
    check(lib().mhd_main(u.ctypes.data, b.ctypes.data, _iref(nx), _iref(ny), _iref(nz), int(nsteps),
                         ctypes.c_float(tf), dts.ctypes.data_as(_FP), cap, ctypes.byref(done)))
    return dts[:min(done.value, cap)]


def mem_info():
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    check(lib().mhd_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


def selfcheck():
    """max |1 - D*rcp(D)| of the FP64 reciprocal instruction over every
    float-valued D: the bound the FP32 limiter quotient relies on."""
    e = ctypes.c_double()
    check(lib().mhd_selfcheck(ctypes.byref(e)))
    return e.value


def selfcheck_division(stride=1):
    """(quotient mismatches, rcp equivariance mismatches) of the FP32
    shared-divisor division: one-correction quotients against IEEE a/d over
    every dividend mantissa and the divisor mantissas 0, stride, 2*stride,
    ...; stride 1 covers all 2^46 pairs.  Both must be 0."""
    m = (ctypes.c_ulonglong * 2)()
    check(lib().mhd_selfcheck_division(int(stride), m))
    return int(m[0]), int(m[1])


def selfcheck_sqrt():
    """mismatches of the shortened FP64 square root of the freezing speed
    against the IEEE root over every positive finite float; must be 0."""
    m = ctypes.c_ulonglong()
    check(lib().mhd_selfcheck_sqrt(ctypes.byref(m)))
    return int(m.value)


def selfcheck_limiter(p, m):
    """(out, out_nb): the FP32 van Leer limiter of the sweeps on the device
    for the pairs (p[i], m[i]), in its branching form (tiles) and its
    branch-free form (streamed tvdb)."""
    p = np.ascontiguousarray(p, dtype=np.float32)
    m = np.ascontiguousarray(m, dtype=np.float32)
    if p.shape != m.shape or p.ndim != 1:
        raise ValueError("selfcheck_limiter: two 1-D arrays of one length")
    out, out_nb = np.empty_like(p), np.empty_like(p)
    check(lib().mhd_selfcheck_limiter(p.ctypes.data, m.ctypes.data, out.ctypes.data, out_nb.ctypes.data, p.size))
    return out, out_nb


class Simulation:
    """A device-resident box advanced by the library's captured timestep."""

    def __init__(self, nx, ny, nz, precision=4, avg_b=False):
        self.shape = (nx, ny, nz)
        self.precision = precision
        self.dtype = np.float32 if precision == 4 else np.float64
        p = ctypes.c_void_p()
        check(lib().mhd_ctx_create(ctypes.byref(p), nx, ny, nz, precision, int(avg_b)))
        self.ctx = p.value
        self.ncell = nx * ny * nz

    def set_state(self, u, b):
        u = np.ascontiguousarray(u, self.dtype)
        b = np.ascontiguousarray(b, self.dtype)
        assert u.size == 5 * self.ncell and b.size == 3 * self.ncell
        check(lib().mhd_ctx_set_state(self.ctx, u.ctypes.data, b.ctypes.data))

    def get_state(self):
        u = np.empty(5 * self.ncell, self.dtype)
        b = np.empty(3 * self.ncell, self.dtype)
        check(lib().mhd_ctx_get_state(self.ctx, u.ctypes.data, b.ctypes.data))
        return u, b

    def reset_clock(self, tf=math.inf):
        check(lib().mhd_ctx_reset_clock(self.ctx, float(tf)))

    # kernel form of the y/z advectbyzx (identical results in every form)
    ADVECT_FORMS = {"auto": 0, "tiles": 1, "two_pass": 2, "fused": 3}

    def set_advect_form(self, form="auto"):
        check(lib().mhd_ctx_set_advect_form(self.ctx, self.ADVECT_FORMS[form]))

    def step(self, nsteps=1):
        check(lib().mhd_ctx_step(self.ctx, int(nsteps)))

    def step_timed(self, nsteps):
        ms = ctypes.c_float()
        check(lib().mhd_ctx_step_timed(self.ctx, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def sync(self):
        check(lib().mhd_ctx_sync(self.ctx))

    def dts(self, maxn=1 << 20):
        cnt = ctypes.c_int()
        out = np.zeros(maxn, np.float64)
        check(lib().mhd_ctx_dts(self.ctx, out.ctypes.data_as(_DP), maxn, ctypes.byref(cnt)))
        return out[:min(cnt.value, maxn)].astype(self.dtype)

    def time(self):
        t, t64 = ctypes.c_double(), ctypes.c_double()
        check(lib().mhd_ctx_time(self.ctx, ctypes.byref(t), ctypes.byref(t64)))
        return t.value, t64.value

    def monitors(self):
        out = np.zeros(6, np.float64)
        check(lib().mhd_ctx_monitors(self.ctx, out.ctypes.data_as(_DP)))
        return out

    def monitor_history(self, maxn=1 << 16):
        """FP64 monitors recorded after every step: rows of (mass, energy,
        momentum x, y, z, max |div B|)."""
        cnt = ctypes.c_int()
        out = np.zeros((maxn, 6), np.float64)
        check(lib().mhd_ctx_monitor_history(self.ctx, out.ctypes.data_as(_DP), maxn, ctypes.byref(cnt)))
        return out[:min(cnt.value, maxn)]

    def profile_step(self, reps=3):
        out = (ctypes.c_float * 13)()
        check(lib().mhd_ctx_profile_step(self.ctx, int(reps), out))
        return np.array(out[:], np.float64)

    def close(self):
        if self.ctx:
            lib().mhd_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent boxes advanced together (the batched form), FP32."""

    def __init__(self, dims, avg_b=False):
        self.dims = [tuple(int(x) for x in d) for d in dims]
        flat = (ctypes.c_int * (3 * len(self.dims)))(*[x for d in self.dims for x in d])
        p = ctypes.c_void_p()
        check(lib().mhd_batch_create(ctypes.byref(p), len(self.dims), flat, int(avg_b)))
        self.bt = p.value

    def ncell(self, i):
        nx, ny, nz = self.dims[i]
        return nx * ny * nz

    def set_state(self, i, u, b):
        u = np.ascontiguousarray(u, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        assert u.size == 5 * self.ncell(i) and b.size == 3 * self.ncell(i)
        check(lib().mhd_batch_set_state(self.bt, i, u.ctypes.data, b.ctypes.data))

    def get_state(self, i):
        u = np.empty(5 * self.ncell(i), np.float32)
        b = np.empty(3 * self.ncell(i), np.float32)
        check(lib().mhd_batch_get_state(self.bt, i, u.ctypes.data, b.ctypes.data))
        return u, b

    def step(self, n=1):
        check(lib().mhd_batch_step(self.bt, int(n)))

    def step_timed(self, n):
        ms = ctypes.c_float()
        check(lib().mhd_batch_step_timed(self.bt, int(n), ctypes.byref(ms)))
        return ms.value

    def sync(self):
        check(lib().mhd_batch_sync(self.bt))

    def dts(self, i, maxn=1 << 14):
        cnt = ctypes.c_int()
        out = np.zeros(maxn, np.float64)
        check(lib().mhd_batch_dts(self.bt, i, out.ctypes.data_as(_DP), maxn, ctypes.byref(cnt)))
        return out[:min(cnt.value, maxn)].astype(np.float32)

    def set_schedule(self, blocks_per_cu, policy):
        check(lib().mhd_batch_set_schedule(self.bt, int(blocks_per_cu), int(policy)))

    def autotune(self, trial_steps=3):
        bpc, pol, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        check(lib().mhd_batch_autotune(self.bt, int(trial_steps), ctypes.byref(bpc), ctypes.byref(pol),
                                       ctypes.byref(ms)))
        return {"blocks_per_cu": bpc.value, "policy": pol.value, "ms_per_step": ms.value}

    def close(self):
        if self.bt:
            lib().mhd_batch_destroy(self.bt)
            self.bt = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The request loop: mode 0 continuous batching, mode 1 FIFO one at a time."""

    def __init__(self, mode=0, hbm_budget=0, max_active=256, max_active_cells=0):
        p = ctypes.c_void_p()
        check(lib().mhd_server_create(ctypes.byref(p), int(mode), int(hbm_budget), int(max_active),
                                      int(max_active_cells)))
        self.sv = p.value
        self._keep = {}

    def submit(self, dims, nsteps, u, b):
        """Queue a request; returns (id, (out_u, out_b)) - the outputs are
        filled when the request shows up in poll()."""
        u = np.ascontiguousarray(u, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        ou, ob = np.empty_like(u), np.empty_like(b)
        rid = ctypes.c_longlong()
        check(lib().mhd_server_submit(self.sv, dims[0], dims[1], dims[2], int(nsteps), u.ctypes.data, b.ctypes.data,
                                      ou.ctypes.data, ob.ctypes.data, ctypes.byref(rid)))
        self._keep[rid.value] = (u, b, ou, ob)
        return rid.value, (ou, ob)

    def poll(self, maxn=1024):
        ids = (ctypes.c_longlong * maxn)()
        lat = np.zeros(maxn)
        qs = np.zeros(maxn)
        rcs = (ctypes.c_int * maxn)()
        n = ctypes.c_int()
        check(lib().mhd_server_poll(self.sv, ids, lat.ctypes.data_as(_DP), qs.ctypes.data_as(_DP), rcs, maxn,
                                    ctypes.byref(n)))
        out = []
        for i in range(n.value):
            out.append({"id": ids[i], "latency_s": float(lat[i]), "queue_s": float(qs[i]), "rc": rcs[i]})
            self._keep.pop(ids[i], None)
        return out

    def pending(self):
        q, a = ctypes.c_int(), ctypes.c_int()
        check(lib().mhd_server_pending(self.sv, ctypes.byref(q), ctypes.byref(a)))
        return q.value, a.value

    def stats(self):
        out = np.zeros(7)
        check(lib().mhd_server_stats(self.sv, out.ctypes.data_as(_DP)))
        keys = ["zone_updates", "span_s", "gpu_busy_s", "steps", "rebuilds", "resident_bytes", "budget_bytes"]
        d = dict(zip(keys, out.tolist()))
        t = np.zeros(4)
        check(lib().mhd_server_timing(self.sv, t.ctypes.data_as(_DP)))
        d.update(zip(["t_upload_s", "t_rebuild_s", "t_step_s", "t_download_s"], t.tolist()))
        return d

    def close(self):
        if self.sv:
            lib().mhd_server_destroy(self.sv)
            self.sv = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
