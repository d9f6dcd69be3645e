"""ctypes binding of libphpc's C ABI (include/phpc_gemm.h, include/phpc_summa.h, ...).

This is the harness the tests and the benchmark drive the library through: every call
goes through the exported C symbols, exactly as a C caller of the reference's
phpc_gemm.cuh / phpc_summa.h entry points would make it. Host arrays are numpy
float64 C-contiguous arrays; device buffers are torch CUDA tensors (HIP on ROCm),
passed as raw pointers.

Loading never falls back to anything else: if lib/libphpc.so is missing the import of
`lib()` raises.
"""
from __future__ import annotations

import ctypes
import os
import re
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
LIB_PATH = HERE / "lib" / "libphpc.so"
BIN_PATH = HERE / "bin" / "main.out"
INCLUDE = ROOT / "include"

PHPC_OK = 0
PHPC_ERR_ARG = 1
PHPC_ERR_HIP = 2
PHPC_ERR_NODEV = 3
PHPC_ERR_MEM = 4
PHPC_ERR_VENDOR = 5
PHPC_ERR_BUDGET = 6

SUMMA_FUSED = 0
SUMMA_STEPWISE = 1
SUMMA_VENDOR = 2
SUMMA_CHAINED = 3
SUMMA_SPLIT = 4

DIST_UNIFORM = 0
DIST_WIDE = 1
DIST_RAMP = 2

_lib = None

c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_int_p = ctypes.POINTER(ctypes.c_int)
vp = ctypes.c_void_p
L = ctypes.c_long
I = ctypes.c_int

class SplitInfo(ctypes.Structure):
    _fields_ = [("slices", ctypes.c_int), ("products", ctypes.c_int), ("launches", ctypes.c_int),
                ("error_estimate", ctypes.c_double), ("method", ctypes.c_int), ("int_bits", ctypes.c_int)]


class GemmItem(ctypes.Structure):
    _fields_ = [("A", ctypes.c_void_p), ("B", ctypes.c_void_p), ("C", ctypes.c_void_p),
                ("lda", ctypes.c_long), ("ldb", ctypes.c_long), ("ldc", ctypes.c_long),
                ("m", ctypes.c_int), ("n", ctypes.c_int), ("k", ctypes.c_int), ("panel", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("accuracy", ctypes.c_int), ("m", ctypes.c_int), ("n", ctypes.c_int), ("k", ctypes.c_int),
                ("dist", ctypes.c_int), ("seed", ctypes.c_uint64), ("result_host", ctypes.c_void_p)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("batch_size", ctypes.c_int), ("latency_s", ctypes.c_double),
                ("queue_s", ctypes.c_double), ("compute_s", ctypes.c_double), ("error_estimate", ctypes.c_double),
                ("fill_s", ctypes.c_double), ("gemm_s", ctypes.c_double)]


GEMM_T = ctypes.CFUNCTYPE(None, c_double_p, I, c_double_p, I, c_double_p, I, I, I, I, I, I, I, I, c_float_p)

_SIGS = {
    # include/phpc_gemm.h
    "phpc_gemm_cuda": (None, [vp, I, vp, I, vp, I, I, I, I, I, I, I, I, c_float_p]),
    "phpc_gemm_cublas": (None, [vp, I, vp, I, vp, I, I, I, I, I, I, I, I, c_float_p]),
    "phpc_last_error": (I, []),
    "phpc_last_error_message": (ctypes.c_char_p, []),
    "phpc_init": (I, []),
    "phpc_kernel_timing": (I, [I]),
    "phpc_kernel_timing_read": (I, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(I)]),
    "phpc_dgemm_device": (I, [vp, L, vp, L, vp, L, I, I, I, I, vp]),
    "phpc_dgemm_vendor_device": (I, [vp, L, vp, L, vp, L, I, I, I, vp]),
    "phpc_set_launch_policy": (I, [I, I, I]),
    "phpc_num_configs": (I, []),
    "phpc_set_tuning": (I, [I, I]),
    "phpc_config_info": (I, [I, c_int_p, c_int_p, c_int_p]),
    "phpc_fill_device": (I, [vp, L, I, I, I, ctypes.c_uint64, vp]),
    "phpc_fill_device_f32": (I, [vp, L, I, I, I, ctypes.c_uint64, vp]),
    "phpc_check_sampled": (I, [vp, L, vp, L, vp, L, vp, L, I, I, I, c_int_p, c_int_p, I,
                               c_double_p, c_double_p, vp]),
    "phpc_check_sampled_host": (I, [vp, L, vp, L, vp, L, I, I, I, I, ctypes.c_uint64, c_double_p, c_double_p]),
    # include/phpc_summa.h
    "phpc_grid_create": (I, [I, ctypes.POINTER(vp)]),
    "phpc_grid_create_dims": (I, [I, I, ctypes.POINTER(vp)]),
    "phpc_grid_destroy": (None, [vp]),
    "phpc_grid_dims": (I, [vp, c_int_p]),
    "phpc_grid_set_panel": (I, [vp, I]),
    "phpc_grid_get_panel": (I, [vp]),
    "phpc_grid_set_mode": (I, [vp, I]),
    "find_lcm": (I, [I, I]),
    "phpc_gemm_summa_cuda": (None, [vp, vp, vp, vp, I, I, I, I, I, c_float_p]),
    "phpc_gemm_summa_cublas": (None, [vp, vp, vp, vp, I, I, c_float_p]),
    "phpc_gemm_summa": (I, [GEMM_T, vp, vp, vp, vp, I, I, I, I, I, c_float_p]),
    "phpc_summa_device": (I, [I, I, I, I, vp, vp, vp, I, I, vp]),
    "phpc_summa_host": (I, [I, I, I, I, vp, vp, vp, I, I, c_float_p]),
    "phpc_pack_device": (I, [vp, L, vp, L, I, I, vp]),
    # include/phpc_split.h
    "phpc_dgemm_split_device": (I, [vp, L, vp, L, vp, L, I, I, I, I, ctypes.c_double, vp,
                                    ctypes.POINTER(SplitInfo)]),
    "phpc_dgemm_split_device_noalloc": (I, [vp, L, vp, L, vp, L, I, I, I, I, ctypes.c_double, vp,
                                            ctypes.POINTER(SplitInfo)]),
    "phpc_split_error_estimate": (ctypes.c_double, [I, ctypes.c_double, ctypes.c_double, I]),
    "phpc_split_set_config": (I, [I]),
    "phpc_split_num_configs": (I, []),
    "phpc_split_set_method": (I, [I]),
    "phpc_sgemm_device": (I, [vp, L, vp, L, vp, L, I, I, I, I, vp]),
    # include/phpc_grouped.h
    "phpc_plan_create": (I, [I, ctypes.POINTER(GemmItem), I, I, vp, ctypes.POINTER(vp)]),
    "phpc_plan_launch": (I, [vp, vp]),
    "phpc_plan_graph_launch": (I, [vp, vp]),
    "phpc_plan_info": (I, [vp, c_int_p, c_int_p, c_int_p, c_double_p]),
    "phpc_plan_destroy": (None, [vp]),
    # include/phpc_serve.h
    "phpc_server_create": (I, [ctypes.c_size_t, I, I, I, ctypes.POINTER(vp)]),
    "phpc_server_submit": (I, [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
    "phpc_server_wait": (I, [vp, ctypes.c_uint64, ctypes.POINTER(Result)]),
    "phpc_server_drain": (I, [vp]),
    "phpc_server_set_split_min": (I, [vp, I]),
    "phpc_server_set_policy": (I, [vp, I, ctypes.c_double]),
    "phpc_server_set_big_limit": (I, [vp, I, ctypes.c_double]),
    "phpc_server_memory": (I, [vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "phpc_server_destroy": (None, [vp]),
    # include/utils.h
    "get_cur_time": (ctypes.c_double, []),
}


def lib():
    """Load libphpc.so (built by `make -C hpc-multigpu-matrixmult_amd`). Raises if absent."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"libphpc is not built: {LIB_PATH} missing (run __graft_entry__.build())")
        # torch ships its own HIP runtime under the same soname. Loading torch first makes
        # libphpc bind to that copy, so a process holds one HIP runtime; loading libphpc
        # first would bring in the system copy and torch would then load a second one.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        h = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def declared_functions():
    """Every function name declared in include/*.h (the C ABI surface)."""
    names = set()
    for h in sorted(INCLUDE.glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{]*?)\)\s*;", text):
            name = m.group(1)
            before = text[: m.start()].rstrip()
            if name in ("if", "while", "for", "return", "sizeof"):
                continue
            # skip function-pointer typedefs: "typedef void (*gemm_t)(...)"
            if before.endswith("(*") or before.endswith("typedef"):
                continue
            names.add(name)
    names.discard("gemm_t")
    return names


def check(rc: int):
    if rc != PHPC_OK:
        msg = lib().phpc_last_error_message()
        raise RuntimeError(f"libphpc error {rc}: {msg.decode() if msg else ''}")


def last_error():
    l = lib()
    msg = l.phpc_last_error_message()
    return l.phpc_last_error(), (msg.decode() if msg else "")


# ---------------- host-pointer helpers (numpy) ----------------

def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def gemm_cuda(a, b, c, m=None, k=None, n=None, lda=None, ldb=None, ldc=None, gpu_count=1,
              grid_width=8, grid_height=8, block_width=32):
    """phpc_gemm_cuda on numpy arrays (C updated in place). Returns kernel seconds."""
    m = a.shape[0] if m is None else m
    k = a.shape[1] if k is None else k
    n = b.shape[1] if n is None else n
    lda = a.shape[1] if lda is None else lda
    ldb = b.shape[1] if ldb is None else ldb
    ldc = c.shape[1] if ldc is None else ldc
    t = ctypes.c_float(0)
    lib().phpc_gemm_cuda(_ptr(a), lda, _ptr(b), ldb, _ptr(c), ldc, m, k, n, gpu_count,
                         grid_width, grid_height, block_width, ctypes.byref(t))
    return t.value


def gemm_cublas(a, b, c, gpu_count=1):
    m, k = a.shape
    n = b.shape[1]
    t = ctypes.c_float(0)
    lib().phpc_gemm_cublas(_ptr(a), k, _ptr(b), n, _ptr(c), n, m, k, n, gpu_count, 0, 0, 0,
                           ctypes.byref(t))
    return t.value


class Grid:
    """Owning wrapper of phpc_grid_t."""

    def __init__(self, nranks=None, dims=None, panel=0, mode=SUMMA_FUSED):
        h = vp()
        if dims is not None:
            check(lib().phpc_grid_create_dims(dims[0], dims[1], ctypes.byref(h)))
        else:
            check(lib().phpc_grid_create(nranks, ctypes.byref(h)))
        self.h = h
        check(lib().phpc_grid_set_panel(h, panel))
        check(lib().phpc_grid_set_mode(h, mode))

    def dims(self):
        d = (ctypes.c_int * 2)()
        check(lib().phpc_grid_dims(self.h, d))
        return d[0], d[1]

    def close(self):
        if self.h:
            lib().phpc_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def summa_cuda(grid: Grid, A, B, C, gw=8, gh=8, bw=32):
    t = ctypes.c_float(0)
    n = A.shape[0]
    lib().phpc_gemm_summa_cuda(grid.h, _ptr(A), _ptr(B), _ptr(C), n, 1, gw, gh, bw, ctypes.byref(t))
    return t.value


def summa_cublas(grid: Grid, A, B, C):
    t = ctypes.c_float(0)
    lib().phpc_gemm_summa_cublas(grid.h, _ptr(A), _ptr(B), _ptr(C), A.shape[0], 1, ctypes.byref(t))
    return t.value


# ---------------- device helpers (torch tensors) ----------------

def dptr(t):
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def fill_device(x, dist, seed, stream=None):
    """Fill a float64 or float32 device matrix with a generated stream (phpc_fill_device)."""
    import torch
    m, n = x.shape
    fn = lib().phpc_fill_device_f32 if x.dtype == torch.float32 else lib().phpc_fill_device
    check(fn(dptr(x), x.stride(0), m, n, dist, seed, stream_ptr(stream)))


def dgemm_device(A, B, C, panel=0, stream=None, m=None, k=None, n=None):
    m = A.shape[0] if m is None else m
    k = A.shape[1] if k is None else k
    n = B.shape[1] if n is None else n
    check(lib().phpc_dgemm_device(dptr(A), A.stride(0), dptr(B), B.stride(0), dptr(C), C.stride(0),
                                  m, k, n, panel, stream_ptr(stream)))


def dgemm_vendor_device(A, B, C, stream=None):
    m, k = A.shape
    n = B.shape[1]
    check(lib().phpc_dgemm_vendor_device(dptr(A), A.stride(0), dptr(B), B.stride(0), dptr(C),
                                         C.stride(0), m, k, n, stream_ptr(stream)))


def summa_device(A, B, C, p, q, panel=0, mode=SUMMA_FUSED, wg=0, stream=None):
    check(lib().phpc_summa_device(p, q, panel, mode, dptr(A), dptr(B), dptr(C), A.shape[0], wg,
                                  stream_ptr(stream)))


def check_sampled(A, B, C, rows, cols, C0=None, stream=None):
    import numpy as np
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    cw = ctypes.c_double(0)
    nw = ctypes.c_double(0)
    c0p = dptr(C0) if C0 is not None else None
    ldc0 = C0.stride(0) if C0 is not None else 0
    check(lib().phpc_check_sampled(dptr(A), A.stride(0), dptr(B), B.stride(0), c0p, ldc0, dptr(C),
                                   C.stride(0), A.shape[0], A.shape[1], B.shape[1],
                                   rows.ctypes.data_as(c_int_p), cols.ctypes.data_as(c_int_p),
                                   len(rows), ctypes.byref(cw), ctypes.byref(nw), stream_ptr(stream)))
    return cw.value, nw.value


def kernel_timing(on=True):
    """Start (or stop) event timing of the main GEMM kernel launches; clears earlier records."""
    check(lib().phpc_kernel_timing(1 if on else 0))


def kernel_timing_read():
    """(summed milliseconds, launch count) of the GEMM kernel launches timed since the last
    read; waits for them to finish."""
    ms, n = ctypes.c_double(), I()
    check(lib().phpc_kernel_timing_read(ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


def set_launch_policy(honor_grid=0, config_id=-1, grid=0):
    check(lib().phpc_set_launch_policy(honor_grid, config_id, grid))


def num_configs():
    return lib().phpc_num_configs()


TUNE_F32_TRANSPOSED_A = 6
TUNE_F32_PLAN_CONFIG = 7


def set_tuning(key, value):
    """One tuning knob (phpc_set_tuning): keys 0-5 pick FP64 configs, key 6 switches large
    single FP32 products between the transposed-A kernel (1) and the register-staged one (0),
    key 7 sets the FP32 configuration of plans created without autotuning (-1 = by work)."""
    check(lib().phpc_set_tuning(key, value))


SPLIT_SLICES = 0
SPLIT_CRT = 1
SPLIT_FP8 = 2


def dgemm_split_device(A, B, C, slices=0, target=1e-14, stream=None, method=None):
    """Split-precision C += A*B on torch float64 CUDA tensors; returns SplitInfo."""
    if method is not None:
        check(lib().phpc_split_set_method(method))
    info = SplitInfo()
    m, k = A.shape
    n = B.shape[1]
    check(lib().phpc_dgemm_split_device(dptr(A), A.stride(0), dptr(B), B.stride(0), dptr(C), C.stride(0),
                                        m, k, n, slices, target, stream_ptr(stream), ctypes.byref(info)))
    return info


F64 = 0
F32 = 1
F32X3 = 2
SGEMM_FP32 = 0
SGEMM_BF16X3 = 1


def sgemm_device(A, B, C, method=1, stream=None):
    """FP32 C += A*B on torch float32 CUDA tensors (method SGEMM_FP32 or SGEMM_BF16X3)."""
    m, k = A.shape
    n = B.shape[1]
    check(lib().phpc_sgemm_device(dptr(A), A.stride(0), dptr(B), B.stride(0), dptr(C), C.stride(0), m, k, n,
                                  method, stream_ptr(stream)))
PLAN_AUTOTUNE = 1


class Plan:
    """Owning wrapper of phpc_plan_t over torch CUDA tensors (A, B, C[, panel]) tuples."""

    def __init__(self, dtype, gemms, autotune=False, stream=None):
        items = (GemmItem * max(1, len(gemms)))()
        self._keep = []
        for i, g in enumerate(gemms):
            A, B, C = g[0], g[1], g[2]
            panel = g[3] if len(g) > 3 else 0
            items[i] = GemmItem(A.data_ptr(), B.data_ptr(), C.data_ptr(), A.stride(0), B.stride(0), C.stride(0),
                                A.shape[0], B.shape[1], A.shape[1], panel)
            self._keep.append((A, B, C))
        h = vp()
        check(lib().phpc_plan_create(dtype, items, len(gemms), PLAN_AUTOTUNE if autotune else 0,
                                     stream_ptr(stream), ctypes.byref(h)))
        self.h = h

    def launch(self, stream=None):
        check(lib().phpc_plan_launch(self.h, stream_ptr(stream)))

    def graph_launch(self, stream):
        check(lib().phpc_plan_graph_launch(self.h, stream_ptr(stream)))

    def info(self):
        c, t, g, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(lib().phpc_plan_info(self.h, ctypes.byref(c), ctypes.byref(t), ctypes.byref(g), ctypes.byref(ms)))
        return {"config": c.value, "tiles": t.value, "grid": g.value, "tuned_ms": ms.value}

    def close(self):
        if self.h:
            lib().phpc_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------- serving loop (include/phpc_serve.h) ----------------

ACC_FP64 = 0
ACC_FP32 = 1
SCHED_FIFO = 0
SCHED_SJF = 1
SCHED_LANES = 2


class Server:
    """Owning wrapper of phpc_server_t. submit() returns a ticket; wait(ticket) blocks
    (without holding the GIL) and returns the result as a dict."""

    def __init__(self, budget_bytes, slots=2, max_dim=4096, batch_max=16):
        h = vp()
        check(lib().phpc_server_create(int(budget_bytes), slots, max_dim, batch_max, ctypes.byref(h)))
        self.h = h

    def submit(self, accuracy, m, n, k, dist=DIST_UNIFORM, seed=1, out=None):
        """out: optional host numpy array (float64 for ACC_FP64, float32 for ACC_FP32, m x n)."""
        req = Request(accuracy, m, n, k, dist, seed, out.ctypes.data if out is not None else None)
        t = ctypes.c_uint64()
        check(lib().phpc_server_submit(self.h, ctypes.byref(req), ctypes.byref(t)))
        return t.value

    def wait(self, ticket):
        r = Result()
        rc = lib().phpc_server_wait(self.h, ticket, ctypes.byref(r))
        check(rc)
        return {"status": r.status, "batch_size": r.batch_size, "latency_s": r.latency_s, "queue_s": r.queue_s,
                "compute_s": r.compute_s, "error_estimate": r.error_estimate, "fill_s": r.fill_s,
                "gemm_s": r.gemm_s}

    def drain(self):
        check(lib().phpc_server_drain(self.h))

    def set_split_min(self, dim):
        check(lib().phpc_server_set_split_min(self.h, dim))

    def set_policy(self, policy, aging=0.0):
        check(lib().phpc_server_set_policy(self.h, policy, aging))

    def set_big_limit(self, limit, min_seconds=0.0):
        check(lib().phpc_server_set_big_limit(self.h, limit, min_seconds))

    def memory(self):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().phpc_server_memory(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def close(self):
        if self.h:
            lib().phpc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
