"""ctypes binding of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the
benchmark's CPU-baseline leg, always as the checker. The product never imports it.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref_iterative.so"
REF_BIN = HERE / "_ref" / "iterative.out"
ITER_O0 = HERE / "_build" / "iterative_O0.out"
ITER_O2 = HERE / "_build" / "iterative_O2.out"

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
I = ctypes.c_int


def build(quiet=True):
    """Compile the oracle (and oracle/_ref when the reference sources are present)."""
    out = subprocess.run(["make", "-C", str(HERE)], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    subprocess.run(["make", "-C", str(HERE), "ref"], capture_output=True, text=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        h = ctypes.CDLL(str(LIB))
        h.orc_splitmix64_at.restype = ctypes.c_uint64
        h.orc_splitmix64_at.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        for f in ("orc_fill_uniform", "orc_fill_wide"):
            getattr(h, f).argtypes = [_dp, ctypes.c_size_t, ctypes.c_uint64]
        h.orc_fill_ramp.argtypes = [_dp, ctypes.c_size_t]
        h.orc_iterative.argtypes = [_dp, _dp, _dp, I]
        h.orc_kernel_emul.argtypes = [_dp, I, _dp, I, _dp, I, I, I, I]
        h.orc_find_lcm.restype = I
        h.orc_find_lcm.argtypes = [I, I]
        h.orc_dims_create.argtypes = [I, _ip]
        h.orc_block_size.restype = I
        h.orc_block_size.argtypes = [I, I, I]
        h.orc_block_start.restype = I
        h.orc_block_start.argtypes = [I, I, I]
        h.orc_summa_emul.argtypes = [_dp, _dp, _dp, I, I, I, I]
        L_ = ctypes.c_long
        h.orc_summa_entries.argtypes = [_dp, L_, _dp, L_, _dp, L_, _ip, _ip, L_, I, I, _dp]
        h.orc_gemm_dd.argtypes = [_dp, I, _dp, I, _dp, I, _dp, I, I, I, I]
        h.orc_gemm_dd_entries.argtypes = [_dp, I, _dp, I, _dp, I, _ip, _ip, _dp, ctypes.c_size_t, I]
        h.orc_dot_q.restype = ctypes.c_double
        h.orc_dot_q.argtypes = [_dp, I, _dp, I, I]
        h.orc_absdot.restype = ctypes.c_double
        h.orc_absdot.argtypes = [_dp, I, _dp, I, I, I, I]
        h.orc_err_full.argtypes = [_dp, I, _dp, I, _dp, I, _dp, I, _dp, I, I, I, I, _dp, _dp]
        h.orc_ramp_closed.restype = ctypes.c_double
        h.orc_ramp_closed.argtypes = [I, I, I]
        h.orc_split_gemm.argtypes = [_dp, I, _dp, I, _dp, I, I, I, I, I]
        h.orc_split_digits.argtypes = [ctypes.c_double, I, I, _ip]
        h.orc_split_gemm_w.argtypes = [_dp, I, _dp, I, _dp, I, I, I, I, I, I, ctypes.c_long]
        h.orc_split_digits_w.argtypes = [ctypes.c_double, I, I, I, _ip]
        h.orc_exp_of_max.restype = I
        h.orc_exp_of_max.argtypes = [ctypes.c_double]
        h.orc_crt_gemm.argtypes = [_dp, I, _dp, I, _dp, I, I, I, I, I, I]
        h.orc_crt_digits.argtypes = [_ip, I, _ip]
        h.orc_freivalds.restype = ctypes.c_double
        h.orc_freivalds.argtypes = [_dp, _dp, _dp, I, ctypes.c_uint64]
        _lib = h
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def uniform(shape, seed):
    x = np.empty(shape, dtype=np.float64)
    lib().orc_fill_uniform(_p(x), x.size, seed)
    return x


def wide(shape, seed):
    x = np.empty(shape, dtype=np.float64)
    lib().orc_fill_wide(_p(x), x.size, seed)
    return x


def ramp(shape):
    x = np.empty(shape, dtype=np.float64)
    lib().orc_fill_ramp(_p(x), x.size)
    return x


def splitmix64_at(seed, i):
    return lib().orc_splitmix64_at(seed, i)


def iterative(A, B, C):
    """src/iterative.c:8-13 on square matrices (C updated in place)."""
    lib().orc_iterative(_p(A), _p(B), _p(C), A.shape[0])
    return C


def kernel_emul(A, B, C, m=None, k=None, n=None, lda=None, ldb=None, ldc=None):
    """Per-element arithmetic of the reference gemm_kernel (C updated in place)."""
    m = A.shape[0] if m is None else m
    k = A.shape[1] if k is None else k
    n = B.shape[1] if n is None else n
    lda = A.shape[1] if lda is None else lda
    ldb = B.shape[1] if ldb is None else ldb
    ldc = C.shape[1] if ldc is None else ldc
    lib().orc_kernel_emul(_p(A), lda, _p(B), ldb, _p(C), ldc, m, k, n)
    return C


def summa_emul(A, B, C, p, q, panel=0):
    lib().orc_summa_emul(_p(A), _p(B), _p(C), A.shape[0], p, q, panel)
    return C


def summa_entries(A, B, rows, cols, panel=0, C0=None):
    """Entries (rows[e], cols[e]) of the reference arithmetic for C0 + A B: per-panel FMA
    chains from zero added to C in panel order (panel <= 0: one chain over all of k)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    r = np.ascontiguousarray(rows, dtype=np.int32)
    c = np.ascontiguousarray(cols, dtype=np.int32)
    out = np.empty(len(r))
    c0 = None
    ldc0 = 0
    if C0 is not None:
        C0 = np.ascontiguousarray(C0, dtype=np.float64)
        c0, ldc0 = _p(C0), C0.shape[1]
    lib().orc_summa_entries(_p(A), A.shape[1], _p(B), B.shape[1], c0, ldc0,
                            r.ctypes.data_as(_ip), c.ctypes.data_as(_ip), len(r), A.shape[1], panel, _p(out))
    return out


def find_lcm(a, b):
    return lib().orc_find_lcm(a, b)


def dims_create(n):
    d = (ctypes.c_int * 2)()
    lib().orc_dims_create(n, d)
    return d[0], d[1]


def block_size(n, parts, i):
    return lib().orc_block_size(n, parts, i)


def block_start(n, parts, i):
    return lib().orc_block_start(n, parts, i)


def gemm_dd(A, B, C0=None):
    m, k = A.shape
    n = B.shape[1]
    out = np.empty((m, n), dtype=np.float64)
    c0p = _p(C0) if C0 is not None else None
    lib().orc_gemm_dd(_p(A), k, _p(B), n, c0p, n, _p(out), n, m, k, n)
    return out


def gemm_dd_entries(A, B, rows, cols, C0=None):
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    out = np.empty(len(rows), dtype=np.float64)
    c0p = _p(C0) if C0 is not None else None
    ldc0 = C0.shape[1] if C0 is not None else 0
    lib().orc_gemm_dd_entries(_p(A), A.shape[1], _p(B), B.shape[1], c0p, ldc0,
                              rows.ctypes.data_as(_ip), cols.ctypes.data_as(_ip), _p(out), len(rows),
                              A.shape[1])
    return out


def dot_q(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().orc_dot_q(_p(a), 1, _p(b), 1, len(a))


def err_full(C, Cref, A, B, C0=None):
    m, k = A.shape
    n = B.shape[1]
    cw = ctypes.c_double()
    nw = ctypes.c_double()
    c0p = _p(C0) if C0 is not None else None
    lib().orc_err_full(_p(C), n, _p(Cref), n, _p(A), k, _p(B), n, c0p, n, m, k, n,
                       ctypes.byref(cw), ctypes.byref(nw))
    return cw.value, nw.value


def ramp_closed(N, r, c):
    return lib().orc_ramp_closed(N, r, c)


def ramp_closed_matrix(N):
    """Exact C of the ramp inputs (float64, exact for N <= 1536)."""
    r = np.arange(N, dtype=object)[:, None]
    c = np.arange(N, dtype=object)[None, :]
    s1 = N * (N - 1) // 2
    s2 = (N - 1) * N * (2 * N - 1) // 6
    v = r * N * N * s1 + r * c * N * N + N * s2 + c * s1
    return v.astype(np.float64)


def freivalds(A, B, C, seed=7):
    return lib().orc_freivalds(_p(A), _p(B), _p(C), A.shape[0], seed)


def ref_iterative_available():
    return REF_LIB.exists()


def ref_iterative(A, B, C):
    """The reference's own phpc_gemm_iterative (oracle/_ref, built from its sources)."""
    h = ctypes.CDLL(str(REF_LIB))
    h.phpc_gemm_iterative.argtypes = [_dp, _dp, _dp, I]
    h.phpc_gemm_iterative(_p(A), _p(B), _p(C), A.shape[0])
    return C


def split_gemm(A, B, C, s, bits=7):
    """Restatement of the library's split-precision GEMM (C updated in place): 7-bit int8
    digits, or bits=4 for the FP8-E4M3 slices (k pieces of at most 74565 products)."""
    m, k = A.shape
    n = B.shape[1]
    if bits == 7:
        lib().orc_split_gemm(_p(A), k, _p(B), n, _p(C), n, m, k, n, s)
    else:
        lib().orc_split_gemm_w(_p(A), k, _p(B), n, _p(C), n, m, k, n, s, bits, 74565 if bits == 4 else 131072)
    return C


def split_digits(x, e, s, bits=7):
    d = (ctypes.c_int * s)()
    if bits == 7:
        lib().orc_split_digits(float(x), e, s, d)
    else:
        lib().orc_split_digits_w(float(x), e, s, bits, d)
    return list(d)


def exp_of_max(mx):
    return lib().orc_exp_of_max(float(mx))


MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179, 173]


def crt_gemm(A, B, C, p, L):
    """Restatement of the library's CRT split GEMM (C updated in place)."""
    m, k = A.shape
    n = B.shape[1]
    lib().orc_crt_gemm(_p(A), k, _p(B), n, _p(C), n, m, k, n, p, L)
    return C


def crt_digits(residues):
    L = len(residues)
    r = (ctypes.c_int * L)(*residues)
    v = (ctypes.c_int * L)()
    lib().orc_crt_digits(r, L, v)
    return list(v)
