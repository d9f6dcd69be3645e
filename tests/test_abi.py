"""The C-ABI library loads and exports every function declared in include/*.h; the
host-only C code (grid handle, lcm, the generic SUMMA loop) behaves like the reference
on the CPU. No GPU compute is called here."""
import ctypes
import subprocess

import numpy as np
import pytest

from conftest import phpc, ROOT
import oracle as O


def test_library_exports_every_declared_symbol(lib):
    declared = phpc.declared_functions()
    assert {"phpc_gemm_cuda", "phpc_gemm_cublas", "phpc_gemm_summa_cuda", "phpc_gemm_summa_cublas",
            "get_cur_time", "log_to_csv", "find_lcm"} <= declared
    out = subprocess.run(["nm", "-D", "--defined-only", str(phpc.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = sorted(declared - exported)
    assert not missing, f"declared but not exported: {missing}"


def test_binding_signatures_cover_declarations(lib):
    # every declared function except log_to_csv (FILE*) has a ctypes signature
    assert phpc.declared_functions() - set(phpc._SIGS) <= {"log_to_csv"}


def test_find_lcm_matches_oracle(lib):
    for a in range(1, 13):
        for b in range(1, 13):
            assert lib.find_lcm(a, b) == O.find_lcm(a, b)


@pytest.mark.parametrize("P", [1, 2, 3, 4, 6, 7, 8, 9, 12, 16, 64])
def test_grid_create_matches_dims_create(lib, P):
    g = phpc.Grid(nranks=P)
    assert g.dims() == O.dims_create(P)


def test_grid_rejects_bad_arguments(lib):
    h = ctypes.c_void_p()
    assert lib.phpc_grid_create(0, ctypes.byref(h)) == phpc.PHPC_ERR_ARG
    assert lib.phpc_grid_create_dims(2, 0, ctypes.byref(h)) == phpc.PHPC_ERR_ARG
    g = phpc.Grid(dims=(2, 2))
    assert lib.phpc_grid_set_panel(g.h, -1) == phpc.PHPC_ERR_ARG
    assert lib.phpc_grid_set_mode(g.h, 7) == phpc.PHPC_ERR_ARG


def test_get_cur_time_is_wall_clock(lib):
    import time
    assert abs(lib.get_cur_time() - time.time()) < 1.0


def _oracle_gemm_t():
    """A gemm_t that runs the oracle's emulation of the reference kernel, so the C
    SUMMA loop can be exercised on the CPU."""
    calls = []

    def f(a, lda, b, ldb, c, ldc, m, k, n, gpu_count, gw, gh, bw, t):
        calls.append((m, k, n, lda, ldb, ldc))
        O.lib().orc_kernel_emul(a, lda, b, ldb, c, ldc, m, k, n)
        t[0] = 0.5

    return phpc.GEMM_T(f), calls


@pytest.mark.parametrize("p,q,N,panel", [(1, 1, 12, 0), (2, 2, 16, 0), (4, 2, 16, 0), (2, 3, 18, 0),
                                         (3, 3, 20, 4), (4, 4, 24, 5)])
def test_host_summa_loop_matches_oracle(lib, p, q, N, panel):
    A, B, C0 = O.uniform((N, N), 1), O.uniform((N, N), 2), O.uniform((N, N), 3)
    cb, calls = _oracle_gemm_t()
    g = phpc.Grid(dims=(p, q), panel=panel)
    C = C0.copy()
    t = ctypes.c_float(0)
    rc = lib.phpc_gemm_summa(cb, g.h, phpc._ptr(A), phpc._ptr(B), phpc._ptr(C), N, 1, 8, 8, 32, ctypes.byref(t))
    assert rc == 0
    want = O.summa_emul(A, B, C0.copy(), p, q, panel)
    assert np.array_equal(C.view(np.uint64), want.view(np.uint64))
    w = panel if panel else N // O.find_lcm(p, q)
    steps = -(-N // w)
    assert len(calls) == steps * p * q
    assert t.value == pytest.approx(0.5 * steps * p * q)
    # every call sees the full matrices as leading dimension (zero-copy aliases)
    assert all(c[3] == N and c[4] == N and c[5] == N for c in calls)


def test_host_summa_rejects_grid_larger_than_n(lib):
    cb, _ = _oracle_gemm_t()
    g = phpc.Grid(dims=(4, 4))
    A = np.zeros((2, 2))
    t = ctypes.c_float(0)
    assert lib.phpc_gemm_summa(cb, g.h, phpc._ptr(A), phpc._ptr(A), phpc._ptr(A), 2, 1, 1, 1, 1,
                               ctypes.byref(t)) == phpc.PHPC_ERR_ARG


def test_reference_header_name_compiles_and_links(tmp_path, lib):
    """A C file written against the reference's headers (it includes "phpc_gemm.cuh", as
    src/phpc_summa.c:5 does, and "phpc_summa.h") compiles with gcc against include/ and links
    against libphpc.so; the program only takes the entry points' addresses (no GPU call)."""
    src = tmp_path / "caller.c"
    src.write_text(
        '#include "phpc_gemm.cuh"\n'
        '#include "phpc_summa.h"\n'
        '#include <stdio.h>\n'
        'typedef void (*summa_fn)(phpc_grid_t, const double *, const double *, double *, int, int, int, int, int,\n'
        '                         float *);\n'
        'int main(void) {\n'
        '  gemm_t f = phpc_gemm_cuda, g = phpc_gemm_cublas;\n'
        '  summa_fn s = phpc_gemm_summa_cuda;\n'
        '  printf("%d\\n", f != 0 && g != 0 && s != 0 && find_lcm(4, 6) == 12);\n'
        '  return 0;\n'
        '}\n')
    exe = tmp_path / "caller"
    libdir = phpc.LIB_PATH.parent
    subprocess.run(["gcc", "-Wall", "-Werror", "-I", str(ROOT / "include"), str(src), "-o", str(exe),
                    "-L", str(libdir), "-lphpc", f"-Wl,-rpath,{libdir}", "-L/opt/rocm/lib",
                    "-Wl,-rpath,/opt/rocm/lib", "-lamdhip64"], check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=60).stdout
    assert out.strip() == "1"


def test_host_summa_loop_ignores_a_stale_error(lib):
    """An error left on this thread by an earlier call must not stop the generic loop after its
    first block: a gemm_t that records no status (this Python callback) runs every step."""
    A = np.zeros((4, 4))
    t = ctypes.c_float(0)
    # block_width 0 is invalid: the call records PHPC_ERR_ARG (or NODEV without a GPU)
    lib.phpc_gemm_cuda(phpc._ptr(A), 4, phpc._ptr(A), 4, phpc._ptr(A), 4, 4, 4, 4, 1, 1, 1, 0, ctypes.byref(t))
    assert lib.phpc_last_error() != 0
    N, p, q = 12, 2, 2
    A, B, C0 = O.uniform((N, N), 1), O.uniform((N, N), 2), O.uniform((N, N), 3)
    cb, calls = _oracle_gemm_t()
    g = phpc.Grid(dims=(p, q))
    C = C0.copy()
    rc = lib.phpc_gemm_summa(cb, g.h, phpc._ptr(A), phpc._ptr(B), phpc._ptr(C), N, 1, 8, 8, 32, ctypes.byref(t))
    assert rc == 0 and lib.phpc_last_error() == 0
    assert len(calls) == (N // (N // O.find_lcm(p, q))) * p * q
    want = O.summa_emul(A, B, C0.copy(), p, q, 0)
    assert np.array_equal(C.view(np.uint64), want.view(np.uint64))
