"""Writes the committed fixtures under tests/golden/.

ref_iterative_n48.npz  A, B: uniform[-1,1) (SplitMix64 seeds 101, 102), 48 x 48;
                       C: the reference's own phpc_gemm_iterative (src/iterative.c:8-13),
                       compiled from its sources into oracle/_ref by `make -C oracle ref`,
                       applied to C = 0.
ramp_n64.npz           C of the reference's ramp inputs A = B = [0..N*N) (src/main.c:85-86),
                       N = 64, from exact Python integer arithmetic.
Run from the repository root: python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent


def main():
    if not O.ref_iterative_available():
        raise SystemExit("oracle/_ref/libref_iterative.so missing: run `make -C oracle ref` where the reference exists")
    A = O.uniform((48, 48), 101)
    B = O.uniform((48, 48), 102)
    C = O.ref_iterative(A, B, np.zeros((48, 48)))
    np.savez(OUT / "ref_iterative_n48.npz", A=A, B=B, C=C)
    N = 64
    M = np.arange(N * N, dtype=object).reshape(N, N)
    np.savez(OUT / "ramp_n64.npz", C=M.dot(M).astype(np.float64))
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
