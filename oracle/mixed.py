"""CPU emulation of the mixed-precision K build (test infrastructure only).

The reference builds K in FP64 only (build_Fock.c:201-347).  The GPU library's
mixed mode forms the same two products,

    X_i = C_occ(partners of i)^T B(rows of i)      (build_Fock.c:249-285)
    K   = X X^T                                     (build_Fock.c:317-341)

from FP32 operands split as x = hi + lo with hi = bf16(x), lo = bf16(x - hi)
(round to nearest even), each product taken as hi*hi + hi*lo + lo*hi.  Here
the split is emulated bit-exactly and the split products are summed in FP64,
so the result differs from the GPU's only by the GPU's FP32 accumulation;
the FP64 K of the oracle measures the method's own error.

Pinned by: test_oracle_mixed.py (bf16 rounding against hand-worked bit
patterns and ties-to-even cases; hi + lo against the float32 value; the
three-term product against the exact product within its 2^-16 bound).
"""
import numpy as np


def bf16_round(x):
    """float32 -> nearest bf16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def split(x):
    """x (any float) -> (hi, lo) bf16 values as float64 arrays, x32 = float32(x)."""
    x32 = np.asarray(x, dtype=np.float32)
    hi = bf16_round(x32)
    lo = bf16_round((x32 - hi).astype(np.float32))
    return hi.astype(np.float64), lo.astype(np.float64)


def split_matmul(a, b):
    """a @ b with both operands split: hi_a hi_b + hi_a lo_b + lo_a hi_b, FP64 sums."""
    ah, al = split(a)
    bh, bl = split(b)
    return ah @ bh + ah @ bl + al @ bh


def abs_x(B_rows, mask, Cocc):
    """A_i = |C_occ(partners of i)|^T |B(rows of i)|: the magnitude scale of the sums that
    form X_i (nbf x n_occ x naux)."""
    nbf, n_occ = Cocc.shape
    A = np.zeros((nbf, n_occ, B_rows.shape[1]))
    for i in range(nbf):
        js = np.nonzero(mask[i] >= 0)[0]
        if len(js):
            A[i] = np.abs(Cocc[js]).T @ np.abs(B_rows[mask[i, js]])
    return A


def k_build(B_rows, mask, Cocc):
    """Mixed-mode K from the reference-layout DF tensor.

    B_rows: nnz x naux (rows in the CSR order of bf_pair_mask); mask: nbf x nbf
    row ids (-1 screened); Cocc: nbf x n_occ.  Returns (K, X32) with X32 the
    float32 X (nbf x n_occ x naux) that the K product splits again."""
    nbf, n_occ = Cocc.shape
    naux = B_rows.shape[1]
    X = np.zeros((nbf, n_occ, naux))
    for i in range(nbf):
        js = np.nonzero(mask[i] >= 0)[0]
        if len(js) == 0:
            continue
        X[i] = split_matmul(Cocc[js].T, B_rows[mask[i, js]])
    X32 = X.astype(np.float32)
    Xf = X32.reshape(nbf, -1)
    K = split_matmul(Xf, Xf.T)
    return K, X32
