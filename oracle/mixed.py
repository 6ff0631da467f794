"""CPU emulation of the mixed-precision K build (test infrastructure only).

The reference builds K in FP64 only (build_Fock.c:201-347).  The GPU library's
mixed mode forms the same two products,

    X_i = C_occ(partners of i)^T B(rows of i)      (build_Fock.c:249-285)
    K   = X X^T                                     (build_Fock.c:317-341)

from FP32 operands split into np BF16 parts, p_0 = bf16(x), p_1 = bf16(x - p_0),
p_2 = bf16(x - p_0 - p_1) (round to nearest even), each product taken as the
sum of the part products p_a q_b with a + b < np: np = 2 ("BF16x3", hi*hi +
hi*lo + lo*hi) or np = 3 ("BF16x6", six products).  Here the split is
emulated bit-exactly and the split products are summed in FP64, so the result
differs from the GPU's only by the GPU's FP32 accumulation; the FP64 K of the
oracle measures the method's own error.

Pinned by: test_oracle_mixed.py (bf16 rounding against hand-worked bit
patterns and ties-to-even cases; the parts of hand-worked float32 values;
hi + lo against the float32 value within 2^-17, the three parts equal to it
exactly; the three- and six-term products against the exact product within
their 2^-16 and 2^-23 bounds).
"""
import numpy as np


def bf16_round(x):
    """float32 -> nearest bf16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def split(x, nparts=2):
    """x (any float) -> its nparts bf16 parts as float64 arrays (x32 = float32(x); each part
    the bf16 rounding of what the earlier ones leave, the remainders formed in float32)."""
    r = np.asarray(x, dtype=np.float32)
    parts = []
    for _ in range(nparts):
        h = bf16_round(r)
        parts.append(h.astype(np.float64))
        r = (r - h).astype(np.float32)
    return tuple(parts)


def split_matmul(a, b, nparts=2):
    """a @ b with both operands split into nparts parts: the part products p_a q_b with
    a + b < nparts (hi_a hi_b + hi_a lo_b + lo_a hi_b for two parts), FP64 sums."""
    pa = split(a, nparts)
    pb = split(b, nparts)
    out = 0.0
    for s in range(nparts):
        for i in range(s + 1):
            out = out + pa[i] @ pb[s - i]
    return out


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


def k_build(B_rows, mask, Cocc, nparts=2):
    """Mixed-mode K from the reference-layout DF tensor (nparts BF16 parts per operand).

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
        X[i] = split_matmul(Cocc[js].T, B_rows[mask[i, js]], nparts)
    X32 = X.astype(np.float32)
    Xf = X32.reshape(nbf, -1)
    K = split_matmul(Xf, Xf.T, nparts)
    return K, X32
