"""Writes sir_fsp.npz: exact marginal laws of the SIR benchmark model.

Model: S + I -> 2 I at 1e-4 * S * I, I -> R at 0.01 * I, from (S, I, R) =
(999, 1, 0); the population S + I + R = 1000 is conserved, so the states
(S, I) with S + I <= 1000 are a closed, finite state space (501 501 states)
and the chemical master equation p' = A p is solved exactly (up to the
Krylov/Taylor tolerance of scipy.sparse.linalg.expm_multiply) with no
projection error.  Marginals of S, I and R are stored at t = 0, 10, ..., 250.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import expm_multiply

HERE = Path(__file__).resolve().parent
N = 1000
K_INF, K_REC = 1e-4, 0.01


def index(s, i):
    # states ordered by s, then i in 0..N-s
    return s * (N + 1) - s * (s - 1) // 2 + i


def main() -> None:
    S_list, I_list = [], []
    for s in range(N + 1):
        for i in range(N - s + 1):
            S_list.append(s)
            I_list.append(i)
    S = np.array(S_list)
    I = np.array(I_list)
    n = len(S)
    assert index(N, 0) == n - 1
    a_inf = K_INF * S * I
    a_rec = K_REC * I
    rows, cols, vals = [], [], []
    src = np.arange(n)
    m = a_inf > 0
    rows.append(index(S[m] - 1, I[m] + 1)); cols.append(src[m]); vals.append(a_inf[m])
    m = a_rec > 0
    rows.append(index(S[m], I[m] - 1)); cols.append(src[m]); vals.append(a_rec[m])
    rows.append(src); cols.append(src); vals.append(-(a_inf + a_rec))
    A = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    p0 = np.zeros(n)
    p0[index(999, 1)] = 1.0
    times = np.arange(0, 251, 10, dtype=np.float64)
    P = expm_multiply(A, p0, start=0.0, stop=250.0, num=len(times), endpoint=True)
    R = N - S - I
    mS = np.zeros((len(times), N + 1)); mI = np.zeros_like(mS); mR = np.zeros_like(mS)
    for k in range(len(times)):
        p = np.clip(P[k], 0, None)
        mS[k] = np.bincount(S, weights=p, minlength=N + 1)
        mI[k] = np.bincount(I, weights=p, minlength=N + 1)
        mR[k] = np.bincount(R, weights=p, minlength=N + 1)
    np.savez_compressed(HERE / "sir_fsp.npz", times=times, S=mS.astype(np.float32),
                        I=mI.astype(np.float32), R=mR.astype(np.float32),
                        mass=P.sum(axis=1))
    print("mass", P.sum(axis=1).min(), P.sum(axis=1).max())


if __name__ == "__main__":
    main()
