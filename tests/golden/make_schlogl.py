"""Writes the Schlogl fixtures used by the tests.

1. tests/golden/schlogl/ : the model of schlogl.xml (reference repo, lines
   39-249) in the input-folder format read by the CLI: species order A, X, B;
   A (1e5) and B (2e5) are boundary species (M_feed); X0 = 250;
   k = 3e-7, 1e-4, 1e-3, 3.5; 1024 identical rows of MX_0 / c_matrix;
   t_vector = 0.1 * i, i = 1..100 (t_end = 10).
2. tests/golden/schlogl_cme.npz : the exact chemical master equation for X
   (A, B constant) with the propensities the simulator uses
   (a1 = c1 A X(X-1)/2, a2 = c2 X(X-1)(X-2)/6, a3 = c3 B, a4 = c4 X),
   truncated at X = 2000 and solved with scipy's expm_multiply; the CDF of X at
   every tenth sample time plus mean/sd at all 100 times.

Run:  python tests/golden/make_schlogl.py
"""
import os

import numpy as np
from scipy.sparse import csr_matrix
from scipy.sparse.linalg import expm_multiply

HERE = os.path.dirname(os.path.abspath(__file__))
LEFT = [[1, 2, 0], [0, 3, 0], [0, 0, 1], [0, 1, 0]]
RIGHT = [[0, 3, 0], [1, 2, 0], [0, 1, 0], [0, 0, 1]]
FEED = [1, 0, 1]
X0 = [100000, 250, 200000]
C = [3e-7, 1e-4, 1e-3, 3.5]
N_ROWS = 1024
TS = [round(0.1 * i, 10) for i in range(1, 101)]


def write_folder(path):
    os.makedirs(path, exist_ok=True)
    tab = lambda row: "\t".join(str(v) for v in row)
    with open(os.path.join(path, "left_side"), "w") as f:
        f.write("\n".join(tab(r) for r in LEFT) + "\n")
    with open(os.path.join(path, "right_side"), "w") as f:
        f.write("\n".join(tab(r) for r in RIGHT) + "\n")
    with open(os.path.join(path, "M_feed"), "w") as f:
        f.write(tab(FEED) + "\n")
    with open(os.path.join(path, "MX_0"), "w") as f:
        f.write("".join(tab(X0) + "\n" for _ in range(N_ROWS)))
    with open(os.path.join(path, "c_matrix"), "w") as f:
        f.write("".join(tab(C) + "\n" for _ in range(N_ROWS)))
    with open(os.path.join(path, "t_vector"), "w") as f:
        f.write("".join(f"{t:g}\n" for t in TS))


def cme(nmax=2000):
    A, B = X0[0], X0[2]
    x = np.arange(nmax + 1, dtype=np.float64)
    up = C[0] * A * x * (x - 1) / 2 + C[2] * B
    dn = C[1] * x * (x - 1) * (x - 2) / 6 + C[3] * x
    up[-1] = 0.0                       # truncation: no birth out of the box
    rows, cols, vals = [], [], []
    for i in range(nmax + 1):
        if i < nmax:
            rows.append(i + 1); cols.append(i); vals.append(up[i])
        if i > 0:
            rows.append(i - 1); cols.append(i); vals.append(dn[i])
        rows.append(i); cols.append(i); vals.append(-(up[i] + (dn[i] if i > 0 else 0.0)))
    Q = csr_matrix((vals, (rows, cols)), shape=(nmax + 1, nmax + 1))   # dp/dt = Q p
    p0 = np.zeros(nmax + 1); p0[X0[1]] = 1.0
    P = expm_multiply(Q, p0, start=0.0, stop=TS[-1], num=len(TS) + 1, endpoint=True)[1:]
    P = np.clip(P, 0, None)
    mean = P @ x
    sd = np.sqrt(np.maximum(P @ (x * x) - mean ** 2, 0))
    return P, mean, sd


if __name__ == "__main__":
    write_folder(os.path.join(HERE, "schlogl"))
    P, mean, sd = cme()
    sel = np.arange(9, 100, 10)
    np.savez_compressed(os.path.join(HERE, "schlogl_cme.npz"), t=np.array(TS), mean=mean, sd=sd,
                        t_cdf=np.array(TS)[sel], cdf=np.cumsum(P[sel], axis=1).astype(np.float64),
                        mass=P.sum(axis=1))
    print("mass min", P.sum(axis=1).min(), "mean@10", mean[-1], "sd@10", sd[-1])
