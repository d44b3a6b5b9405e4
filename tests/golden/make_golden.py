"""Writes tests/golden/golden.json: the reference tests' own input recipes and hand
tensors with expected outputs computed by brute force in exact (integer / Fraction)
arithmetic from the index definitions of the reference kernels.  The oracle and the
HIP library are both checked against these files.

Inputs:
  * test/test_mttkrp.cpp:15-27 -- srand(2); X, A, B, C as float(int(10*(rand()/RAND_MAX - 0.5)))
    (regenerated with the C library's own rand(), not a reimplementation)
  * test/test_t2m.cpp:10-13   -- srand(2); A[i] = rand()*0.1/RAND_MAX, 4 x 3 x 2
  * test/test_ttm.cpp:21-34   -- srand(2); X, U1, U2, U3 = rand()*0.1/RAND_MAX - 0.5
  * decomposition/GPU/large/de_cp.cpp:16-19      -- 3 x 4 x 2 integer X, A, B, C
  * decomposition/GPU/normal/cp/de_cp.cpp:16-19  -- 3 x 2 x 2 X = 0..11, B, C
  * decomposition/GPU/normal/tucker/de_tucker.cpp:18-21 -- A, B, C, G of a 3x4x2 Tucker model
Run:  python tests/golden/make_golden.py
"""
import ctypes
import json
import os
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
libc = ctypes.CDLL("libc.so.6")
libc.rand.restype = ctypes.c_int
RAND_MAX = 2147483647


def srand(s):
    libc.srand(ctypes.c_uint(s))


def rand_int_float():
    # float(int(10*(((float) rand())/RAND_MAX - 0.5))): float division, then double arithmetic
    f = np.float32(libc.rand()) / np.float32(RAND_MAX)
    return float(int(10.0 * (float(f) - 0.5)))


def rand_scaled(offset):
    # rand()*0.1/(RAND_MAX) [- 0.5] in double, stored to float
    v = libc.rand() * 0.1 / RAND_MAX - offset
    return float(np.float32(v))


def idx3(i, j, k, I, J):
    return i + j * I + k * I * J


def mttkrp_exact(X, dims, U, mode, R):
    I, J, K = dims
    out = [[Fraction(0)] * R for _ in range(dims[mode])]
    for k in range(K):
        for j in range(J):
            for i in range(I):
                x = Fraction(X[idx3(i, j, k, I, J)])
                ii = (i, j, k)
                for r in range(R):
                    p = x
                    for m in range(3):
                        if m != mode:
                            p *= Fraction(U[m][ii[m] + r * dims[m]])
                    out[ii[mode]][r] += p
    return [float(out[i][r]) for r in range(R) for i in range(dims[mode])]  # column-major


def t2m_exact(T, flag, m, n, k):
    out = [None] * (m * n * k)
    for tube in range(k):
        for col in range(n):
            for row in range(m):
                src = tube * m * n + col * m + row
                if flag == 1:
                    dst = src
                elif flag == 2:
                    dst = tube * m * n + row * n + col
                else:
                    dst = k * (col * m + row) + tube
                out[dst] = T[src]
    return out


def ttm_exact(X, U, flag, a, b, c, d):
    size = {1: d * b * c, 2: a * d * c, 3: a * b * d}[flag]
    out = [Fraction(0)] * size
    for k in range(c):
        for j in range(b):
            for i in range(a):
                x = Fraction(X[i + j * a + k * a * b])
                for p in range(d):
                    if flag == 1:
                        out[p + j * d + k * d * b] += Fraction(U[p + i * d]) * x
                    elif flag == 2:
                        out[i + p * a + k * a * d] += Fraction(U[p + j * d]) * x
                    else:
                        out[i + j * a + p * a * b] += Fraction(U[p + k * d]) * x
    return [float(v) for v in out]


def gencp_exact(U, dims, R):
    I, J, K = dims
    return [float(sum(Fraction(U[0][i + r * I]) * Fraction(U[1][j + r * J]) * Fraction(U[2][k + r * K])
                      for r in range(R)))
            for k in range(K) for j in range(J) for i in range(I)]


def gentucker_exact(G, U, dims, ranks):
    I, J, K = dims
    r1, r2, r3 = ranks
    out = []
    for k in range(K):
        for j in range(J):
            for i in range(I):
                s = Fraction(0)
                for c3 in range(r3):
                    for c2 in range(r2):
                        for c1 in range(r1):
                            s += (Fraction(G[c1 + c2 * r1 + c3 * r1 * r2]) * Fraction(U[0][i + c1 * I]) *
                                  Fraction(U[1][j + c2 * J]) * Fraction(U[2][k + c3 * K]))
                out.append(float(s))
    return out


def main():
    g = {}
    # test_mttkrp.cpp
    m = n = k = 3
    r = 2
    srand(2)
    X = [rand_int_float() for _ in range(m * n * k)]
    A = [rand_int_float() for _ in range(m * r)]
    B = [rand_int_float() for _ in range(n * r)]
    C = [rand_int_float() for _ in range(k * r)]
    g["mttkrp_srand2"] = {
        "dims": [m, n, k], "rank": r, "X": X, "A": A, "B": B, "C": C,
        "M": [mttkrp_exact(X, [m, n, k], [A, B, C], mode, r) for mode in range(3)],
    }
    # test_t2m.cpp
    m, n, k = 4, 3, 2
    srand(2)
    T = [rand_scaled(0.0) for _ in range(m * n * k)]
    g["t2m_srand2"] = {"dims": [m, n, k], "T": T, "unfold": [t2m_exact(T, f, m, n, k) for f in (1, 2, 3)]}
    # test_ttm.cpp
    a = b = c = 3
    d = 2
    srand(2)
    X = [rand_scaled(0.5) for _ in range(a * b * c)]
    U1 = [rand_scaled(0.5) for _ in range(d * a)]
    U2 = [rand_scaled(0.5) for _ in range(d * b)]
    U3 = [rand_scaled(0.5) for _ in range(d * c)]
    g["ttm_srand2"] = {"dims": [a, b, c], "d": d, "X": X, "U": [U1, U2, U3],
                       "XU": [ttm_exact(X, U, f, a, b, c, d) for f, U in zip((1, 2, 3), (U1, U2, U3))]}
    # decomposition/GPU/large/de_cp.cpp hand tensor
    X = [2, 0, 4, 6, -1, 1, 1, 2, 6, 3, 8, 2, 12, 4, 8, 22, 21, 1, 3, 4, 2, 6, 4, 2]
    A = [3, 2, 12, 5, 0, 3]
    B = [7, 2, 8, 9, 13, 2, 12, 11]
    C = [4, 5, 8, 2]
    dims = [3, 4, 2]
    g["hand_cp_3x4x2"] = {"dims": dims, "rank": 2, "X": X, "A": A, "B": B, "C": C,
                          "M": [mttkrp_exact(X, dims, [A, B, C], mode, 2) for mode in range(3)],
                          "gencp": gencp_exact([A, B, C], dims, 2)}
    # decomposition/GPU/normal/cp/de_cp.cpp hand tensor (A is all zeros there; it is output)
    X = list(range(12))
    B = [2, 1, 6, 3]
    C = [0, 7, 1, 2]
    A = [0, 0, 0, 0, 0, 0]
    dims = [3, 2, 2]
    g["hand_cp_3x2x2"] = {"dims": dims, "rank": 2, "X": X, "B": B, "C": C,
                          "M0": mttkrp_exact(X, dims, [A, B, C], 0, 2)}
    # decomposition/GPU/normal/tucker/de_tucker.cpp model
    A = [1, 3, 2, 0, -1, 2]
    B = [2, 0, 4, 6, -1, 1, 1, 2]
    C = [3, -1, 0, 2]
    G = [-2, -1, 0, 3, 4, 1, 2, 3]
    g["hand_tucker_3x4x2"] = {"dims": [3, 4, 2], "ranks": [2, 2, 2], "A": A, "B": B, "C": C, "G": G,
                              "X": gentucker_exact(G, [A, B, C], [3, 4, 2], [2, 2, 2])}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(g, f, indent=1)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
