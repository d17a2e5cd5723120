"""Exact-diagonalisation golden energies (independent of the library: scipy
sparse Hamiltonians in the S^z basis).  Run: python tests/golden/make_ed.py"""
import itertools
import json
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def heisenberg_sz0(N, bonds, couplings):
    states = [s for s in itertools.product((0, 1), repeat=N) if sum(s) == N // 2]
    index = {s: k for k, s in enumerate(states)}
    rows, cols, vals = [], [], []
    for k, s in enumerate(states):
        diag = 0.0
        for (i, j), J in zip(bonds, couplings):
            si, sj = 0.5 - s[i], 0.5 - s[j]
            diag += J * si * sj
            if s[i] != s[j]:
                t = list(s)
                t[i], t[j] = t[j], t[i]
                rows.append(index[tuple(t)])
                cols.append(k)
                vals.append(0.5 * J)
        rows.append(k)
        cols.append(k)
        vals.append(diag)
    n = len(states)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def ground(H):
    return float(sla.eigsh(H, k=1, which="SA", tol=1e-14, maxiter=100000)[0][0])


def main():
    out = {}
    N = 16
    bonds = [(i, i + 1) for i in range(N - 1)]
    out["heisenberg_obc_L16_sz0"] = ground(heisenberg_sz0(N, bonds, [1.0] * len(bonds)))
    N = 12
    bonds = [(i, i + 1) for i in range(N - 1)]
    out["heisenberg_obc_L12_sz0"] = ground(heisenberg_sz0(N, bonds, [1.0] * len(bonds)))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ed_energies.json")
    json.dump(out, open(path, "w"), indent=1)
    print(out)


if __name__ == "__main__":
    main()
