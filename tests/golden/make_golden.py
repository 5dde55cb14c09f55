"""Writes the golden fixtures under tests/golden/ (run from the repo root).

Expected outputs come from dense float64 products and boolean pattern
products in numpy and from scipy.io.mmread, never from the oracle, so the
fixtures pin the oracle independently:
  prod_<name>.npz  A, B as COO triplets (stored entries, duplicates kept),
                   C structure (pattern product), C values (dense float64)
                   and sum |a*b| per entry
  mtx_<name>.mtx   Matrix Market inputs; mtx_<name>.npz the COO scipy reads
"""
import os

import numpy as np
import scipy.io
import scipy.sparse as sp

OUT = os.path.dirname(os.path.abspath(__file__))


def truth(Ar, Ac, Av, Ash, Br, Bc, Bv, Bsh):
    m, K = Ash
    K2, n = Bsh
    assert K == K2
    PA = np.zeros((m, K), np.int64)
    PB = np.zeros((K, n), np.int64)
    np.add.at(PA, (Ar, Ac), 1)
    np.add.at(PB, (Br, Bc), 1)
    DA = np.zeros((m, K), np.float64)
    DB = np.zeros((K, n), np.float64)
    np.add.at(DA, (Ar, Ac), Av.astype(np.float64))
    np.add.at(DB, (Br, Bc), Bv.astype(np.float64))
    AA = np.zeros((m, K), np.float64)
    AB = np.zeros((K, n), np.float64)
    np.add.at(AA, (Ar, Ac), np.abs(Av.astype(np.float64)))
    np.add.at(AB, (Br, Bc), np.abs(Bv.astype(np.float64)))
    S = PA @ PB
    V = DA @ DB
    SA = AA @ AB
    ptr = np.zeros(m + 1, np.int64)
    idx, val, sab = [], [], []
    for i in range(m):
        js = np.nonzero(S[i])[0]
        ptr[i + 1] = ptr[i] + len(js)
        idx.append(js)
        val.append(V[i, js])
        sab.append(SA[i, js])
    cat = lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(0, dt)
    return ptr, cat(idx, np.int32), cat(val, np.float64), cat(sab, np.float64), int(S.sum())


def grid_vals(rng, size):
    # values on a 2^-10 grid in [-1, 1): FP32 products and short sums exact
    return (rng.integers(-1024, 1024, size) / 1024.0).astype(np.float32)


def rand_coo(rng, m, n, dens, grid=True):
    M = sp.random(m, n, density=dens, random_state=rng, format="coo")
    r, c = M.row.astype(np.int32), M.col.astype(np.int32)
    v = grid_vals(rng, len(r)) if grid else rng.uniform(-1, 1, len(r)).astype(np.float32)
    return r, c, v


def save_prod(name, A, Ash, B, Bsh):
    ptr, idx, val, sab, nprod = truth(*A, Ash, *B, Bsh)
    np.savez_compressed(os.path.join(OUT, f"prod_{name}.npz"),
                        A_rows=A[0], A_cols=A[1], A_vals=A[2], A_shape=np.array(Ash),
                        B_rows=B[0], B_cols=B[1], B_vals=B[2], B_shape=np.array(Bsh),
                        C_ptr=ptr, C_idx=idx, C_val=val, C_sabs=sab, nprod=np.int64(nprod))


def main():
    rng = np.random.default_rng(20240601)
    # tiny: hand-made, duplicate entry in A, an exact cancellation (C(0,3) = 0
    # is structural), an empty row of A, an empty column of B
    A = (np.array([0, 0, 2, 3, 3, 3, 5, 5], np.int32), np.array([1, 2, 0, 4, 4, 2, 1, 4], np.int32),
         np.array([1.0, 1.0, 0.5, 0.25, 0.25, -1.5, 2.0, 1.0], np.float32))
    B = (np.array([1, 2, 0, 0, 4, 4, 2], np.int32), np.array([3, 3, 0, 6, 1, 6, 5], np.int32),
         np.array([2.0, -2.0, 4.0, 0.5, -1.0, 3.0, 0.125], np.float32))
    save_prod("tiny", A, (6, 5), B, (5, 7))
    save_prod("rect", rand_coo(rng, 40, 30, 0.1), (40, 30), rand_coo(rng, 30, 50, 0.1), (30, 50))
    # power-law rows/cols drawn with numpy (not the product generator)
    n = 128
    e = 900
    r = np.minimum((rng.pareto(1.2, e) * 4).astype(np.int64), n - 1).astype(np.int32)
    c = np.minimum((rng.pareto(1.2, e) * 4).astype(np.int64), n - 1).astype(np.int32)
    key = np.unique(r.astype(np.int64) * n + c)
    r, c = (key // n).astype(np.int32), (key % n).astype(np.int32)
    Pv = grid_vals(rng, len(r))
    save_prod("powerlaw", (r, c, Pv), (n, n), (r, c, Pv), (n, n))
    # one dense row and one dense column
    n = 64
    rr = list(range(n)) + [7] * n
    cc = [9] * n + list(range(n))
    key = np.unique(np.array(rr, np.int64) * n + np.array(cc))
    r, c = (key // n).astype(np.int32), (key % n).astype(np.int32)
    Dv = grid_vals(rng, len(r))
    save_prod("dense_rowcol", (r, c, Dv), (n, n), (r, c, Dv), (n, n))
    save_prod("n1", (np.array([0], np.int32), np.array([0], np.int32), np.array([0.75], np.float32)),
              (1, 1), (np.array([0], np.int32), np.array([0], np.int32), np.array([-2.0], np.float32)),
              (1, 1))
    z = np.zeros(0, np.int32)
    save_prod("empty", (z, z, np.zeros(0, np.float32)), (5, 4), (z, z, np.zeros(0, np.float32)), (4, 6))
    # uniform-in-[-1,1] values with cancellation (not on the grid)
    save_prod("rand_real", rand_coo(rng, 60, 60, 0.08, grid=False), (60, 60),
              rand_coo(rng, 60, 60, 0.08, grid=False), (60, 60))

    # Matrix Market files
    mtx = {
        "general": "%%MatrixMarket matrix coordinate real general\n% comment line\n%\n"
                   "4 3 5\n1 1 1.5\n2 3 -2.25\n4 2 3e-1\n3 1 7\n2 3 0.5\n",
        "symmetric": "%%MatrixMarket matrix coordinate real symmetric\n"
                     "4 4 5\n1 1 2.0\n2 1 -1.0\n3 2 0.5\n4 4 1.0\n4 1 3.0\n",
        "skew": "%%MatrixMarket matrix coordinate real skew-symmetric\n"
                "3 3 2\n2 1 4.0\n3 2 -1.5\n",
        "pattern": "%%MatrixMarket matrix coordinate pattern general\n"
                   "3 5 4\n1 2\n3 5\n2 1\n3 3\n",
        "integer": "%%MatrixMarket matrix coordinate integer general\n"
                   "2 2 3\n1 1 3\n2 1 -4\n2 2 12\n",
        "long_line": "%%MatrixMarket matrix coordinate real general\n%" + "x" * 300 + "\n"
                     "2 2 2\n1 2 " + "1." + "0" * 150 + "\n2 1 2\n",
    }
    for name, text in mtx.items():
        path = os.path.join(OUT, f"mtx_{name}.mtx")
        with open(path, "w") as f:
            f.write(text)
        M = scipy.io.mmread(path)
        M = sp.coo_matrix(M)
        np.savez_compressed(os.path.join(OUT, f"mtx_{name}.npz"), rows=M.row.astype(np.int32),
                            cols=M.col.astype(np.int32), vals=M.data.astype(np.float32),
                            shape=np.array(M.shape))


if __name__ == "__main__":
    main()
