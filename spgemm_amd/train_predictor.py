"""Trains the row-count predictor on seeded synthetic products and saves it.

python -m spgemm_amd.train_predictor [--out spgemm_amd/models/rmat_mix32.mlp]

The training set is the exact row counts (from the symbolic pass) of 32
seeded products: 16 R-MAT (scales 12-15, edge factors 4-16, (a, b, c) =
(0.57, 0.19, 0.19), every other one rectangular with 4x the rows of B),
8 with uniform random rows (n 4096-32768, 4-32 nnz per row) and 8 27-point
stencils (grids 10^3 to 24^3).  Training is deterministic for a given seed,
so the saved file can be regenerated bit for bit.
"""
import argparse
import os

import spgemm_amd as spg

DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models",
                             "mix32.mlp")


def training_set(n_mats=32):
    for j in range(n_mats):
        kind = j % 4
        if kind in (0, 1):  # R-MAT, square or rectangular
            s = 12 + (j // 2) % 4
            ef = (4, 8, 12, 16)[(j // 8) % 4]
            if kind:
                A = spg.gen_rmat(s + 2, s, ef << (s + 2), seed=1000 + j)
                B = spg.gen_rmat(s, s, ef << s, seed=2000 + j)
            else:
                A = spg.gen_rmat(s, s, ef << s, seed=1000 + j)
                B = A
        elif kind == 2:  # uniform rows
            n = 4096 << ((j // 4) % 4)
            A = spg.gen_uniform_rows(n, n, 4 + 4 * ((j // 4) % 8), 3000 + j)
            B = A
        else:  # 27-point stencil
            L = 10 + 2 * (j // 4)
            A = spg.gen_stencil27(L, L, L, 4000 + j)
            B = A
        yield A, B


def train(ctx, seed=0, n_mats=32, epochs=20, batch=4096, lr=2e-3):
    """Returns (predictor, info)."""
    P = spg.Predictor(0, seed=seed)
    rows = 0
    for A, B in training_set(n_mats):
        ctx.set_operands_from_coo(A, B)
        ctx.multiply()
        P.add_samples(ctx)
        rows += A.num_rows
    loss = P.train(epochs=epochs, batch=batch, lr=lr)
    return P, {"matrices": n_mats, "rows": rows, "epochs": epochs, "final_loss": round(loss, 6)}


def load_or_train(ctx, path=DEFAULT_MODEL):
    """The saved predictor if path exists, else a freshly trained one."""
    if path and os.path.exists(path):
        P = spg.Predictor(0, seed=0)
        P.load(path)
        return P, {"loaded": os.path.relpath(path, os.path.dirname(os.path.dirname(DEFAULT_MODEL)))}
    return train(ctx)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    ctx = spg.Context(0)
    P, info = train(ctx, seed=args.seed, epochs=args.epochs)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    P.save(args.out)
    print(info, "->", args.out, flush=True)
    P.close()
    ctx.close()


if __name__ == "__main__":
    main()
