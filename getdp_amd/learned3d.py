"""The learned preconditioner on the 3D problem: the U-Net with 3x3x3
convolutions (``gdp_unet_create_3d``) on the node grid of
``problems.poisson3d`` (P1 on 6-tet Kuhn cells, checkerboard or lognormal
reluctivity), trained on residual/error pairs from the plain solver's own
GMRES iterates and used as the right preconditioner of FGMRES (Sparskit
``fgmres``, iters.f:2522-2807), with an ILU(0) correction stage and the
ILU(0)-GMRES fallback of :func:`learned.solve_learned`.

    python -m getdp_amd.learned3d train --n 32 --weights unet_poisson3d.npz
    python -m getdp_amd.learned3d eval --weights ... --eval-n 32 64
"""
from __future__ import annotations

import time

import numpy as np

from . import learned as L2
from .learned import FEAT_SCALE, GridMaps, UNet, diag_scale, from_grid, to_grid


def grid_maps(pr, dof_table: np.ndarray, A=None) -> GridMaps:
    """Voxel (z, y, x) <- node (i = x, j = y, k = z) of the (n+1)^3 node grid
    for 0 <= i, j, k < n (the nodes of the faces i, j, k = 0 and = n are
    Dirichlet nodes without a DOF; the grid is n^3, n a multiple of 8).
    Feature: log(nu) of the 8 cells around a node, averaged, centred over the
    DOFs and scaled.  With A the network works in the diagonally scaled space."""
    n = int(pr.meta["n"])
    kk, jj, ii = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    node = ((kk * (n + 1) + jj) * (n + 1) + ii + 1).ravel()
    d = dof_table[node]
    pix2dof = np.where(d > 0, d - 1, -1).astype(np.int32)
    ndof = int((dof_table > 0).sum())
    if int((pix2dof >= 0).sum()) != ndof:
        raise ValueError("some DOFs are not on the voxel grid")
    dof2pix = np.zeros(ndof, np.int64)
    dof2pix[pix2dof[pix2dof >= 0]] = np.nonzero(pix2dof >= 0)[0]
    K = int(pr.meta["ncls"])
    nt = int(pr.meta["ntet"])
    cls = (pr.region[:nt] - 1) % K
    lognu_cell = np.log(np.asarray(pr.meta["cls_val"])[cls[0::6]]).reshape(n, n, n)  # [ck, cj, ci]
    pad = np.pad(lognu_cell, 1, mode="edge")
    node_lognu = sum(pad[a:a + n + 1, b:b + n + 1, c:c + n + 1] for a in (0, 1) for b in (0, 1)
                     for c in (0, 1)) / 8.0
    f = node_lognu[:n, :n, :n]
    mask = (pix2dof >= 0).reshape(n, n, n).astype(np.float32)
    f = (f - f[mask > 0].mean()) * FEAT_SCALE
    ds = diag_scale(A) if A is not None else None
    return GridMaps(n, n, ndof, pix2dof, dof2pix, mask, f.astype(np.float32), ds)


def random_field_problem(n: int, seed: int):
    """Seeded field: half checkerboard {1, 1e-3} over 4^3..16^3 subcubes
    (a random 30-70 % of them at 1e-3), half lognormal with sigma in [0.5, 2]."""
    from . import problems as P
    rng = np.random.default_rng(20_000 + seed)
    if seed % 2 == 0:
        nsub = int(rng.choice([4, 8, 16]))
        return P.poisson3d(n, seed=seed, nsub=min(nsub, n), frac=float(rng.uniform(0.3, 0.7)))
    return P.poisson3d(n, seed=seed, field_kind="lognormal", sigma=float(rng.uniform(0.5, 2.0)))


def random_multiscale(n: int, rng) -> np.ndarray:
    """Random n^3 field with a power-law spectrum (1 + |k|^2/k0^2)^(-alpha/2)."""
    w = rng.standard_normal((n, n, n))
    k = np.fft.fftfreq(n) * n
    kr = np.fft.rfftfreq(n) * n
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + kr[None, None, :] ** 2
    alpha = rng.uniform(0.0, 3.0)
    k0 = rng.uniform(1.0, 6.0)
    f = np.fft.irfftn(np.fft.rfftn(w) * (1.0 + k2 / k0 ** 2) ** (-alpha / 2), s=(n, n, n), axes=(0, 1, 2))
    return f / max(f.std(), 1e-30)


def generate_pairs(n: int, nfields: int, seed0: int = 0, ks=(0, 3, 10, 30), rhs_per_field: int = 5,
                   log=None, **solver_kw):
    """(r_k, e_k) pairs of the plain ILU(0)-GMRES(30) stopped after k
    iterations, for known multi-scale solutions, in the network's (diagonally
    scaled) space; returns r, e, feat (P, n, n, n) float32 and the mask."""
    from . import linalg as LA
    rs, es, fs = [], [], []
    mask = None
    t0 = time.perf_counter()
    for f in range(nfields):
        pr = random_field_problem(n, seed0 + f)
        fe = LA.FESystem(pr)
        A = LA.Matrix(fe.N)
        b0 = LA.Vector(fe.N)
        fe.generate(A, b0)
        maps = grid_maps(pr, fe.dof_table(0), A)
        ds = maps.dscale
        mask = maps.mask
        rng = np.random.default_rng(seed0 * 7919 + f)
        s = L2._solver(**solver_kw)
        first = True
        y = LA.Vector(fe.N)
        for q in range(rhs_per_field):
            xs = from_grid(random_multiscale(n, rng), maps) * ds
            LA.prod_matrix_vector(A, LA.Vector(fe.N, xs), y)
            bh = y.get()
            b = LA.Vector(fe.N, bh)
            for k in ks:
                if k == 0:
                    xk = np.zeros(fe.N)
                else:
                    s.set("Nb_Iter_Max", k)
                    xv = LA.Vector(fe.N)
                    (LA.solve if first else LA.solve_again)(A, b, s, xv)
                    first = False
                    xk = xv.get()
                LA.prod_matrix_vector(A, LA.Vector(fe.N, xk), y)
                rs.append(to_grid(ds * (bh - y.get()), maps).astype(np.float32))
                es.append(to_grid((xs - xk) / ds, maps).astype(np.float32))
                fs.append(maps.feat)
        if log and (f + 1) % max(1, nfields // 10) == 0:
            log(f"pairs: field {f + 1}/{nfields}, {len(rs)} pairs, {time.perf_counter() - t0:.1f} s")
    return np.stack(rs), np.stack(es), np.stack(fs), mask


def train(net: UNet, r, e, feat, mask, steps: int, batch: int, lr: float = 1e-3, seed: int = 0, log=None,
          log_every: int = 100):
    """Adam on the device with cosine decay; each batch is shown under a random
    permutation of the three axes (the Kuhn split, whose tetrahedra share the
    cell diagonal, and the mask are invariant under it)."""
    rng = np.random.default_rng(seed)
    rn, en = L2.normalise_pairs(r, e, mask)
    hist = []
    t0 = time.perf_counter()
    for s in range(steps):
        idx = rng.integers(0, rn.shape[0], size=batch)
        perm = (0,) + tuple(1 + rng.permutation(3))
        rb, eb, fb = (np.ascontiguousarray(x[idx].transpose(perm)) for x in (rn, en, feat))
        cur = lr * 0.5 * (1 + np.cos(np.pi * s / steps))
        loss = net.train_step(rb, fb, eb, mask, cur)
        hist.append(loss)
        if log and ((s + 1) % log_every == 0 or s == 0):
            log(f"train step {s + 1}/{steps}: loss {np.mean(hist[-log_every:]):.4f} lr {cur:.2e} "
                f"({time.perf_counter() - t0:.1f} s)")
    return np.asarray(hist)


def evaluate(params: np.ndarray, n: int, seeds, log=None, precisions=(0, 1), lsq: bool = True):
    """Learned two-stage FGMRES (FP64, and mixed precision: FP32 inner FGMRES
    with FP64 refinement) against plain ILU(0)-GMRES(30) in the reference's
    RCMK ordering and in multicolour ordering, on unseen fields, all to 1e-10
    relative residual (true FP64 residual).  Wall-clock of each solve call
    after one untimed warm-up."""
    from . import linalg as LA
    rows = []
    for seed in seeds:
        pr = random_field_problem(n, seed)
        fe, A, b = LA.assemble(pr)
        maps = grid_maps(pr, fe.dof_table(0), A)
        net = UNet(n, n, batch=1, seed=0, D=n)
        net.set_params(params)
        net.set_mode(lsq=lsq)
        row = {"n": n, "seed": seed, "field": "checker" if seed % 2 == 0 else "lognormal", "dofs": fe.N}
        for ordering, key in ((0, "plain_rcmk"), (1, "plain_multicolour")):
            L2.plain_solve(A, b, fe.N, ordering)
            x, info = L2.plain_solve(A, b, fe.N, ordering)
            row[key] = {"its": info["its"], "ms": round(info["t_wall_ms"], 2), "rel_res": info["rel_res"]}
        for prec in precisions:
            for ordering in (0, 1):
                key = f"learned_{'rcmk' if ordering == 0 else 'multicolour'}" + ("_mixed" if prec else "")
                x = LA.Vector(fe.N)
                L2.solve_learned(A, b, x, net, maps, maxits=2000, two_stage=True, ilu_ordering=ordering,
                                 precision=prec)
                x = LA.Vector(fe.N)
                out = L2.solve_learned(A, b, x, net, maps, maxits=2000, two_stage=True, ilu_ordering=ordering,
                                       precision=prec)
                row[key] = {"its": out["learned"]["its"], "ms": round(1e3 * out["t_total_s"], 2),
                            "fallback_its": None if out["fallback"] is None else out["fallback"]["its"],
                            "rel_res": out["rel_res"]}
        learned = {k: v for k, v in row.items() if k.startswith("learned") and v["rel_res"] <= 1.5e-10}
        best_l = min(learned, key=lambda k: learned[k]["ms"]) if learned else None
        row["best_learned"] = best_l
        row["speedup_vs_plain_rcmk"] = (round(row["plain_rcmk"]["ms"] / learned[best_l]["ms"], 2)
                                        if best_l else None)
        conv = [row[k]["ms"] for k in ("plain_rcmk", "plain_multicolour") if row[k]["rel_res"] <= 1.5e-10]
        row["speedup_vs_best_plain"] = (round(min(conv) / learned[best_l]["ms"], 2)
                                        if best_l and conv else None)
        rows.append(row)
        if log:
            log(f"eval3d n={n} seed={seed} ({row['field']}): " + ", ".join(
                f"{k} {v['its']} its {v['ms']} ms" for k, v in row.items() if isinstance(v, dict))
                + f"; best {best_l}: {row['speedup_vs_plain_rcmk']}x / {row['speedup_vs_best_plain']}x")
    return rows


def main(argv=None):
    import argparse
    import json
    import sys
    from . import linalg as LA
    ap = argparse.ArgumentParser(description="train / evaluate the 3D learned preconditioner")
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--fields", type=int, default=120)
    ap.add_argument("--rhs", type=int, default=5)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--weights", default=None)
    ap.add_argument("--eval-n", type=int, nargs="*", default=[32, 64])
    ap.add_argument("--eval-seeds", type=int, nargs="*", default=[90000, 90001])
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)

    def log(*m):
        print(*m, file=sys.stderr, flush=True)

    LA.init(0)
    if a.cmd == "train":
        r, e, f, mask = generate_pairs(a.n, a.fields, seed0=a.seed, rhs_per_field=a.rhs, log=log,
                                       ILU_Ordering=1)
        log(f"{r.shape[0]} pairs")
        net = UNet(a.n, a.n, batch=a.batch, seed=a.seed, D=a.n)
        hist = train(net, r, e, f, mask, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed, log=log,
                     log_every=250)
        del r, e, f
        out = a.weights or "unet_poisson3d.npz"
        net.set_mode(lsq=True)  # trained on the error direction only: the scale is fitted at use
        net.save(out)
        log(f"saved {out}; final loss {hist[-200:].mean():.4f}")
        params = net.get_params()
    else:
        with np.load(a.weights, allow_pickle=False) as z:
            params = z["params"]
    rows = []
    for n in a.eval_n:
        rows += evaluate(params, n, a.eval_seeds, log=log, lsq=True)
    if a.json:
        with open(a.json, "w") as fh:
            json.dump(rows, fh, indent=1)


if __name__ == "__main__":
    main()
