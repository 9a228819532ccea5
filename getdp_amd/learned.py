"""Learned preconditioner for 2D nodal problems: host driver of the device
U-Net (``gdp_unet_*`` in include/gdp_hip.h).

Nothing here has a counterpart in the reference; the network is used as the
right preconditioner of flexible GMRES (Sparskit ``fgmres``, iters.f:2522-2807)
and plain ILU(0)-GMRES (iters.f:1070-1300 with ilu0, ilut.F:1833-1951) is the
fallback that also judges the final accuracy (true FP64 residual).

Pieces:
  * :class:`UNet` -- handle of the device network (create / parameters /
    forward / train step / attach to a solver).
  * :func:`grid_maps` -- pixel <-> DOF map and the log-coefficient feature of
    a structured-grid 2D problem (``problems.poisson2d``).
  * :func:`generate_pairs` -- residual/error pairs from the plain solver's own
    GMRES iterates.
  * :func:`train` -- Adam on the scale-invariant loss, on the device.
  * :func:`solve_learned` -- FGMRES with the network, ILU(0)-GMRES fallback.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from typing import Optional

import numpy as np

from ._lib import check, dptr, fptr, iptr, lib

FEAT_SCALE = 0.25  # log(nu) feature scaling (keeps the input O(1))


class UNet:
    """Device U-Net (16-32-64-128 channels) on an H x W grid, or (D > 1) on a
    D x H x W grid with 3x3x3 convolutions."""

    def __init__(self, H: int, W: int, batch: int = 1, seed: int = 0, D: int = 1):
        self.H, self.W, self.batch, self.D = int(H), int(W), int(batch), int(D)
        self.shape = (self.D, self.H, self.W) if self.D > 1 else (self.H, self.W)
        self.h = C.c_void_p()
        if self.D > 1:
            check(lib().gdp_unet_create_3d(self.D, self.H, self.W, self.batch, seed, C.byref(self.h)),
                  "gdp_unet_create_3d")
        else:
            check(lib().gdp_unet_create(self.H, self.W, self.batch, seed, C.byref(self.h)),
                  "gdp_unet_create")

    def close(self):
        if self.h and self.h.value:
            lib().gdp_unet_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self) -> int:
        n = C.c_long()
        check(lib().gdp_unet_num_params(self.h, C.byref(n)), "num_params")
        return n.value

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.num_params, np.float32)
        check(lib().gdp_unet_get_params(self.h, fptr(p)), "get_params")
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params, (p.size, self.num_params)
        check(lib().gdp_unet_set_params(self.h, fptr(p)), "set_params")

    def get_grads(self) -> np.ndarray:
        g = np.zeros(self.num_params, np.float32)
        check(lib().gdp_unet_get_grads(self.h, fptr(g)), "get_grads")
        return g

    def forward(self, r: np.ndarray, feat: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(r, np.float32).reshape(-1, *self.shape)
        feat = np.ascontiguousarray(np.broadcast_to(feat, r.shape), np.float32)
        z = np.zeros_like(r)
        check(lib().gdp_unet_forward(self.h, r.shape[0], fptr(r), fptr(feat), fptr(z)), "forward")
        return z

    def train_step(self, r, feat, target, mask, lr: float) -> float:
        r = np.ascontiguousarray(r, np.float32).reshape(-1, *self.shape)
        feat = np.ascontiguousarray(np.broadcast_to(feat, r.shape), np.float32)
        target = np.ascontiguousarray(target, np.float32).reshape(r.shape)
        mask = np.ascontiguousarray(mask, np.float32).reshape(self.shape)
        loss = C.c_float()
        check(lib().gdp_unet_train_step(self.h, r.shape[0], fptr(r), fptr(feat), fptr(target),
                                        fptr(mask), C.c_float(lr), C.byref(loss)), "train_step")
        return loss.value

    def activation(self, t: int, batch: int = 1) -> np.ndarray:
        n = C.c_long()
        check(lib().gdp_unet_get_activation(self.h, t, batch, None, C.byref(n)), "activation")
        a = np.zeros(n.value, np.float32)
        check(lib().gdp_unet_get_activation(self.h, t, batch, fptr(a), C.byref(n)), "activation")
        return a

    def attach(self, solver, maps: "GridMaps"):
        p2d = np.ascontiguousarray(maps.pix2dof, np.int32)
        feat = np.ascontiguousarray(maps.feat, np.float32)
        ds = None
        if maps.dscale is not None:
            ds = np.ascontiguousarray(maps.dscale, np.float64)
        check(lib().gdp_unet_attach(self.h, solver.h, int(maps.ndof), iptr(p2d), fptr(feat),
                                    None if ds is None else dptr(ds)), "gdp_unet_attach")

    def set_mode(self, odd: bool = False, lsq: bool = False):
        """odd: apply the odd part of the network; lsq: least-squares scaling
        of its output in the two-stage preconditioner."""
        self.lsq = bool(lsq)
        check(lib().gdp_unet_set_mode(self.h, (1 if odd else 0) | (2 if lsq else 0)), "gdp_unet_set_mode")

    @staticmethod
    def detach(solver):
        check(lib().gdp_unet_detach(solver.h), "gdp_unet_detach")

    def save(self, path):
        np.savez_compressed(path, params=self.get_params(), H=self.H, W=self.W, D=self.D,
                            lsq=int(getattr(self, "lsq", False)))

    def load(self, path):
        """Parameters, and the output-scaling mode the weights were evaluated with."""
        with np.load(path, allow_pickle=False) as z:
            self.set_params(z["params"])
            self.set_mode(lsq=bool(z["lsq"]) if "lsq" in z.files else False)


# ---------------------------------------------------------------------------
# grids


@dataclass
class GridMaps:
    H: int
    W: int
    ndof: int
    pix2dof: np.ndarray   # (H*W,) or (D*H*W,) int32, -1 where no DOF
    dof2pix: np.ndarray   # (ndof,) int64
    mask: np.ndarray      # (H, W) or (D, H, W) float32, 1 on DOF pixels
    feat: np.ndarray      # like mask: scaled log(nu) per node
    dscale: Optional[np.ndarray] = None  # (ndof,) 1/sqrt(diag A): the network's space

    @property
    def shape(self):
        return self.mask.shape


def diag_scale(A) -> np.ndarray:
    """1/sqrt(A_ii) per row of a device matrix (host copy of the CSR)."""
    ia, ja, a = A.get_csr()
    rows = np.repeat(np.arange(A.n), np.diff(ia))
    d = np.zeros(A.n)
    on = ja == rows
    d[rows[on]] = a[on]
    if not (d > 0).all():
        raise ValueError("diagonal scaling needs a positive diagonal")
    return 1.0 / np.sqrt(d)


def grid_maps(pr, dof_table: np.ndarray, A=None) -> GridMaps:
    """Pixel (y, x) <- node (i=x, j=y) of the (n+1)^2 node grid of
    ``problems.poisson2d`` for 0 <= i, j < n (the last row/column are
    Dirichlet nodes and carry no DOF).  ``dof_table`` is
    :meth:`FESystem.dof_table` of the nodal space (1-based DOF numbers > 0).
    With the matrix ``A`` the network works in the diagonally scaled space
    (input D^{-1/2} r, output D^{1/2} e), where the coefficient's dynamic
    range no longer sets the scale of the error."""
    n = int(pr.meta["n"])
    H = W = n
    jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    node = (jj * (n + 1) + ii + 1).ravel()
    d = dof_table[node]
    pix2dof = np.where(d > 0, d - 1, -1).astype(np.int32)
    ndof = int((dof_table > 0).sum())
    if int((pix2dof >= 0).sum()) != ndof:
        raise ValueError("some DOFs are not on the pixel grid")
    dof2pix = np.zeros(ndof, np.int64)
    dof2pix[pix2dof[pix2dof >= 0]] = np.nonzero(pix2dof >= 0)[0]
    # log(nu) per cell, averaged to nodes, centred over the DOFs
    K = int(pr.meta["ncls"])
    ntri = int(pr.meta["ntri"])
    cls = (pr.region[:ntri] - 1) % K
    lognu_cell = np.log(np.asarray(pr.meta["cls_val"])[cls[0::2]]).reshape(n, n)  # [cj, ci]
    pad = np.pad(lognu_cell, 1, mode="edge")
    node_lognu = 0.25 * (pad[:-1, :-1] + pad[1:, :-1] + pad[:-1, 1:] + pad[1:, 1:])  # (n+1)^2
    f = node_lognu[:n, :n]
    mask = (pix2dof >= 0).reshape(H, W).astype(np.float32)
    f = (f - f[mask > 0].mean()) * FEAT_SCALE
    ds = diag_scale(A) if A is not None else None
    return GridMaps(H, W, ndof, pix2dof, dof2pix, mask, f.astype(np.float32), ds)


def to_grid(v: np.ndarray, maps: GridMaps) -> np.ndarray:
    g = np.zeros(int(np.prod(maps.shape)), np.float64)
    g[maps.dof2pix] = v
    return g.reshape(maps.shape)


def from_grid(g: np.ndarray, maps: GridMaps) -> np.ndarray:
    return np.asarray(g, np.float64).ravel()[maps.dof2pix]


# ---------------------------------------------------------------------------
# data


def random_field_problem(n: int, seed: int):
    """Seeded coefficient field: half checkerboard {1, 1e-3} (8..32 blocks),
    half lognormal with sigma drawn from [0.5, 2]."""
    from . import problems as P
    rng = np.random.default_rng(10_000 + seed)
    if seed % 2 == 0:
        nb = int(rng.choice([8, 16, 32]))
        return P.poisson2d(n, seed=seed, nblocks=nb, field_kind="checker")
    sigma = float(rng.uniform(0.5, 2.0))
    return P.poisson2d(n, seed=seed, field_kind="lognormal", sigma=sigma, nbins=64)


def _solver(**kw):
    from . import linalg as LA
    opts = dict(Algorithm=8, Preconditioner=6, Krylov_Size=30, Nb_Iter_Max=20000,
                Stopping_Test=1e-10, Renumbering_Technique=1, Orthogonalization=1)
    opts.update(kw)
    return LA.Solver(**opts)


def random_multiscale(H: int, W: int, rng) -> np.ndarray:
    """Random field with a power-law spectrum (1 + |k|^2/k0^2)^(-alpha/2),
    alpha in [0, 3] (white noise to smooth), k0 in [1, 8] grid wavenumbers."""
    w = rng.standard_normal((H, W))
    ky = np.fft.fftfreq(H) * H
    kx = np.fft.rfftfreq(W) * W
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    alpha = rng.uniform(0.0, 3.0)
    k0 = rng.uniform(1.0, 8.0)
    f = np.fft.irfft2(np.fft.rfft2(w) * (1.0 + k2 / k0 ** 2) ** (-alpha / 2), s=(H, W))
    return f / max(f.std(), 1e-30)


def generate_pairs(n: int, nfields: int, seed0: int = 0, ks=(0, 3, 10, 30),
                   rhs_per_field: int = 10, log=None, **solver_kw):
    """Residual/error pairs from the plain solver's own iterates.

    For each seeded field and each of ``rhs_per_field`` known solutions x*
    (random multi-scale fields of D^{1/2} x*), b = A x* is formed on the device and the
    plain ILU(0)-GMRES(30) is stopped after k iterations (factorisation
    reused); the pair is (r_k, e_k) = (b - A x_k, x* - x_k), k = 0 meaning
    x_0 = 0.  Returns grid arrays r, e, feat (P, H, W) float32 and the
    mask (H, W)."""
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
        s = _solver(**solver_kw)
        first = True
        y = LA.Vector(fe.N)
        for q in range(rhs_per_field):
            xs = from_grid(random_multiscale(maps.H, maps.W, rng), maps) * ds
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
                # in the network's (diagonally scaled) space
                rs.append(to_grid(ds * (bh - y.get()), maps).astype(np.float32))
                es.append(to_grid((xs - xk) / ds, maps).astype(np.float32))
                fs.append(maps.feat)
        if log and (f + 1) % max(1, nfields // 10) == 0:
            log(f"pairs: field {f + 1}/{nfields}, {len(rs)} pairs, "
                f"{time.perf_counter() - t0:.1f} s")
    return np.stack(rs), np.stack(es), np.stack(fs), mask


def normalise_pairs(r, e, mask):
    """r -> unit RMS over the DOF pixels; e -> unit 2-norm (the loss target)."""
    m = mask[None]
    nd = float(mask.sum())
    ax = tuple(range(1, r.ndim))
    rn = np.sqrt((r * r * m).sum(axis=ax, keepdims=True) / nd)
    en = np.sqrt((e * e * m).sum(axis=ax, keepdims=True))
    rn[rn == 0] = 1
    en[en == 0] = 1
    return (r * m / rn).astype(np.float32), (e * m / en).astype(np.float32)


def train(net: UNet, r, e, feat, mask, steps: int, batch: int, lr: float = 1e-3,
          seed: int = 0, log=None, log_every: int = 100, augment: bool = True):
    """Adam on the device; cosine learning-rate decay; random batches with
    optional flips (the operator's grid symmetries up to the diagonal of the
    triangulation).  Returns the loss history."""
    rng = np.random.default_rng(seed)
    rn, en = normalise_pairs(r, e, mask)
    hist = []
    t0 = time.perf_counter()
    for s in range(steps):
        idx = rng.integers(0, rn.shape[0], size=batch)
        rb, eb, fb = rn[idx], en[idx], feat[idx]
        if augment and rng.random() < 0.5:
            # transpose (swap x and y) maps the triangulation n00-n10-n11 /
            # n00-n11-n01 onto itself; the mask is symmetric
            rb, eb, fb = rb.transpose(0, 2, 1), eb.transpose(0, 2, 1), fb.transpose(0, 2, 1)
        cur = lr * 0.5 * (1 + np.cos(np.pi * s / steps))
        loss = net.train_step(rb, fb, eb, mask, cur)
        hist.append(loss)
        if log and ((s + 1) % log_every == 0 or s == 0):
            log(f"train step {s + 1}/{steps}: loss {np.mean(hist[-log_every:]):.4f} "
                f"lr {cur:.2e} ({time.perf_counter() - t0:.1f} s)")
    return np.asarray(hist)


# ---------------------------------------------------------------------------
# solve


def solve_learned(A, b, x, net: UNet, maps: GridMaps, rtol: float = 1e-10,
                  maxits: int = 1000, krylov: int = 64, precision: int = 0,
                  fallback: bool = True, two_stage: bool = False, ilu_ordering: int = 0):
    """FGMRES(krylov) right-preconditioned by the network (two_stage: the
    network, then an ILU(0) correction of the residual it leaves); if it stops
    above ``rtol`` (true FP64 residual), ILU(0)-GMRES(30) continues on the
    residual equation A d = b - A x.  Returns a dict with both stages'
    statistics."""
    from . import linalg as LA
    opts = dict(Algorithm=9, Preconditioner=6 if two_stage else 0, Krylov_Size=krylov,
                Nb_Iter_Max=maxits, Stopping_Test=rtol, Renumbering_Technique=1 if two_stage else 0,
                Scaling=0, Orthogonalization=1, Precision=precision, ILU_Ordering=ilu_ordering)
    if precision:
        opts.update(Inner_Tolerance=1e-5, Refinement_Steps=30)
    s = LA.Solver(**opts)
    net.attach(s, maps)
    out = {"learned": None, "fallback": None}
    t0 = time.perf_counter()
    try:
        info = LA.solve(A, b, s, x)
    except LA.GdpError as err:
        info = s.info()
        info["error"] = str(err)
    out["learned"] = info
    t1 = time.perf_counter()
    out["t_learned_s"] = t1 - t0
    ok = info.get("rel_res", 1.0) <= rtol and info.get("status", -1) == 0
    if not ok and fallback:
        N = maps.ndof
        y = LA.Vector(N)
        LA.prod_matrix_vector(A, x, y)
        r = LA.Vector(N, b.get() - y.get())
        d = LA.Vector(N)
        bn = np.linalg.norm(b.get())
        rn = np.linalg.norm(r.get())
        fb = _solver(Orthogonalization=1, Stopping_Test=max(rtol * bn / max(rn, 1e-300), 1e-14))
        finfo = LA.solve(A, r, fb, d)
        x.set(x.get() + d.get())
        out["fallback"] = finfo
    out["t_total_s"] = time.perf_counter() - t0
    y = LA.Vector(maps.ndof)
    LA.prod_matrix_vector(A, x, y)
    out["rel_res"] = float(np.linalg.norm(b.get() - y.get()) / np.linalg.norm(b.get()))
    return out


# ---------------------------------------------------------------------------
# evaluation and command line


def plain_solve(A, b, N, ordering: int = 0):
    """The plain method: ILU(0)-GMRES(30) to 1e-10 (ordering 0: the
    reference's RCMK renumbering; 1: multicolour ILU ordering)."""
    from . import linalg as LA
    s = _solver(ILU_Ordering=ordering, Orthogonalization=1)
    x = LA.Vector(N)
    t0 = time.perf_counter()
    info = LA.solve(A, b, s, x)
    info["t_wall_ms"] = 1e3 * (time.perf_counter() - t0)
    return x, info


def evaluate(params: np.ndarray, n: int, seeds, log=None, odd: bool = False):
    """Learned two-stage FGMRES against plain ILU(0)-GMRES on unseen fields
    (the problem's own source term).  Times are wall-clock of the solve call
    (set-up included), each after one untimed warm-up solve."""
    from . import linalg as LA
    rows = []
    for seed in seeds:
        pr = random_field_problem(n, seed)
        fe, A, b = LA.assemble(pr)
        maps = grid_maps(pr, fe.dof_table(0), A)
        net = UNet(n, n, batch=2, seed=0)
        net.set_params(params)
        net.set_mode(odd=odd)
        row = {"n": n, "seed": seed, "field": "checker" if seed % 2 == 0 else "lognormal",
               "dofs": fe.N}
        for ordering, key in ((0, "plain_rcmk"), (1, "plain_multicolour")):
            plain_solve(A, b, fe.N, ordering)
            x, info = plain_solve(A, b, fe.N, ordering)
            row[key] = {"its": info["its"], "ms": round(info["t_wall_ms"], 2),
                        "rel_res": info["rel_res"]}
        for ordering, key in ((0, "learned_rcmk"), (1, "learned_multicolour")):
            x = LA.Vector(fe.N)
            solve_learned(A, b, x, net, maps, maxits=1000, two_stage=True, ilu_ordering=ordering)
            x = LA.Vector(fe.N)
            out = solve_learned(A, b, x, net, maps, maxits=1000, two_stage=True,
                                ilu_ordering=ordering)
            row[key] = {"its": out["learned"]["its"], "ms": round(1e3 * out["t_total_s"], 2),
                        "fallback_its": None if out["fallback"] is None else out["fallback"]["its"],
                        "rel_res": out["rel_res"]}
        conv = [row[k]["ms"] for k in ("plain_rcmk", "plain_multicolour") if row[k]["rel_res"] <= 1e-10]
        best_plain = min(conv) if conv else float("nan")
        best_learned = min(row["learned_rcmk"]["ms"], row["learned_multicolour"]["ms"])
        row["speedup_vs_plain_rcmk"] = round(row["plain_rcmk"]["ms"] / best_learned, 2)
        row["speedup_vs_best_plain"] = round(best_plain / best_learned, 2)
        rows.append(row)
        if log:
            log(f"eval n={n} seed={seed} ({row['field']}): " + ", ".join(
                f"{k} {row[k]['its']} its {row[k]['ms']} ms" for k in
                ("plain_rcmk", "plain_multicolour", "learned_rcmk", "learned_multicolour"))
                + f"; speedup {row['speedup_vs_plain_rcmk']}x / {row['speedup_vs_best_plain']}x")
    return rows


def main(argv=None):
    import argparse
    import json
    import sys
    from . import linalg as LA
    ap = argparse.ArgumentParser(description="train / evaluate the learned preconditioner")
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--fields", type=int, default=500)
    ap.add_argument("--rhs", type=int, default=10)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--weights", default=None)
    ap.add_argument("--eval-n", type=int, nargs="*", default=[128, 256])
    ap.add_argument("--eval-seeds", type=int, nargs="*", default=[90000, 90001, 90002, 90003])
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)

    def log(*m):
        print(*m, file=sys.stderr, flush=True)

    LA.init(0)
    if a.cmd == "train":
        r, e, f, mask = generate_pairs(a.n, a.fields, seed0=a.seed, rhs_per_field=a.rhs, log=log,
                                       ILU_Ordering=1)
        log(f"{r.shape[0]} pairs")
        net = UNet(a.n, a.n, batch=a.batch, seed=a.seed)
        hist = train(net, r, e, f, mask, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed,
                     log=log, log_every=500)
        del r, e, f
        out = a.weights or "unet_poisson2d.npz"
        net.save(out)
        log(f"saved {out}; final loss {hist[-200:].mean():.4f}")
        params = net.get_params()
    else:
        with np.load(a.weights, allow_pickle=False) as z:
            params = z["params"]
    rows = []
    for n in a.eval_n:
        rows += evaluate(params, n, a.eval_seeds, log=log)
    if a.json:
        with open(a.json, "w") as fh:
            json.dump(rows, fh, indent=1)


if __name__ == "__main__":
    main()
