"""Batched assembly and ILU(0)-GMRES over hundreds of independent 2D problems
(host driver of ``gdp_batch_*`` in include/gdp_hip.h).

The instances' meshes are concatenated into one formulation, so the device
assembles all of them in one pass and the DOFs of instance i form the
diagonal block [off[i], off[i+1]) of one matrix.  ``gdp_batch_run`` replays
one captured hipGraph: assembly, renumbering of values, ILU(0) factorisation,
one persistent solver kernel per launch configuration, unpermutation.

The launch configuration of an instance (workgroup size, Gram-Schmidt
variant, ILU ordering) comes from :func:`autotune`, which times candidate
configurations per size class on the device and stores the winners in
``tuning/batch_gfx950.json``.
"""
from __future__ import annotations

import ctypes as C
import json
import time
from pathlib import Path
from typing import List, Optional

import numpy as np

from . import linalg as LA
from . import problems as P
from ._lib import check, dptr, iptr, lib

TUNING = Path(__file__).resolve().parent / "tuning" / "batch_gfx950.json"
BUCKETS = [(32, 63), (64, 95), (96, 127), (128, 159), (160, 192)]
CANDIDATES = [(t, o, d) for t in (256, 512, 1024) for o in (0, 1) for d in (0, 1)]
DEFAULT_CFG = (256, 1, 0)


def make_instances(count: int = 512, seed: int = 0, nmin: int = 32, nmax: int = 192) -> List:
    """``count`` 2D Poisson problems (the unit-square checkerboard problem of
    ``problems.poisson2d``) with sides drawn uniformly from [nmin, nmax] and
    independently seeded 16x16 checkerboards."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=count)
    return [P.poisson2d(int(n), seed=seed * 100_003 + i, nblocks=16, field_kind="checker")
            for i, n in enumerate(sizes)]


def concat(problems) -> "P.Problem":
    """One formulation over all instances: nodes and elements appended in
    instance order (so each instance's DOFs are consecutive), shared regions
    and coefficient tables (every instance uses the same class values)."""
    p0 = problems[0]
    for p in problems[1:]:
        if p.meta["ncls"] != p0.meta["ncls"] or not np.array_equal(p.meta["cls_val"], p0.meta["cls_val"]):
            raise ValueError("instances must share the coefficient classes")
    xyz, etype, region, conn, fixed = [], [], [], [], []
    noff = 0
    for p in problems:
        xyz.append(p.xyz)
        etype.append(p.etype)
        region.append(p.region)
        c = p.conn.copy()
        c[c > 0] += noff
        conn.append(c)
        fixed.append(p.spaces[0].fixed_entity + noff)
        noff += p.nnodes
    fixed = np.concatenate(fixed).astype(np.int32)
    s0 = p0.spaces[0]
    space = P.Space(s0.bf, s0.numtype, list(s0.support), fixed, np.zeros(fixed.shape[0]))
    return P.Problem(f"batch_{len(problems)}", np.concatenate(xyz), np.concatenate(etype),
                     np.concatenate(region), np.concatenate(conn), [space], list(p0.terms),
                     meta=dict(ninst=len(problems)))


def bucket_of(n: int) -> int:
    for k, (lo, hi) in enumerate(BUCKETS):
        if lo <= n <= hi:
            return k
    return 0 if n < BUCKETS[0][0] else len(BUCKETS) - 1


def load_tuning() -> Optional[dict]:
    if TUNING.exists():
        return json.loads(TUNING.read_text())
    return None


class Batch:
    """Device batch: concatenated formulation + ``gdp_batch`` handle."""

    def __init__(self, problems, config=None, krylov: int = 30, rtol: float = 1e-10,
                 maxits: int = 20000, renumbering: int = 1, fe=None):
        self.problems = problems
        self.ninst = len(problems)
        t0 = time.perf_counter()
        self.pr = concat(problems)
        self.fe = fe or LA.FESystem(self.pr)
        self.t_setup_host = time.perf_counter() - t0
        dt = self.fe.dof_table(0)
        off = [0]
        nstart = 0
        for p in problems:
            off.append(off[-1] + int((dt[nstart + 1:nstart + p.nnodes + 1] > 0).sum()))
            nstart += p.nnodes
        self.off = np.asarray(off, np.int32)
        assert self.off[-1] == self.fe.N
        self.N = self.fe.N
        self.A = LA.Matrix(self.N)
        self.b = LA.Vector(self.N)
        self.x = LA.Vector(self.N)
        self.solver = LA.Solver(Algorithm=8, Preconditioner=6, Krylov_Size=krylov,
                                Nb_Iter_Max=maxits, Stopping_Test=rtol,
                                Renumbering_Technique=renumbering)
        if config is None:
            tun = load_tuning()
            config = [self._cfg_from(tun, int(p.meta["n"])) for p in problems]
        cfg = np.ascontiguousarray(np.asarray(config, np.int32).reshape(-1))
        assert cfg.size == 3 * self.ninst
        self.config = cfg.reshape(-1, 3)
        self.h = C.c_void_p()
        check(lib().gdp_batch_create(self.solver.h, self.ninst, iptr(self.off), iptr(cfg),
                                     C.byref(self.h)), "gdp_batch_create")

    @staticmethod
    def _cfg_from(tun, n):
        if not tun:
            return DEFAULT_CFG
        e = tun["buckets"][bucket_of(n)]
        return (e["threads"], e["orth"], e["ordering"])

    def close(self):
        if self.h and self.h.value:
            lib().gdp_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, assemble: bool = True):
        check(lib().gdp_batch_run(self.h, self.fe.h if assemble else None, self.A.h, self.b.h,
                                  self.x.h), "gdp_batch_run")

    def info(self) -> dict:
        n = self.ninst
        its = np.zeros(n, np.int32)
        st = np.zeros(n, np.int32)
        r0 = np.zeros(n)
        r = np.zeros(n)
        tb = np.zeros(n)
        te = np.zeros(n)
        gms = C.c_double()
        check(lib().gdp_batch_info(self.h, iptr(its), iptr(st), dptr(r0), dptr(r), dptr(tb),
                                   dptr(te), C.byref(gms)), "gdp_batch_info")
        return dict(its=its, status=st, res0=r0, res=r, t_begin_us=tb, t_end_us=te,
                    graph_ms=gms.value)

    def solution(self, i: int) -> np.ndarray:
        return self.x.get()[self.off[i]:self.off[i + 1]]

    def true_rel_residuals(self) -> np.ndarray:
        """|b_i - A_i x_i| / |b_i| per instance, FP64 on the host."""
        ia, ja, a = self.A.get_csr()
        x = self.x.get()
        b = self.b.get()
        rows = np.repeat(np.arange(self.N), np.diff(ia))
        ax = np.bincount(rows, weights=a * x[ja], minlength=self.N)
        r = b - ax
        out = np.zeros(self.ninst)
        for i in range(self.ninst):
            s = slice(self.off[i], self.off[i + 1])
            nb = np.linalg.norm(b[s])
            out[i] = np.linalg.norm(r[s]) / nb if nb > 0 else np.linalg.norm(r[s])
        return out


def autotune(per_bucket: int = 24, seed: int = 7, reps: int = 1, log=None,
             candidates=CANDIDATES) -> dict:
    """Time every candidate (workgroup size, Gram-Schmidt, ILU ordering) on a
    batch of ``per_bucket`` instances of each size class and keep the fastest
    (device time of the whole graph, best of ``reps`` replays)."""
    result = {"device": "gfx950", "per_bucket": per_bucket, "buckets": []}
    for k, (lo, hi) in enumerate(BUCKETS):
        probs = make_instances(per_bucket, seed=seed + k, nmin=lo, nmax=hi)
        fe = LA.FESystem(concat(probs))
        best = None
        timings = []
        for cand in candidates:
            b = Batch(probs, config=[cand] * len(probs), fe=fe)
            b.run()
            ms = []
            for _ in range(reps):
                b.run()
                ms.append(b.info()["graph_ms"])
            inf = b.info()
            ok = bool((inf["status"] == 0).all())
            t = float(min(ms))
            timings.append({"threads": cand[0], "orth": cand[1], "ordering": cand[2], "ms": round(t, 3),
                            "converged": ok, "its_mean": float(inf["its"].mean())})
            if log:
                log(f"bucket {lo}-{hi} cand {cand}: {t:.2f} ms, its {inf['its'].mean():.0f}, ok {ok}")
            if ok and (best is None or t < best["ms"]):
                best = timings[-1]
            b.close()
        entry = dict(best or timings[0])
        entry["range"] = [lo, hi]
        entry["all"] = timings
        result["buckets"].append(entry)
    return result


def bench(count: int = 512, seed: int = 0, reps: int = 3, config=None, log=None) -> dict:
    """Assemble + solve ``count`` uneven instances; returns throughput and
    per-instance latency statistics (device clock)."""
    t0 = time.perf_counter()
    probs = make_instances(count, seed=seed)
    b = Batch(probs, config=config)
    if log:
        log(f"batch: {count} instances, {b.N} DOFs, {int((b.pr.etype == P.TRI).sum())} triangles, "
            f"host set-up {time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    b.run()
    t_first = time.perf_counter() - t1
    runs = []
    for _ in range(reps):
        t1 = time.perf_counter()
        b.run()
        runs.append((time.perf_counter() - t1, b.info()["graph_ms"]))
    inf = b.info()
    rel = b.true_rel_residuals()
    lat = inf["t_end_us"] - inf["t_begin_us"]
    wall = min(r[0] for r in runs)
    gms = min(r[1] for r in runs)
    out = {
        "instances": count,
        "dofs": int(b.N),
        "elements": int((b.pr.etype == P.TRI).sum()),
        "graph_ms": round(gms, 3),
        "wall_ms": round(1e3 * wall, 3),
        "first_run_ms": round(1e3 * t_first, 1),
        "solves_per_s": round(count / (gms * 1e-3), 1),
        "latency_us_p50": round(float(np.percentile(lat, 50)), 1),
        "latency_us_p99": round(float(np.percentile(lat, 99)), 1),
        "completion_us_p50": round(float(np.percentile(inf["t_end_us"], 50)), 1),
        "completion_us_p99": round(float(np.percentile(inf["t_end_us"], 99)), 1),
        "its_mean": float(inf["its"].mean()),
        "its_max": int(inf["its"].max()),
        "converged": int((inf["status"] == 0).sum()),
        "max_true_rel_res": float(rel.max()),
        "configs": {f"{t}/{o}/{d}": int(c) for (t, o, d), c in
                    zip(*np.unique(b.config, axis=0, return_counts=True))},
    }
    return out


def main(argv=None):
    import argparse
    import sys
    ap = argparse.ArgumentParser(description="batched assembly + ILU(0)-GMRES")
    ap.add_argument("cmd", choices=["tune", "bench"])
    ap.add_argument("--count", type=int, default=512)
    ap.add_argument("--per-bucket", type=int, default=24)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)

    def log(*m):
        print(*m, file=sys.stderr, flush=True)

    LA.init(0)
    if a.cmd == "tune":
        res = autotune(per_bucket=a.per_bucket, log=log)
        txt = json.dumps(res, indent=1)
        Path(a.out or TUNING).parent.mkdir(parents=True, exist_ok=True)
        Path(a.out or TUNING).write_text(txt)
        log(txt)
    else:
        print(json.dumps(bench(a.count, log=log)))


if __name__ == "__main__":
    main()
