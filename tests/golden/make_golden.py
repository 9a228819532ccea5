"""Write the golden fixtures of the Sparskit path from the reference's own
Fortran kernels (oracle/_ref, compiled here from contrib/Sparskit).

    python tests/golden/make_golden.py

The inputs are the small assembled systems of the three problem families
(poisson2d(8), elasticity3d(2), curlcurl3d(2)) as the oracle assembles them;
the outputs are what the reference's kernels return for them: the RCMK
permutation (cmkreord + sortcol), the FP32 ILU(0) factors in MSR form (ilu0),
and the solutions / iteration counts / status of solve_matrix's sequence for
GMRES(30), FGMRES(30), CG with diagonal scaling, GMRES(10) without
renumbering, BCGSTAB + ILU(0), BCGSTAB with diagonal scaling, TFQMR + ILU(0),
unpreconditioned TFQMR, GetDP's default ILUTP(20)-GMRES(40) and ILUT(5, 1e-3).
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from getdp_amd import problems as P  # noqa: E402
from oracle import pyoracle as O  # noqa: E402

CASES = {
    "poisson2d_8": lambda: P.poisson2d(8),
    "elasticity3d_2": lambda: P.elasticity3d(2),
    "curlcurl3d_2": lambda: P.curlcurl3d(2),
}
SOLVES = {
    "gmres_ilu0": dict(algorithm=8, preconditioner=6, krylov=30),
    "fgmres_ilu0": dict(algorithm=9, preconditioner=6, krylov=30),
    "cg_jacobi": dict(algorithm=1, preconditioner=0, scaling=1),
    "gmres_none_natural": dict(algorithm=8, preconditioner=0, renumbering=0, krylov=10),
    "bcgstab_ilu0": dict(algorithm=5, preconditioner=6),
    "bcgstab_jacobi": dict(algorithm=5, preconditioner=0, scaling=1),
    "tfqmr_ilu0": dict(algorithm=6, preconditioner=6),
    "tfqmr_none": dict(algorithm=6, preconditioner=0),
    "gmres40_ilutp_default": dict(algorithm=8, preconditioner=2, krylov=40),
    "gmres40_ilut_fill5": dict(algorithm=8, preconditioner=1, krylov=40, nb_fill=5, droptol=1e-3),
}


def main():
    if O.ref() is None:
        raise SystemExit("oracle/_ref not built (needs the reference sources)")
    out = {}
    for name, make in CASES.items():
        pr = make()
        ia, ja, a, b, tab = O.assemble(pr)
        out[f"{name}/ia"], out[f"{name}/ja"], out[f"{name}/a"], out[f"{name}/b"] = ia, ja, a, b
        perm, rperm, ib, jb, bb, nlev = O.ref_rcmk(ia, ja, a)
        out[f"{name}/rcmk_perm"] = perm
        alu, jlu, ju, ierr = O.ref_ilu0(ib, jb, bb)
        out[f"{name}/ilu_alu"], out[f"{name}/ilu_jlu"], out[f"{name}/ilu_ju"] = alu, jlu, ju
        for sname, kw in SOLVES.items():
            x, info, r = O.ref_solve(ia, ja, a, b, maxits=20000, **kw)
            out[f"{name}/{sname}/x"] = x
            out[f"{name}/{sname}/its"] = np.int64(info["its"])
            out[f"{name}/{sname}/status"] = np.int64(info["status"])
            out[f"{name}/{sname}/true_res"] = np.float64(info["true_res"])
    path = Path(__file__).with_name("sparskit_ref.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({path.stat().st_size} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
