"""Write the FP64 reference solution of the curl-curl 48^3 system (663,552
tetrahedra, 753,552 edge DOFs) as a small fixture.

    python tests/golden/make_c2_solution.py

The system is the oracle's assembly of ``problems.curlcurl3d(48)``; the solve
is the reference's own Sparskit Fortran (oracle/_ref, compiled here from
contrib/Sparskit) driven through solve_matrix's sequence: CG (Algorithm 1)
with diagonal scaling (Scaling 1), RCMK, to 1e-10 -- the method that reaches
1e-10 on this operator.  ILU(0)-GMRES stagnates near 6e-7 on it in the
reference's arithmetic too.  The full vector is 6 MB, so the fixture keeps
every 97th entry, the norms and the iteration count; the device solutions are
compared against these.
"""
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from getdp_amd import problems as P  # noqa: E402
from oracle import pyoracle as O  # noqa: E402

STRIDE = 97


def main():
    if O.ref() is None:
        raise SystemExit("oracle/_ref not built (needs the reference sources)")
    pr = P.curlcurl3d(48)
    ia, ja, a, b, tab = O.assemble(pr)
    t0 = time.perf_counter()
    x, info, r = O.ref_solve(ia, ja, a, b, algorithm=1, preconditioner=0, scaling=1, renumbering=1,
                             maxits=200000)
    dt = time.perf_counter() - t0
    res = O.amux(ia, ja, a, x) - b
    out = dict(stride=np.int64(STRIDE), x_sampled=x[::STRIDE].copy(), n=np.int64(x.shape[0]),
               norm2=np.float64(np.linalg.norm(x)), norm_inf=np.float64(np.abs(x).max()),
               its=np.int64(info["its"]), status=np.int64(info["status"]),
               rel_res=np.float64(np.linalg.norm(res) / np.linalg.norm(b)), seconds=np.float64(dt))
    np.savez_compressed(Path(__file__).with_name("c2_curlcurl48_cg.npz"), **out)
    print({k: (v if np.ndim(v) == 0 else v.shape) for k, v in out.items()})


if __name__ == "__main__":
    main()
