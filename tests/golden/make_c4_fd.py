"""Write tests/golden/c4_fd_oracle.npz: the oracle's objective and central
finite-difference gradients for the differentiation workload.

For every design row p of necpp_amd.decks.c4_params() (256 rows of 44
parameters, 15-element Yagi, 31 segments per element, N = 465, 144 MHz):
  J(p)  = -G(theta=90, phi=0) + 0.01 |Z_in - 50|^2 from the oracle's RP
          gain (pg_tot) and input impedance;
  fd    = central differences with the step h = 1e-6 |p| (Euclidean norm of
          the row) for every component;
  fd2   = the same with 2h;
  fd4   = (4 fd - fd2) / 3, the fourth-order central difference of the same
          five points: the radii (~0.003 wavelengths) take a step of 5e-4 of
          their size, and J's third derivative in a radius makes fd's own
          h^2 truncation error reach 1e-6 of the gradient, which fd4 cancels
          (a smaller step is worse: J carries ~4e-11 relative rounding noise).
The fill's geometry-dependent decisions (the near/far switch r >= rkh of the
field kernel per segment pair, every accept/refine choice of the adaptive
quadrature, the self-term integrand's small-argument branch) are recorded at
p and replayed at every p +- h: the differences are those of J with the
decisions held fixed, the function whose derivative an adjoint of the same
code computes (without it, a difference that straddles a decision flip is
off by the jump / 2h).  Each replay must consume exactly the recorded tape.
Run: python tests/golden/make_c4_fd.py [--designs N] [--procs P]
"""
import argparse
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from necpp_amd import decks  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

REL_STEP = 1e-6
WEIGHT, Z0 = 0.01, 50.0


def oracle_J(p):
    r = po.run_deck(decks.c4_deck(p), nthreads=1, with_rp=True)
    return -float(r["pattern"]["pg_tot"][0]) + WEIGHT * abs(complex(r["zin"][0]) - Z0) ** 2


def design_task(args):
    """J(p) with the decisions recorded, then J at p +- h, p +- 2h for every
    component with them replayed."""
    d, p = args
    p = np.array(p, dtype=float)
    po.tape("record")
    J0 = oracle_J(p)
    n_rec = po.tape_stats()[0]
    out = {}
    h = REL_STEP * float(np.linalg.norm(p))
    for k in range(len(p)):
        for s in (1, -1, 2, -2):
            q = p.copy()
            q[k] += s * h
            po.tape("replay")
            v = oracle_J(q)
            rec, used, bad = po.tape_stats()
            if bad or used != rec or rec != n_rec:
                raise RuntimeError(f"design {d} component {k}: the replay did not follow the recorded decisions")
            out[(k, s)] = (v, q[k] - p[k])
    po.tape("off")
    # the point itself again with the tape replayed: the same J
    po.tape("replay")
    J1 = oracle_J(p)
    po.tape("off")
    if J1 != J0:
        raise RuntimeError(f"design {d}: replay at p changed J ({J0} -> {J1})")
    fd = np.zeros(len(p))
    fd2 = np.zeros(len(p))
    for k in range(len(p)):
        (vp, hp), (vm, hm) = out[(k, 1)], out[(k, -1)]
        fd[k] = (vp - vm) / (hp - hm)
        (vp, hp), (vm, hm) = out[(k, 2)], out[(k, -2)]
        fd2[k] = (vp - vm) / (hp - hm)
    return d, J0, fd, fd2, n_rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--designs", type=int, default=256)
    ap.add_argument("--procs", type=int, default=8)
    ap.add_argument("--out", default=os.path.join(HERE, "c4_fd_oracle.npz"))
    a = ap.parse_args()
    po.build()
    P = decks.c4_params(count=a.designs)
    J = np.zeros(len(P))
    fd = np.zeros(P.shape)
    fd2 = np.zeros(P.shape)
    ndec = np.zeros(len(P), np.int64)
    t0 = time.time()
    with Pool(a.procs) as pool:
        for i, (d, j0, f1, f2, nd) in enumerate(pool.imap_unordered(design_task, list(enumerate(P)))):
            J[d], fd[d], fd2[d], ndec[d] = j0, f1, f2, nd
            if (i + 1) % 16 == 0:
                print(f"{i + 1}/{len(P)} designs, {time.time() - t0:.0f} s", flush=True)
    fd4 = (4.0 * fd - fd2) / 3.0
    np.savez_compressed(a.out, J=J, fd=fd, fd2=fd2, fd4=fd4, rel_step=REL_STEP, seed=decks.ADJOINT_SEED,
                        params_sum=P.sum(axis=1), decisions=ndec, frozen=True)
    cons = np.linalg.norm(fd - fd2, axis=1) / np.linalg.norm(fd, axis=1)
    print("wrote", a.out, "in", round(time.time() - t0), "s; |fd(h) - fd(2h)| / |fd| max", cons.max(),
          "median", np.median(cons))


if __name__ == "__main__":
    main()
