"""Write tests/golden/c4_fd_oracle.npz: the oracle's objective and central
finite-difference gradients for the differentiation workload.

For every design row p of necpp_amd.decks.c4_params() (256 rows of 44
parameters, 15-element Yagi, 31 segments per element, N = 465, 144 MHz):
  J(p)  = -G(theta=90, phi=0) + 0.01 |Z_in - 50|^2 from the oracle's RP
          gain (pg_tot) and input impedance;
  fd    = central differences with the step h = 1e-6 |p| (Euclidean norm of
          the row) for every component;
  fd2   = the same with 2h (where fd and fd2 disagree, J is not smooth at
          the step's scale: a segment pair crossing the r = rkh switch of the
          field kernel, or a quadrature decision flipping).
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


def task(args):
    d, k, s, p = args  # k < 0: J itself; s: step multiple (+-1, +-2)
    q = np.array(p, dtype=float)
    if k >= 0:
        q[k] += s * REL_STEP * float(np.linalg.norm(p))
    return d, k, s, oracle_J(q), (q[k] - p[k]) if k >= 0 else 0.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--designs", type=int, default=256)
    ap.add_argument("--procs", type=int, default=8)
    ap.add_argument("--out", default=os.path.join(HERE, "c4_fd_oracle.npz"))
    a = ap.parse_args()
    po.build()
    P = decks.c4_params(count=a.designs)
    tasks = []
    for d, p in enumerate(P):
        tasks.append((d, -1, 0, p))
        for k in range(P.shape[1]):
            for s in (1, -1, 2, -2):
                tasks.append((d, k, s, p))
    J = np.zeros(len(P))
    val = {}
    t0 = time.time()
    with Pool(a.procs) as pool:
        for i, (d, k, s, v, dp) in enumerate(pool.imap_unordered(task, tasks, chunksize=8)):
            if k < 0:
                J[d] = v
            else:
                val[(d, k, s)] = (v, dp)
            if (i + 1) % 2000 == 0:
                print(f"{i + 1}/{len(tasks)} evaluations, {time.time() - t0:.0f} s", flush=True)
    fd = np.zeros(P.shape)
    fd2 = np.zeros(P.shape)
    for d in range(len(P)):
        for k in range(P.shape[1]):
            (vp, hp), (vm, hm) = val[(d, k, 1)], val[(d, k, -1)]
            fd[d, k] = (vp - vm) / (hp - hm)
            (vp, hp), (vm, hm) = val[(d, k, 2)], val[(d, k, -2)]
            fd2[d, k] = (vp - vm) / (hp - hm)
    np.savez_compressed(a.out, J=J, fd=fd, fd2=fd2, rel_step=REL_STEP, seed=decks.ADJOINT_SEED,
                        params_sum=P.sum(axis=1))
    print("wrote", a.out, "in", round(time.time() - t0), "s")


if __name__ == "__main__":
    main()
