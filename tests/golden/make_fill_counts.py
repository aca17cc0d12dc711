"""Operation counts of the fill and the pattern sum for the two wire-grid
workloads the bench quotes (41 x 41 nodes, N = 3,280, and 90 x 90 nodes,
N = 16,020, both at 299.792458 MHz with the 91 x 361 pattern).

The counts are the oracle's own tally (oracle/necpp_oracle.c, TALLY, kept
per basic block as the code runs): every +, -, *, / and square root is one
flop, divisions and square roots are also given alone, and sin, cos and log
are "specials" outside the flop count.  The fill counts the oracle's
ora_fill (pairs, branches, Romberg steps and refinements as they happen);
the pattern counts the direction-segment sums of ffld.  The bench divides
them by the measured kernel time for the achieved FLOP/s.

Writes fill_counts.json.  About a minute on 8 threads.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import numpy as np  # noqa: E402

from necpp_amd import decks  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

KEYS = ("pairs", "far_pairs", "near_pairs", "gx_calls", "gxx_calls", "intx_calls", "gf_calls", "scatter_terms",
        "intx_steps", "intx_refines", "intx_self", "flops", "divs", "sqrts", "specials")


def counts(nodes, nthreads=8):
    deck = decks.grid(nodes)
    g = po.geometry_from_deck(deck)
    g.scale(deck["freq_mhz"])
    t = time.time()
    _, fc = g.fill(nthreads=nthreads)
    fill_s = time.time() - t
    rp = po.rp_from_card(deck["rp"])
    t = time.time()
    pc = g.pattern_tally(np.zeros(6 * g.n), deck["freq_mhz"], rp, nthreads=nthreads)
    pat_s = time.time() - t
    return {"n": g.n, "freq_mhz": deck["freq_mhz"],
            "pattern_directions": rp["n_theta"] * rp["n_phi"],
            "fill": {k: int(fc[k]) for k in KEYS},
            "pattern": {k: int(v) for k, v in pc.items()},
            "oracle_wall_s": {"fill": fill_s, "pattern_tally": pat_s}}


def main():
    out = {"convention": "flops = every +, -, *, /, sqrt executed (divs and sqrts also alone); "
                         "specials = sin, cos, log (not in flops); negation, fabs, comparisons and "
                         "expressions of constants count nothing",
           "source": "oracle/necpp_oracle.c tally via tests/golden/make_fill_counts.py",
           "grid41": counts(41), "grid90": counts(90)}
    json.dump(out, open(os.path.join(HERE, "fill_counts.json"), "w"), indent=1)
    print(json.dumps(out, indent=1), flush=True)


if __name__ == "__main__":
    main()
