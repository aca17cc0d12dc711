"""Oracle fixture for the 90 x 90-node wire grid (N = 16,020).

Runs oracle/necpp_oracle.c (the CPU restatement; about 4 minutes on 8
threads) once with the 91 x 361 pattern (theta 0-90, phi 0-360 in 1 degree
steps) and stores:
  c5_oracle.json          Z_in, the source-segment current, norms of the
                          current vector, the input power, the pattern's
                          statistics (max/min/mean/sd of total, RHCP, LHCP gain)
  c5_oracle_current.npy   segment currents (complex128, N)
  c5_oracle_abc.npy       current coefficients A, B, C per segment (6 N
                          doubles, the oracle's own), from which the tests
                          recompute the oracle pattern on any host
  c5_oracle_pattern.npy   total power gain (dB) of the 91 x 361 pattern,
                          phi_index * 91 + theta_index
The GPU tests and the bench's mixed-precision leg compare the FP64 and
mixed-precision paths with them.
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

deck = decks.grid(90)
t = time.time()
r = po.run_deck(deck, nthreads=8, with_rp=True)
cur = np.asarray(r["current"])
sg = r["geometry"].segment_number(*deck["source"])
out = {"n": len(deck["wires"]), "zin": [float(r["zin"][0].real), float(r["zin"][0].imag)],
       "current_sum_abs": float(np.sum(np.abs(cur))), "current_l2": float(np.linalg.norm(cur)),
       "source_segment_number": sg,
       "current_src": [float(cur[sg - 1].real), float(cur[sg - 1].imag)],
       "input_power": float(r["input_power"]),
       "rp": list(deck["rp"]),
       "pattern_stats": [float(v) for v in r["stats"]],
       "max_gain": float(r["max_gain"]),
       "timing": r["timing"], "wall_s": time.time() - t,
       "source": "oracle/necpp_oracle.c via tests/golden/make_c5_oracle.py (FP64, 8 OpenMP threads)"}
np.save(os.path.join(HERE, "c5_oracle_current.npy"), cur)
np.save(os.path.join(HERE, "c5_oracle_abc.npy"), np.asarray(r["abc"], np.float64))
np.save(os.path.join(HERE, "c5_oracle_pattern.npy"), np.asarray(r["pattern"]["pg_tot"], np.float64))
json.dump(out, open(os.path.join(HERE, "c5_oracle.json"), "w"), indent=1)
print(out, flush=True)
