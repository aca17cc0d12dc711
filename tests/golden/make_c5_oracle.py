"""Oracle fixture for the 90 x 90-node wire grid (N = 16,020).

Runs oracle/necpp_oracle.c (the CPU restatement; about 4 minutes on 8
threads) once and stores Z_in, the source-segment current and norms of the
current vector in c5_oracle.json, and the currents in c5_oracle_current.npy.
The GPU tests compare the FP64 and mixed-precision paths with them.
"""
import json, os, sys, time
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import numpy as np
from necpp_amd import decks
from oracle import pyoracle as po
deck = decks.grid(90)
t = time.time()
r = po.run_deck(deck, nthreads=8, with_rp=False)
cur = np.asarray(r["current"])
out = {"n": len(deck["wires"]), "zin": [float(r["zin"][0].real), float(r["zin"][0].imag)],
       "current_sum_abs": float(np.sum(np.abs(cur))), "current_l2": float(np.linalg.norm(cur)),
       "source_segment_number": (sg := r["geometry"].segment_number(*deck["source"])),
       "current_src": [float(cur[sg - 1].real), float(cur[sg - 1].imag)],
       "timing": r["timing"], "wall_s": time.time() - t}
np.save(os.path.join(HERE, 'c5_oracle_current.npy'), cur)
json.dump(out, open(os.path.join(HERE, 'c5_oracle.json'), 'w'), indent=1)
print(out, flush=True)
