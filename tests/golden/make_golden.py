"""Write the golden fixtures of the putative matcher (run from the repository root).

The reference ships no test data, so these fixtures are produced by this project's
deterministic generator (eth-v3d-lgpl_amd/csrc/v3dm_synth.c) and its exact CPU oracle
(oracle/v3dm_oracle.c).  They pin both against drift:
  pairs_n{32,96,512}.npz  uint8 left/right descriptors, the match lists at
                          (0.9, 0.8) and (0.9, no ratio), and colMax / rowMax
  hashes.json             SHA-256 of the 4096 x 4096 inputs (uint8) and of their match
                          list, for the 4096-feature single-pair workload (seed 0)
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "eth-v3d-lgpl_amd")]

from oracle import oracle  # noqa: E402
from v3dmatch import synth  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    for n, seed in [(32, 11), (96, 12), (512, 13)]:
        qL, qR, truth = synth.make_pair(n, n, 0.4, config=0, pair=seed)
        L, R = synth.dequantise(qL), synth.dequantise(qR)
        m = oracle.find_putative_matches(L, R, 0.9, 0.8)
        m_nr = oracle.find_putative_matches(L, R, 0.9, -1.0)
        cm, rm = oracle.maxima(L, R, 0.9, 0.8)
        np.savez_compressed(os.path.join(HERE, f"pairs_n{n}.npz"), qL=qL, qR=qR, truth=truth, matches=m,
                            matches_no_ratio=m_nr, colMax=cm, rowMax=rm)
        print(f"n={n}: {len(m)} matches, {len(m_nr)} without ratio test")
    qL, qR, _ = synth.make_pair(4096, 4096, 0.4, config=1, pair=0)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    m = oracle.find_putative_matches(L, R, 0.9, 0.8)
    h = {
        "workload": "single pair 4096 x 4096, copy fraction 0.4, config 1, pair 0",
        "sha256_qL": hashlib.sha256(qL.tobytes()).hexdigest(),
        "sha256_qR": hashlib.sha256(qR.tobytes()).hexdigest(),
        "sha256_L_float": hashlib.sha256(L.tobytes()).hexdigest(),
        "n_matches": int(len(m)),
        "sha256_matches_int32": hashlib.sha256(np.ascontiguousarray(m, np.int32).tobytes()).hexdigest(),
    }
    with open(os.path.join(HERE, "hashes.json"), "w") as f:
        json.dump(h, f, indent=1)
    print(h)


if __name__ == "__main__":
    main()
