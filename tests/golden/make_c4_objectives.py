"""FP64 host-oracle objectives for the mixed-precision configuration (V 16384 x 8192 =
W0 H0 + U[0,0.01) noise, r = 64, the bench's default workload and seed), written to
c4_fp64_objectives.json.

For each V -- the FP32 V, its BF16 rounding and its FP8 (E4M3, 32 x 32 tile scales)
rounding -- the Euclidean rules run in FP64 (oracle.np_euc_iter, numpy/BLAS) from the
seeded initial W, H, and the FP64 Frobenius objective against that V is recorded after
100, 500 and 1000 iterations.  The GPU test compares the BF16 / FP8 / FP32 paths with
these numbers (both gap definitions: same V, and FP32 V).  Takes minutes on a few cores.

    python tests/golden/make_c4_objectives.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from gpumlib_amd import nmf as G  # noqa: E402
from oracle import oracle as orc  # noqa: E402

N, M, R, SEED = 16384, 8192, 64, 0x5EED + 4
CHECKPOINTS = (100, 500, 1000)


def run(v, w, h, label):
    out = {}
    t0 = time.time()
    for it in range(1, CHECKPOINTS[-1] + 1):
        w, h = orc.np_euc_iter(v, w, h, G.REFERENCE_EPS)
        if it in CHECKPOINTS:
            out[str(it)] = orc.np_frob(v, w, h)
            print(f"{label}: {it} iterations, objective {out[str(it)]:.10e} ({time.time() - t0:.0f} s)", flush=True)
    return out


def main():
    V = G.gen_euclidean(N, M, R, SEED)
    W, H = G.gen_init(N, M, R, SEED)
    v = G.as_matrix(V, N, M).astype(np.float64)
    w0 = G.as_matrix(W, N, R).astype(np.float64)
    h0 = G.as_matrix(H, R, M).astype(np.float64)
    res = {"config": {"n": N, "m": M, "r": R, "seed": SEED, "eps": G.REFERENCE_EPS,
                      "generator": "gpumlib_nmf_gen_euclidean_v / gen_init (seeded SplitMix64)"},
           "objective_at": {}}
    res["objective_at"]["fp32_v"] = run(v, w0, h0, "fp32 V")
    res["objective_at"]["bf16_v"] = run(orc.bf16_round(G.as_matrix(V, N, M)), w0, h0, "bf16 V")
    res["objective_at"]["fp8_v"] = run(orc.fp8_tile_round(G.as_matrix(V, N, M)), w0, h0, "fp8 V")
    with open(os.path.join(HERE, "c4_fp64_objectives.json"), "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
