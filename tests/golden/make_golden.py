"""Writes the C1 golden fixtures (tests/golden/c1_*.npz).

C1: V 256 x 128, r = 16, 200 iterations from the seeded generator (seed 0x5EED + 1).
Euclidean: V = W0 H0 + U[0,0.01); KL: V ~ Poisson(W0 H0) with Exp(1) factors.
Expected outputs come from the FP64 CPU oracle (oracle/nmf_oracle.c), which the
CPU tests pin by properties (tests/test_oracle.py); the reference itself has no
golden vectors for this path.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gpumlib_amd import nmf as G  # noqa: E402
from oracle import oracle as orc  # noqa: E402

N, M, R, SEED, ITERS, EVERY = 256, 128, 16, 0x5EED + 1, 200, 20


def make(rule):
    V = G.gen_euclidean(N, M, R, SEED) if rule == G.EUCLIDEAN else G.gen_poisson(N, M, R, SEED)
    W0, H0 = G.gen_init(N, M, R, SEED)
    v, w, h = (np.ascontiguousarray(x, dtype=np.float64) for x in (V, W0, H0))
    objf = orc.frob if rule == G.EUCLIDEAN else orc.kl_div
    step = orc.euc_iter if rule == G.EUCLIDEAN else orc.kl_iter
    traj = [objf(N, M, R, v, w, h)]
    for _ in range(ITERS // EVERY):
        step(N, M, R, v, w, h, G.REFERENCE_EPS, iters=EVERY)
        traj.append(objf(N, M, R, v, w, h))
    name = "c1_euclidean.npz" if rule == G.EUCLIDEAN else "c1_kl.npz"
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", name), V=V, W0=W0, H0=H0, W=w, H=h,
                        objective=np.array(traj), every=EVERY, iters=ITERS, n=N, m=M, r=R, seed=SEED,
                        eps=G.REFERENCE_EPS)


if __name__ == "__main__":
    make(G.EUCLIDEAN)
    make(G.KL)
