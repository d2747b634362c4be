"""The committed C1 fixtures: the generator reproduces the inputs bit-exactly and
the oracle reproduces the expected outputs (regression of both)."""
import os

import numpy as np
import pytest

from conftest import rel_fro
from gpumlib_amd import nmf as G

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name,rule", [("c1_euclidean.npz", G.EUCLIDEAN), ("c1_kl.npz", G.KL)])
def test_fixture_regenerates(native, orc, name, rule):
    f = np.load(os.path.join(GOLDEN, name))
    n, m, r, seed = int(f["n"]), int(f["m"]), int(f["r"]), int(f["seed"])
    V = G.gen_euclidean(n, m, r, seed) if rule == G.EUCLIDEAN else G.gen_poisson(n, m, r, seed)
    W0, H0 = G.gen_init(n, m, r, seed)
    assert np.array_equal(V, f["V"]) and np.array_equal(W0, f["W0"]) and np.array_equal(H0, f["H0"])
    v, w, h = (np.ascontiguousarray(x, dtype=np.float64) for x in (V, W0, H0))
    (orc.euc_iter if rule == G.EUCLIDEAN else orc.kl_iter)(n, m, r, v, w, h, float(f["eps"]), iters=int(f["iters"]))
    assert rel_fro(w, f["W"]) < 1e-12 and rel_fro(h, f["H"]) < 1e-12
    traj = f["objective"]
    assert np.all(np.diff(traj) <= 0)  # monotone (Lee & Seung)


@pytest.mark.gpu
@pytest.mark.parametrize("name,rule", [("c1_euclidean.npz", G.EUCLIDEAN), ("c1_kl.npz", G.KL)])
def test_gpu_matches_fixture(native, cuda, name, rule):
    f = np.load(os.path.join(GOLDEN, name))
    n, m, r = int(f["n"]), int(f["m"]), int(f["r"])
    d = G.DeviceNMF(rule, f["V"], f["W0"], f["H0"], n, m, r, eps=float(f["eps"]))
    kind = G.OBJ_FROBENIUS if rule == G.EUCLIDEAN else G.OBJ_KL
    every = int(f["every"])
    for k, expect in enumerate(f["objective"][:-1]):
        got = d.iterate(every, objective=kind)
        assert abs(got - expect) / abs(expect) < 1e-4, (k, got, expect)
    assert rel_fro(d.get_w(), f["W"]) < 1e-4
    assert rel_fro(d.get_h(), f["H"]) < 1e-4
