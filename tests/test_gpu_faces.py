"""The headless FacesNMF driver (gpumlib_amd/examples/faces_nmf.cpp) over the C++
drop-in API: training, test projection with W fixed, groups with a block-diagonal
H, and the CSV format of the reference demo (one line per column, 20 digits)."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_fro
from gpumlib_amd import nmf as G

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gpumlib_amd", "bin", "faces_nmf")


def run(args, out):
    p = subprocess.run([EXE, *args, "--out", str(out)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    return json.loads(p.stdout.strip().splitlines()[-1])


def read_csv(path, rows):
    lines = open(path).read().strip().splitlines()
    for ln in lines[:2]:
        assert len(ln.split(",")[0].split(".")[1]) == 20  # fixed, 20 digits
    cols = [np.array([float(x) for x in ln.split(",")]) for ln in lines]
    assert all(len(c) == rows for c in cols)
    return np.stack(cols, axis=1)  # rows x columns


def uniform(seed, stream, count):
    import ctypes

    from gpumlib_amd import capi

    out = np.empty(count, np.float32)
    capi.load().gpumlib_nmf_gen_uniform_f32(seed, stream, count, out.ctypes.data_as(ctypes.c_void_p))
    return out


@pytest.mark.parametrize("method", ["euclidean", "kl"])
def test_train_and_project(native, orc, cuda, tmp_path, method):
    n, m, r0, seed, r, iters, mtest = 361, 500, 20, 77, 49, 40, 100
    res = run(["--synthetic", str(n), str(m), str(r0), str(seed), "--rank", str(r), "--method", method,
               "--max-iterations", str(iters), "--tol", "-1e30", "--test-images", str(mtest)], tmp_path)
    assert res["iterations"] == iters and res["test_iterations"] == iters
    rule = G.EUCLIDEAN if method == "euclidean" else G.KL
    V = G.gen_euclidean(n, m, r0, seed) if rule == G.EUCLIDEAN else G.gen_poisson(n, m, r0, seed)
    Vm = G.as_matrix(V, n, m).astype(np.float64)
    mt = m - mtest
    w = G.as_matrix(uniform(seed, 4, n * r), n, r).astype(np.float64)
    h = G.as_matrix(uniform(seed, 5, r * mt), r, mt).astype(np.float64)
    step = orc.np_euc_iter if rule == G.EUCLIDEAN else orc.np_kl_iter
    for _ in range(iters):
        w, h = step(Vm[:, :mt], w, h, G.REFERENCE_EPS)
    assert rel_fro(read_csv(tmp_path / "W.csv", n), w) < 1e-4
    assert rel_fro(read_csv(tmp_path / "Htrain.csv", r), h) < 1e-4
    ht = G.as_matrix(uniform(seed, 6, r * mtest), r, mtest).astype(np.float64)
    wt = read_csv(tmp_path / "W.csv", n)
    for _ in range(iters):
        _, ht = step(Vm[:, mt:], wt, ht, G.REFERENCE_EPS, update_w=False)
    assert rel_fro(read_csv(tmp_path / "Htest.csv", r), ht) < 1e-4


def test_groups_block_diagonal(native, orc, cuda, tmp_path):
    n, m, r0, seed, r, group, iters = 361, 300, 12, 5, 48, 100, 30
    res = run(["--synthetic", str(n), str(m), str(r0), str(seed), "--rank", str(r), "--groups", str(group),
               "--max-iterations", str(iters), "--tol", "-1e30"], tmp_path)
    subs, rr = m // group, r // (m // group)
    assert res["iterations"] == subs * iters
    W = read_csv(tmp_path / "W.csv", n)
    H = read_csv(tmp_path / "Htrain.csv", r)
    Vm = G.as_matrix(G.gen_euclidean(n, m, r0, seed), n, m).astype(np.float64)
    for sm in range(subs):
        rows, cols = slice(sm * rr, (sm + 1) * rr), slice(sm * group, (sm + 1) * group)
        mask = np.ones_like(H, bool)
        mask[rows, cols] = False
        assert np.all(H[:, cols][mask[:, cols]] == 0.0)  # block-diagonal
        w = G.as_matrix(uniform(seed + sm, 4, n * rr), n, rr).astype(np.float64)
        h = G.as_matrix(uniform(seed + sm, 5, rr * group), rr, group).astype(np.float64)
        for _ in range(iters):
            w, h = orc.np_euc_iter(Vm[:, cols], w, h, G.REFERENCE_EPS)
        assert rel_fro(W[:, rows], w) < 1e-4 and rel_fro(H[rows, cols], h) < 1e-4


def test_quality_stopping(native, cuda, tmp_path):
    """The demo's stopping rule: stops once the quality improvement drops below tol."""
    res = run(["--synthetic", "361", "400", "10", "3", "--rank", "10", "--max-iterations", "5000", "--tol", "1e-3"],
              tmp_path)
    assert 3 <= res["iterations"] < 5000


@pytest.mark.parametrize("method,equalize", [("euclidean", False), ("kl", True)])
def test_cbcl_faces_train_and_project(native, orc, cuda, tmp_path, method, equalize):
    """The demo's own data: the 2429 CBCL faces (V 361 x 2429 as LoadImage builds it), rank
    49 (Definitions.h:29), 2000 training and 429 test images, against the FP64 oracle."""
    from test_faces_loader import cbcl_blobs, write_files

    blobs = cbcl_blobs()
    write_files(tmp_path / "img", blobs)
    r, iters, mtest, seed = 49, 60, 429, 0x5EED
    args = ["--pgm", str(tmp_path / "img"), "--images", str(len(blobs)), "--rank", str(r), "--method", method,
            "--max-iterations", str(iters), "--tol", "-1e30", "--test-images", str(mtest), "--seed", str(seed)]
    res = run(args + (["--equalize"] if equalize else []), tmp_path)
    assert res["n"] == 361 and res["m_train"] == 2000 and res["iterations"] == iters
    Vm = np.stack([orc.pgm_column(b, equalize) for b in blobs], axis=1).astype(np.float64)
    n, mt = 361, 2000
    step = orc.np_euc_iter if method == "euclidean" else orc.np_kl_iter
    w = G.as_matrix(uniform(seed, 4, n * r), n, r).astype(np.float64)
    h = G.as_matrix(uniform(seed, 5, r * mt), r, mt).astype(np.float64)
    obj = orc.np_frob if method == "euclidean" else orc.np_kl
    f0 = obj(Vm[:, :mt], w, h)
    for _ in range(iters):
        w, h = step(Vm[:, :mt], w, h, G.REFERENCE_EPS)
    assert obj(Vm[:, :mt], w, h) < 0.5 * f0  # it learns
    assert rel_fro(read_csv(tmp_path / "W.csv", n), w) < 1e-4
    assert rel_fro(read_csv(tmp_path / "Htrain.csv", r), h) < 1e-4
    wt = read_csv(tmp_path / "W.csv", n)
    ht = G.as_matrix(uniform(seed, 6, r * mtest), r, mtest).astype(np.float64)
    for _ in range(iters):
        _, ht = step(Vm[:, mt:], wt, ht, G.REFERENCE_EPS, update_w=False)
    assert rel_fro(read_csv(tmp_path / "Htest.csv", r), ht) < 1e-4
