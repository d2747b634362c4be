"""FP32X storage path: V in FP32 (V and V^T tiled for the passes), the pass products on
the BF16 matrix cores from three-way BF16 splits of both operands, FP32 updates.

The path claims FP32 results, so it is held to the FP32 path's gate: after a fixed
number of iterations from the same seeded inputs, W, H and the objective match the
FP64 restatement of the reference update rules (nmf_multiplicative_euclidian.cpp:29-67,
nmf_kernels.cu) within 1e-4 relative Frobenius error; and its distance to the FP64
result must stay of the order of the FP32 path's own.
"""
import numpy as np
import pytest

from conftest import rel_fro
from gpumlib_amd import nmf as G

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def _oracle(orc, n, m, r, V, W, H, eps, iters, update_w=True):
    v, w, h = _f64(V), _f64(W), _f64(H)
    orc.euc_iter(n, m, r, v, w, h, eps, update_w=update_w, iters=iters)
    return v, w, h


def _problem(n, m, r, seed, r0=None):
    V = G.gen_euclidean(n, m, r0 or r, seed)
    W, H = G.gen_init(n, m, r, seed)
    return V, W, H


def untile(raw, K, cols):
    """One orientation of the FP32X V store -> its padded matrix M[k, col]: blocks of 64
    columns x 128 contraction elements, block (k // 128, col // 64) at index
    k // 128 * ntiles + col // 64, column-major inside the block."""
    ldk, nt = (K + 127) // 128 * 128, (cols + 63) // 64
    blocks = raw[:ldk * 64 * nt].reshape(ldk // 128, nt, 64, 128)  # [chunk, tile, column, k in chunk]
    return blocks.transpose(0, 3, 1, 2).reshape(ldk, nt * 64)


@pytest.mark.parametrize("nm", [(64, 48), (97, 61), (1000, 333)])
def test_pack_bit_exact(native, cuda, nm):
    n, m = nm
    V = (np.random.default_rng(n).random(n * m, dtype=np.float32) * 37).astype(np.float32)
    W, H = G.gen_init(n, m, 4, 1)
    p = G.DeviceNMF(G.EUCLIDEAN, V, W, H, n, m, 4, precision=G.FP32X)
    raw = p.V.download(np.float32)
    ldv, ldt, tm, tn = (n + 127) // 128 * 128, (m + 127) // 128 * 128, (m + 63) // 64, (n + 63) // 64
    off_t = ((ldv * 64 * tm * 4 + 255) // 256 * 256) // 4
    assert raw.size * 4 == off_t * 4 + (ldt * 64 * tn * 4 + 255) // 256 * 256
    Vm = G.as_matrix(V, n, m)
    vx = untile(raw, n, m)           # vx[i, u] = V(i, u)
    vt = untile(raw[off_t:], m, n)   # vt[u, i] = V(i, u)
    assert np.array_equal(vx[:n, :m], Vm)
    assert np.array_equal(vt[:m, :n], Vm.T)
    assert not vx[n:].any() and not vx[:, m:].any() and not vt[m:].any() and not vt[:, n:].any()


def test_c1_200_iterations(native, orc, cuda):
    """C1 (256 x 128, r = 16, 200 iterations) under the FP32 gate."""
    n, m, r, seed = 256, 128, 16, 0x5EED + 1
    V, W, H = _problem(n, m, r, seed)
    eps = G.REFERENCE_EPS
    d = G.DeviceNMF(G.EUCLIDEAN, V, W, H, n, m, r, eps=eps, precision=G.FP32X)
    d.iterate(200)
    v, w, h = _oracle(orc, n, m, r, V, W, H, eps, 200)
    assert rel_fro(d.get_w(), w) < TOL
    assert rel_fro(d.get_h(), h) < TOL
    ref = orc.frob(n, m, r, v, w, h)
    assert abs(d.objective(G.OBJ_FROBENIUS) - ref) / ref < TOL


SHAPES = [
    (97, 61, 5),      # nothing aligned: padding of chunks, tiles and rank rows
    (130, 257, 33),   # r just above one MFMA block: two rank blocks, four gram tiles
    (64, 64, 64),
    (300, 200, 100),  # four rank blocks, 64-column tiles, sixteen gram tiles
    (128, 384, 128),  # largest fused rank
    (1, 7, 1),
    (4100, 36, 12),   # tall: many K slices in the H pass
    (36, 4100, 12),   # wide: many K slices in the W pass
    (16500, 90, 64),  # n >= 16384: the 32-row update workgroups, a partial last one
    (16390, 70, 20),  # the same for one rank block, n not a multiple of 4
    (2048, 1536, 64),  # whole tiles and chunks, slice grams summed by the updates
]


@pytest.mark.parametrize("shape", SHAPES)
def test_shapes_vs_oracle(native, orc, cuda, shape):
    n, m, r = shape
    print(f"case fp32x shape={shape}", flush=True)
    V, W, H = _problem(n, m, r, seed=1234 + n + m + r, r0=min(r, 8))
    eps = G.REFERENCE_EPS
    d = G.DeviceNMF(G.EUCLIDEAN, V, W, H, n, m, r, eps=eps, precision=G.FP32X)
    d.iterate(20)
    _, w, h = _oracle(orc, n, m, r, V, W, H, eps, 20)
    assert rel_fro(d.get_w(), w) < TOL
    assert rel_fro(d.get_h(), h) < TOL


def test_update_h_only(native, orc, cuda):
    """DoIteration(false): H changes, W stays bit-identical."""
    n, m, r = 200, 150, 24
    V, W, H = _problem(n, m, r, seed=77)
    d = G.DeviceNMF(G.EUCLIDEAN, V, W, H, n, m, r, precision=G.FP32X)
    d.iterate(5, update_w=False)
    _, w, h = _oracle(orc, n, m, r, V, W, H, G.REFERENCE_EPS, 5, update_w=False)
    assert np.array_equal(d.get_w(), W)
    assert rel_fro(d.get_h(), h) < TOL


def test_objectives_match_oracle(native, orc, cuda):
    """FP64 objectives over the tiled FP32 store, standalone and on entry to an iteration."""
    n, m, r = 333, 190, 20
    V, W, H = _problem(n, m, r, seed=5)
    v, w, h = _f64(V), _f64(W), _f64(H)
    d = G.DeviceNMF(G.EUCLIDEAN, V, W, H, n, m, r, precision=G.FP32X)
    frob = orc.frob(n, m, r, v, w, h)
    kl = orc.kl_div(n, m, r, v, w, h)
    wh = np.zeros(n * m)
    orc.gemm("N", "N", n, m, r, 1.0, w, n, h, r, 0.0, wh, n)
    qual = orc.quality(v, wh, G.REFERENCE_EPS)
    assert abs(d.objective(G.OBJ_FROBENIUS) - frob) / frob < 1e-5
    assert abs(d.objective(G.OBJ_KL) - kl) / abs(kl) < 1e-5
    assert abs(d.objective(G.OBJ_QUALITY) - qual) / abs(qual) < 1e-5
    assert abs(d.iterate(1, objective=G.OBJ_FROBENIUS) - frob) / frob < 1e-5


def test_deterministic(native, cuda):
    """Fixed reduction order: two runs give identical bits."""
    n, m, r = 1000, 700, 32
    V, W, H = _problem(n, m, r, seed=9)
    outs = []
    for _ in range(2):
        d = G.DeviceNMF(G.EUCLIDEAN, V, W, H, n, m, r, precision=G.FP32X)
        d.iterate(10)
        outs.append((d.get_w(), d.get_h()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("nmr", [(1500, 1100, 64), (700, 900, 100)])
def test_error_of_the_fp32_path_order(native, orc, cuda, nmr):
    """Distance to the FP64 iteration after 30 iterations, FP32X against FP32 storage: the split
    products must not lose more than a small factor of the FP32 path's accuracy."""
    n, m, r = nmr
    V, W, H = _problem(n, m, r, seed=31, r0=12)
    eps = G.REFERENCE_EPS
    _, w, h = _oracle(orc, n, m, r, V, W, H, eps, 30)
    errs = {}
    for name, prec in (("fp32", G.FP32), ("fp32x", G.FP32X)):
        d = G.DeviceNMF(G.EUCLIDEAN, V, W, H, n, m, r, eps=eps, precision=prec)
        d.iterate(30)
        errs[name] = (rel_fro(d.get_w(), w), rel_fro(d.get_h(), h))
    print(f"{n}x{m} r={r}: fp32 W {errs['fp32'][0]:.2e} H {errs['fp32'][1]:.2e}  "
          f"fp32x W {errs['fp32x'][0]:.2e} H {errs['fp32x'][1]:.2e}", flush=True)
    for k in range(2):
        assert errs["fp32x"][k] < TOL
        assert errs["fp32x"][k] <= 4 * errs["fp32"][k] + 1e-6
