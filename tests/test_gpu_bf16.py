"""BF16 storage path: V packed to BF16 (V and V^T), BF16 MFMA passes with FP32
accumulation, FP32 updates.  Two objective gaps are checked (the north-star
protocol): (i) against the FP64 oracle on the same BF16-quantised V (pure
arithmetic error of the BF16 operands) and (ii) against the FP64 oracle on the
FP32 V (storage error included)."""
import numpy as np
import pytest

from conftest import rel_fro
from gpumlib_amd import nmf as G

pytestmark = pytest.mark.gpu


def bf16_round(x):
    """float32 -> nearest-even bfloat16, returned as float32 (finite inputs)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_bits(x):
    return (bf16_round(x).view(np.uint32) >> 16).astype(np.uint16)


def BF16Problem(V, W, H, n, m, r, eps=G.REFERENCE_EPS):
    return G.DeviceNMF(G.EUCLIDEAN, V, W, H, n, m, r, eps=eps, precision=G.BF16)


@pytest.mark.parametrize("nm", [(64, 48), (97, 61), (1000, 333)])
def test_pack_bit_exact(native, cuda, nm):
    n, m = nm
    V = (np.random.default_rng(n).random(n * m, dtype=np.float32) * 37).astype(np.float32)
    W, H = G.gen_init(n, m, 4, 1)
    p = BF16Problem(V, W, H, n, m, 4)
    raw = p.V.download(np.uint16)
    ldv, ldt = (n + 7) // 8 * 8, (m + 7) // 8 * 8
    off_t = ((m * ldv * 2 + 255) // 256 * 256) // 2
    Vm = G.as_matrix(V, n, m)
    vb = raw[:m * ldv].reshape(m, ldv)[:, :n]          # column u of V
    vt = raw[off_t:off_t + n * ldt].reshape(n, ldt)[:, :m]  # column i of V^T
    assert np.array_equal(vb, bf16_bits(Vm.T))
    assert np.array_equal(vt, bf16_bits(Vm))


def _gaps(n, m, r, iters, seed):
    V, W, H = G.gen_euclidean(n, m, r, seed), *G.gen_init(n, m, r, seed)
    p = BF16Problem(V, W, H, n, m, r)
    p.iterate(iters)
    gobj = p.objective(G.OBJ_FROBENIUS)  # against the BF16 V it factorises
    vq = G.as_matrix(bf16_round(V), n, m).astype(np.float64)
    v = G.as_matrix(V, n, m).astype(np.float64)
    w0 = G.as_matrix(W, n, r).astype(np.float64)
    h0 = G.as_matrix(H, r, m).astype(np.float64)
    from oracle import oracle as orc

    wq, hq = w0, h0
    w, h = w0, h0
    for _ in range(iters):
        wq, hq = orc.np_euc_iter(vq, wq, hq, G.REFERENCE_EPS)
        w, h = orc.np_euc_iter(v, w, h, G.REFERENCE_EPS)
    ref_q = orc.np_frob(vq, wq, hq)
    ref = orc.np_frob(v, w, h)
    gw, gh = G.as_matrix(p.get_w(), n, r), G.as_matrix(p.get_h(), r, m)
    gobj_fp32v = orc.np_frob(v, gw.astype(np.float64), gh.astype(np.float64))
    # relative to the objective, floored at 1e-6 of |V|^2 (a problem fitted exactly has objective ~ 0)
    floor = 1e-6 * float(np.sum(vq * vq))
    return (abs(gobj - ref_q) / max(ref_q, floor), abs(gobj_fp32v - ref) / max(ref, floor), rel_fro(gw, wq),
            rel_fro(gh, hq))


def test_bf16_objective_gap(native, orc, cuda):
    gap_same_v, gap_fp32_v, ew, eh = _gaps(2048, 1024, 64, 100, seed=0x5EED + 5)
    print(f"bf16 gaps: same-V {gap_same_v:.2e}  fp32-V {gap_fp32_v:.2e}  W {ew:.2e}  H {eh:.2e}")
    assert gap_same_v < 1e-3
    assert gap_fp32_v < 1e-3
    assert ew < 3e-2 and eh < 3e-2


@pytest.mark.parametrize("shape", [(97, 61, 5), (130, 257, 33), (300, 200, 100), (1, 9, 1)])
def test_bf16_shapes(native, orc, cuda, shape):
    n, m, r = shape
    gap_same_v, _, _, _ = _gaps(n, m, r, 20, seed=n + m)
    assert gap_same_v < 1e-3


def test_bf16_deterministic(native, cuda):
    n, m, r = 900, 700, 32
    V, W, H = G.gen_euclidean(n, m, r, 3), *G.gen_init(n, m, r, 3)
    outs = []
    for _ in range(2):
        p = BF16Problem(V, W, H, n, m, r)
        p.iterate(10)
        outs.append((p.get_w(), p.get_h()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
