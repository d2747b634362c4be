"""BF16 storage path: V packed to BF16 (V and V^T), BF16 MFMA passes with FP32
accumulation, FP32 updates.  Two objective gaps are checked: (i) against the FP64 oracle on the same BF16-quantised V (pure
arithmetic error of the BF16 operands) and (ii) against the FP64 oracle on the
FP32 V (storage error included)."""
import numpy as np
import pytest

from conftest import objective_allowance, rel_fro
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


def untile(raw, K, cols):
    """One orientation of the BF16 V store -> its padded matrix M[k, col]: blocks of 64
    columns x 256 contraction elements, block (k // 256, col // 64) at index
    k // 256 * ntiles + col // 64, column-major inside the block."""
    ldk, nt = (K + 255) // 256 * 256, (cols + 63) // 64
    blocks = raw[:ldk * 64 * nt].reshape(ldk // 256, nt, 64, 256)  # [chunk, tile, column, k in chunk]
    return blocks.transpose(0, 3, 1, 2).reshape(ldk, nt * 64)


@pytest.mark.parametrize("nm", [(64, 48), (97, 61), (1000, 333)])
def test_pack_bit_exact(native, cuda, nm):
    n, m = nm
    V = (np.random.default_rng(n).random(n * m, dtype=np.float32) * 37).astype(np.float32)
    W, H = G.gen_init(n, m, 4, 1)
    p = BF16Problem(V, W, H, n, m, 4)
    raw = p.V.download(np.uint16)
    ldv, ldt, tm, tn = (n + 255) // 256 * 256, (m + 255) // 256 * 256, (m + 63) // 64, (n + 63) // 64
    off_t = ((ldv * 64 * tm * 2 + 255) // 256 * 256) // 2
    assert raw.size * 2 == off_t * 2 + (ldt * 64 * tn * 2 + 255) // 256 * 256
    Vm = G.as_matrix(V, n, m)
    vb = untile(raw, n, m)           # vb[i, u] = V(i, u)
    vt = untile(raw[off_t:], m, n)   # vt[u, i] = V(i, u)
    assert np.array_equal(vb[:n, :m], bf16_bits(Vm))
    assert np.array_equal(vt[:m, :n], bf16_bits(Vm.T))
    assert not vb[n:].any() and not vb[:, m:].any() and not vt[m:].any() and not vt[:, n:].any()


def _gaps(n, m, r, iters, seed):
    """Run the BF16 path and return its errors against three FP64 references:
    same_v  the oracle on the BF16-rounded V (the format's effect on W and H operands, stated);
    fp32_v  the oracle on the FP32 V, objectives of both results taken against the FP32 V;
    emu_*   the restatement of the path's own arithmetic (BF16 V and BF16 shadows of W and the
            new H feeding the products, FP32 masters in the update) in FP64: only the FP32
            accumulation of the kernels separates the two."""
    from oracle import oracle as orc

    V, W, H = G.gen_euclidean(n, m, r, seed), *G.gen_init(n, m, r, seed)
    p = BF16Problem(V, W, H, n, m, r)
    p.iterate(iters)
    gobj = p.objective(G.OBJ_FROBENIUS)  # against the BF16 V it factorises
    vq = G.as_matrix(bf16_round(V), n, m).astype(np.float64)
    v = G.as_matrix(V, n, m).astype(np.float64)
    w0 = G.as_matrix(W, n, r).astype(np.float64)
    h0 = G.as_matrix(H, r, m).astype(np.float64)
    wq, hq = w0, h0
    w, h = w0, h0
    we, he = w0, h0
    for _ in range(iters):
        wq, hq = orc.np_euc_iter(vq, wq, hq, G.REFERENCE_EPS)
        w, h = orc.np_euc_iter(v, w, h, G.REFERENCE_EPS)
        we, he = orc.np_euc_iter_shadow(vq, we, he, G.REFERENCE_EPS, orc.bf16_round, orc.bf16_round)
    ref_q = orc.np_frob(vq, wq, hq)
    ref = orc.np_frob(v, w, h)
    ref_e = orc.np_frob(vq, we, he)
    gw, gh = G.as_matrix(p.get_w(), n, r), G.as_matrix(p.get_h(), r, m)
    gobj_fp32v = orc.np_frob(v, gw.astype(np.float64), gh.astype(np.float64))
    floor = 1e-9 * float(np.sum(vq * vq))  # an exactly fitted problem has objective ~ 0
    # emu_obj: objective difference beyond what the factors' distance and FP32 residuals explain
    d_wh = float(np.linalg.norm(gw.astype(np.float64) @ gh.astype(np.float64) - we @ he))
    excess = max(0.0, abs(gobj - ref_e) - objective_allowance(ref_e, float(np.sum(vq * vq)), r, d_wh))
    out = dict(same_v=abs(gobj - ref_q) / max(ref_q, floor), fp32_v=abs(gobj_fp32v - ref) / max(ref, floor),
               emu_obj=excess / max(ref_e, floor), emu_w=rel_fro(gw, we), emu_h=rel_fro(gh, he))
    print(f"bf16 {n}x{m} r={r} after {iters} iterations: " + "  ".join(f"{k} {v:.2e}" for k, v in out.items()))
    return out


def test_bf16_objective_gap(native, orc, cuda):
    g = _gaps(2048, 1024, 64, 100, seed=0x5EED + 5)
    assert g["same_v"] < 1e-3
    assert g["fp32_v"] < 1e-3
    # the kernels against the restatement of their own arithmetic: FP32 accumulation only
    assert g["emu_obj"] < 1e-4
    assert g["emu_w"] < 1e-3 and g["emu_h"] < 1e-3


# r = 65 is the first rank on 64-column tiles, r = 128 the largest rank of the path;
# (16500, 600, 48) and (300, 1100, 100) carry the gram spread over column tiles into the
# in-pass reduction of the slice grams (more than 8 slices, or r > 64)
@pytest.mark.parametrize("shape", [(97, 61, 5), (130, 257, 33), (300, 200, 100), (1, 9, 1), (16500, 90, 64),
                                   (70, 16390, 20), (129, 300, 65), (128, 384, 128), (16500, 600, 48),
                                   (300, 1100, 100)])
def test_bf16_shapes(native, orc, cuda, shape):
    n, m, r = shape
    g = _gaps(n, m, r, 20, seed=n + m)
    assert g["emu_obj"] < 1e-4
    assert g["emu_w"] < 1e-3 and g["emu_h"] < 1e-3


def test_bf16_deterministic(native, cuda):
    n, m, r = 900, 700, 32
    V, W, H = G.gen_euclidean(n, m, r, 3), *G.gen_init(n, m, r, 3)
    outs = []
    for _ in range(2):
        p = BF16Problem(V, W, H, n, m, r)
        p.iterate(10)
        outs.append((p.get_w(), p.get_h()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("rule,precision", [(G.EUCLIDEAN, G.BF16), (G.KL, G.BF16), (G.EUCLIDEAN, G.FP8)])
def test_bf16_stream_policy_bit_identical(native, cuda, rule, precision):
    """The BF16 passes read V with the non-temporal cache policy when the store is too large to stay
    cached (16384 x 8192 and up) and with the default policy below: both policies forced on one
    problem give the same bits (only the loads' cache policy differs); the FP8 passes take the
    forced policy too."""
    from gpumlib_amd import capi
    n, m, r = 900, 700, 32
    V = G.gen_euclidean(n, m, r, 7) if rule == G.EUCLIDEAN else G.gen_poisson(n, m, r, 7)
    W, H = G.gen_init(n, m, r, 7)
    outs = []
    try:
        for policy in (capi.STREAM_NONTEMPORAL, capi.STREAM_CACHED):
            capi.check(native.gpumlib_nmf_set_stream_policy(policy))
            p = G.DeviceNMF(rule, V, W, H, n, m, r, eps=G.REFERENCE_EPS, precision=precision)
            p.iterate(10)
            outs.append((p.get_w(), p.get_h()))
    finally:
        capi.check(native.gpumlib_nmf_set_stream_policy(capi.STREAM_AUTO))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    assert native.gpumlib_nmf_set_stream_policy(3) != 0


def test_bf16_partial_stream_policy_bit_identical(native, cuda):
    """At 16384 x 8192 (a 537 MB BF16 store) the default policy mixes the two cache policies by
    K slice (the H pass's first slices cached, the rest of V and all of V^T non-temporal): the
    same bits as the default policy everywhere."""
    from gpumlib_amd import capi
    n, m, r = 16384, 8192, 64
    V = G.gen_euclidean(n, m, r, 11)
    W, H = G.gen_init(n, m, r, 11)
    outs = []
    try:
        for policy in (capi.STREAM_AUTO, capi.STREAM_CACHED):
            capi.check(native.gpumlib_nmf_set_stream_policy(policy))
            p = BF16Problem(V, W, H, n, m, r)
            p.iterate(3)
            outs.append((p.get_w(), p.get_h()))
            del p
    finally:
        capi.check(native.gpumlib_nmf_set_stream_policy(capi.STREAM_AUTO))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
