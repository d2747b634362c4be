"""KL rule with BF16 storage: V packed to BF16 (V and V^T), the KL passes on BF16 MFMAs
(W H formed per block from the BF16 shadows, the FP32 ratio V / (W H + eps) rounded to BF16
for the second product, FP32 accumulation), FP32 updates over the clamped sums.

Gaps are stated as for the Euclidean BF16 path: against the FP64 restatement of the path's
own arithmetic (oracle.np_kl_iter_shadow: only the kernels' FP32 accumulation and reciprocal
separate the two), against the FP64 KL iteration (nmf_kernels.cu:63-341) on the same
BF16 V, and at 4096 x 4096 r = 32 (the KL benchmark configuration) against the FP32 path.
Poisson counts below 256 are exact in BF16, so the BF16 V of these problems is V itself."""
import numpy as np
import pytest

from conftest import rel_fro
from gpumlib_amd import nmf as G

pytestmark = pytest.mark.gpu

EPS = 1e-9


def _kl_bf16(V, W, H, n, m, r):
    return G.DeviceNMF(G.KL, V, W, H, n, m, r, eps=EPS, precision=G.BF16)


def _gaps(orc, n, m, r, iters, seed, r0=None, update_w=True):
    V = G.gen_poisson(n, m, r0 or min(r, 16), seed)
    W, H = G.gen_init(n, m, r, seed)
    p = _kl_bf16(V, W, H, n, m, r)
    p.iterate(iters, update_w=update_w)
    gobj = p.objective(G.OBJ_KL)
    v = G.as_matrix(V, n, m).astype(np.float64)
    assert np.array_equal(orc.bf16_round(v), v)  # counts below 256: the BF16 store is exact
    w0 = G.as_matrix(W, n, r).astype(np.float64)
    h0 = G.as_matrix(H, r, m).astype(np.float64)
    we, he, wq, hq = w0, h0, w0, h0
    for _ in range(iters):
        we, he = orc.np_kl_iter_shadow(v, we, he, EPS, orc.bf16_round, orc.bf16_round, update_w=update_w)
        wq, hq = orc.np_kl_iter(v, wq, hq, EPS, update_w=update_w)
    gw = G.as_matrix(p.get_w(), n, r).astype(np.float64)
    gh = G.as_matrix(p.get_h(), r, m).astype(np.float64)
    ref_e, ref_q = orc.np_kl(v, we, he), orc.np_kl(v, wq, hq)
    host = orc.np_kl(v, gw, gh)  # the GPU factors' objective, in FP64 on the host
    # a problem fitted exactly has objective ~ 0: relative figures are taken against at least
    # 1e-7 of sum(V), the objective pass's figure against 1e-4 of it (its W H is FP32-accurate,
    # and near a fit each KL term is a difference of two terms of the size of V)
    floor = 1e-7 * float(v.sum())
    out = dict(emu_w=rel_fro(gw, we), emu_h=rel_fro(gh, he), emu_obj=abs(host - ref_e) / max(ref_e, floor),
               same_v=abs(gobj - ref_q) / max(ref_q, floor), obj_pass=abs(gobj - host) / max(host, 1e3 * floor))
    print(f"kl bf16 {n}x{m} r={r} after {iters} iterations: " + "  ".join(f"{k} {x:.2e}" for k, x in out.items()),
          flush=True)
    return out, p, W


# ranks on one, two and four rank blocks of 32 (r = 65 and 128 bound the four-block kernel)
@pytest.mark.parametrize("shape", [(97, 61, 5), (130, 257, 33), (300, 200, 64), (1, 9, 1), (4100, 90, 32),
                                   (70, 4100, 20), (129, 300, 48), (600, 700, 64), (2000, 130, 17),
                                   (129, 300, 65), (300, 200, 100), (600, 500, 128)])
def test_kl_bf16_shapes(native, orc, cuda, shape):
    n, m, r = shape
    g, _, _ = _gaps(orc, n, m, r, 20, seed=n + 3 * m + r)
    assert g["emu_w"] < 1e-3 and g["emu_h"] < 1e-3
    assert g["emu_obj"] < 1e-4
    assert g["obj_pass"] < 1e-6  # the objective pass over the BF16 store, against the host


def test_kl_bf16_objective_gap(native, orc, cuda):
    """Against the FP64 KL iteration on the same V: the BF16 operands' effect, stated."""
    g, _, _ = _gaps(orc, 1536, 1024, 32, 100, seed=0x5EED + 3, r0=32)
    assert g["same_v"] < 1e-3
    assert g["emu_obj"] < 1e-4 and g["emu_w"] < 1e-3 and g["emu_h"] < 1e-3


def test_kl_bf16_update_h_only(native, orc, cuda):
    """DoIteration(false): W untouched, H as the restatement's H half."""
    g, p, W = _gaps(orc, 700, 500, 24, 5, seed=11, update_w=False)
    assert np.array_equal(p.get_w(), W)
    assert g["emu_h"] < 1e-4


def test_kl_bf16_deterministic_and_single_pass(native, cuda):
    """Bitwise the same run to run; the single-pass entry leaves W and H alone."""
    n, m, r = 900, 1100, 32
    V, (W, H) = G.gen_poisson(n, m, 16, 5), G.gen_init(n, m, r, 5)
    outs = []
    for _ in range(2):
        p = _kl_bf16(V, W, H, n, m, r)
        p.iterate(10)
        outs.append((p.get_w(), p.get_h()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    p.run_pass(0)
    p.run_pass(1)
    assert np.array_equal(p.get_w(), outs[1][0]) and np.array_equal(p.get_h(), outs[1][1])


def test_kl_bf16_c3_against_fp32_path(native, cuda):
    """4096 x 4096 r = 32 Poisson (rate W0 H0, W0, H0 ~ Gamma(1,1)), eps 1e-9, 300 iterations:
    the BF16 path's KL objective within 1e-3 of the FP32 path's (which the FP32 tests hold
    within 1e-4 of the FP64 oracle at this size), at 100 and 300 iterations."""
    n = m = 4096
    r = 32
    V = G.gen_poisson(n, m, r, 0x5EED + 3)
    W, H = G.gen_init(n, m, r, 0x5EED + 3)
    objs = {}
    for prec in (G.FP32, G.BF16):
        p = G.DeviceNMF(G.KL, V, W, H, n, m, r, eps=EPS, precision=prec)
        p.iterate(100)
        o100 = p.objective(G.OBJ_KL)
        p.iterate(200)
        objs[prec] = (o100, p.objective(G.OBJ_KL))
        del p
    gaps = [abs(b - f) / f for f, b in zip(objs[G.FP32], objs[G.BF16])]
    print(f"kl 4096x4096 r=32: FP32 {objs[G.FP32]}, BF16 {objs[G.BF16]}, gaps {gaps[0]:.2e} / {gaps[1]:.2e}",
          flush=True)
    assert gaps[0] < 1e-3 and gaps[1] < 1e-3
