// C-ABI of the NMF library: argument checking, workspace planning and the
// launch sequence of one multiplicative-update iteration.
#include <atomic>
#include <cmath>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "../../include/gpumlib_nmf.h"
#include "nmf_common.hpp"
#include "nmf_internal.hpp"
#include "nmf_plan.hpp"

using namespace gpumlib_nmf;

namespace gpumlib_nmf {

#define CK(x)                               \
    do {                                    \
        const hipError_t ck_e = (x);        \
        if (ck_e != hipSuccess) return ck_e; \
    } while (0)

static int choose_slices(int tiles, int K, int kc, int target, int* kslice) {
    int S = (target + tiles - 1) / tiles;
    const int maxS = (K + kc - 1) / kc;
    if (S > maxS) S = maxS;
    if (S < 1) S = 1;
    int ks = round_up((K + S - 1) / S, kc);
    if (ks < kc) ks = kc;
    S = (K + ks - 1) / ks;
    if (S < 1) S = 1;
    *kslice = ks;
    return S;
}

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

size_t vstore_bf16_offset_t(int n, int m) { return align256(bf16_tiled_elems(n, m) * 2); }
size_t vstore_x3_offset_t(int n, int m) { return align256(x3_tiled_elems(n, m) * 4); }

Fp8Store fp8_store_layout(int n, int m) {
    Fp8Store l;
    l.off_t = align256((size_t)fp8_ld(n) * 64 * lp_tiles(m));
    l.off_sc = l.off_t + align256((size_t)fp8_ld(m) * 64 * lp_tiles(n));
    l.bytes = l.off_sc + align256((size_t)(fp8_ld(n) / 32) * (fp8_ld(m) / 32));
    return l;
}

size_t vstore_bytes(int precision, int n, int m) {
    if (n < 1 || m < 1) return 0;
    if (precision == GPUMLIB_NMF_FP32) return (size_t)n * m * 4;
    if (precision == GPUMLIB_NMF_BF16) return vstore_bf16_offset_t(n, m) + align256(bf16_tiled_elems(m, n) * 2);
    if (precision == GPUMLIB_NMF_FP8) return fp8_store_layout(n, m).bytes;
    if (precision == GPUMLIB_NMF_FP32X) return vstore_x3_offset_t(n, m) + align256(x3_tiled_elems(m, n) * 4);
    return 0;
}

// Ranks above kMaxFusedRank (FP32 storage): the unfused iteration's workspace.
static bool make_unfused_plan(int rule, int n, int m, int r, Plan* pl) {
    memset(pl, 0, sizeof(*pl));
    pl->rule = rule;
    pl->precision = GPUMLIB_NMF_FP32;
    pl->n = n;
    pl->m = m;
    pl->r = r;
    pl->unfused = 1;
    pl->rpad = round_up(r, 4);
    const size_t nm = (size_t)n * m, rm = (size_t)r * m, nr = (size_t)n * r;
    const size_t big = rm > nr ? rm : nr;
    size_t off = 0;
    pl->off_uWH = off;
    off += align256(sizeof(float) * nm);
    pl->off_uA = off;
    off += align256(sizeof(float) * big);
    pl->off_uB = off;
    off += align256(sizeof(float) * big);
    pl->off_uG = off;
    off += align256(sizeof(float) * (size_t)r * r);
    pl->off_usum = off;
    off += align256(sizeof(float) * (size_t)r);
    pl->off_uobj = off;
    off += align256(sizeof(double) * (size_t)objective_wh_partials((long)nm));
    pl->bytes = off;
    return true;
}

// Cache policy of the BF16 passes' V reads (gpumlib_nmf_set_stream_policy).
static std::atomic<int> g_stream_policy{GPUMLIB_NMF_STREAM_AUTO};

// From which contraction index the narrow passes read V with the non-temporal policy: k0[0] for
// the H pass (over the n rows of V), k0[1] for the W pass (over the m columns); n or m: none.
static void stream_policy_k0(int precision, int n, int m, int (&k0)[2]) {
    k0[0] = n;
    k0[1] = m;
    if (precision != GPUMLIB_NMF_BF16 && precision != GPUMLIB_NMF_FP8) return;
    const int policy = g_stream_policy.load(std::memory_order_relaxed);
    if (policy == GPUMLIB_NMF_STREAM_NONTEMPORAL) {
        k0[0] = k0[1] = 0;
    } else if (policy == GPUMLIB_NMF_STREAM_AUTO && precision == GPUMLIB_NMF_BF16 &&
               vstore_bytes(precision, n, m) > kStreamNtBytes) {
        // the first kStreamCacheBytes of V (whole 256-row chunks of the tiled store) stay cached
        const size_t row_bytes = (size_t)2 * 64 * lp_tiles(m);
        const size_t rows = kStreamCacheBytes / row_bytes / 256 * 256;
        k0[0] = rows < (size_t)n ? (int)rows : n;
        k0[1] = 0;
    }
}

bool make_plan(int rule, int precision, int n, int m, int r, Plan* pl, int target, int tile) {
    if (rule < GPUMLIB_NMF_EUCLIDEAN || rule > GPUMLIB_NMF_ADDITIVE_EUCLIDEAN_WITH_H) return false;
    if (precision != GPUMLIB_NMF_FP32 && precision != GPUMLIB_NMF_BF16 && precision != GPUMLIB_NMF_FP8 &&
        precision != GPUMLIB_NMF_FP32X)
        return false;
    // the narrow storage paths run the Euclidean rule; BF16 storage also runs the
    // multiplicative KL rule (up to r = 128, as the Euclidean rule)
    const bool kl_bf16 = precision == GPUMLIB_NMF_BF16 && rule == GPUMLIB_NMF_KL;
    if (precision != GPUMLIB_NMF_FP32 && rule != GPUMLIB_NMF_EUCLIDEAN && !kl_bf16) return false;
    if (n < 1 || m < 1 || r < 1) return false;
    if (tile != TILE && (tile != 64 || precision != GPUMLIB_NMF_FP32 || r > 64)) return false;
    if (r > kMaxFusedRank) {
        if (precision != GPUMLIB_NMF_FP32) return false;
        return make_unfused_plan(rule, n, m, r, pl);
    }
    if (precision == GPUMLIB_NMF_FP8 && r > 64) return false;
    memset(pl, 0, sizeof(*pl));
    pl->rule = rule;
    pl->precision = precision;
    pl->n = n;
    pl->m = m;
    pl->r = r;
    pl->npad = round_up(n, 4);
    pl->rpad = round_up(r, 4);
    pl->ldr = pl->rpad;
    pl->ldp = pl->npad + pl->rpad;
    pl->ldv8 = bf16_ld(n);
    pl->ldt8 = bf16_ld(m);
    pl->ld16n = fp8_ld(n);
    pl->ld16m = fp8_ld(m);
    pl->ldx_n = x3_ld(n);
    pl->ldx_m = x3_ld(m);
    pl->nbi = pl->ld16n / 32;
    pl->nbu = pl->ld16m / 32;
    const int euc = !rule_is_kl(rule);
    const int x3 = precision == GPUMLIB_NMF_FP32X;
    const int kc = (precision == GPUMLIB_NMF_FP32) ? KC
                   : (x3 || kl_bf16)               ? 128
                                                   : (precision == GPUMLIB_NMF_BF16 ? 256 : 512);
    // Euclidean grams: the FP32 passes spread the ceil(r/32)^2 gram blocks over the
    // first tiles of every slice, one per tile (tiny problems get gram-only tiles), and so
    // do the FP32X passes; the BF16 / FP8 passes put block t on column tile t when there
    // are enough column tiles (each block's k steps split over the waves), else every block
    // on tile 0.
    const int nbk = (r + 31) / 32;
    const int gram_tiles = ((precision == GPUMLIB_NMF_FP32 && euc) || x3) ? nbk * nbk : 0;
    // The BF16 passes take two 64-column blocks per workgroup of 8 waves for r <= 64: each A
    // chunk (the W or H^T shadow, re-read from L2 by every workgroup) then serves twice the
    // columns of V, halving the A loads; one such workgroup fills a CU (101 KB of LDS, the
    // same 8 waves as two 4-wave workgroups), so the pass aims at half the workgroups.  The
    // FP32X passes do the same (one 8-wave workgroup per CU for every rank, up to 153 KB of LDS).
    pl->lp_nt = (((precision == GPUMLIB_NMF_BF16 || x3) && r <= 64) || kl_bf16) ? 2 : 1;
    const int tw = (precision == GPUMLIB_NMF_FP32) ? tile : TILE_LP * pl->lp_nt;  // columns (rows) per workgroup
    const int tgt = x3 ? target / 2 : target / pl->lp_nt;
    const int vt1 = (m + tw - 1) / tw, vt3 = (n + tw - 1) / tw;
    pl->tiles1 = vt1 > gram_tiles ? vt1 : gram_tiles;
    pl->S1 = choose_slices(pl->tiles1, n, kc, tgt, &pl->ks1);
    pl->pstride1 = (long)(m + r) * pl->ldr;
    pl->tiles3 = vt3 > gram_tiles ? vt3 : gram_tiles;
    pl->S3 = choose_slices(pl->tiles3, m, kc, tgt, &pl->ks3);
    // FP32 W pass writes P[s][a*ldp + i]; the BF16 / FP8 / FP32X passes over V^T write P[s][i*ldr + a].
    pl->pstride3 = (precision == GPUMLIB_NMF_FP32) ? (long)r * pl->ldp : (long)(n + r) * pl->ldr;
    const size_t p1 = (size_t)pl->S1 * pl->pstride1, p3 = (size_t)pl->S3 * pl->pstride3;
    const size_t pbytes = align256(sizeof(float) * (p1 > p3 ? p1 : p3));
    const int smax = pl->S1 > pl->S3 ? pl->S1 : pl->S3;
    const size_t sbytes = align256(sizeof(float) * (size_t)smax * pl->rpad);
    // objective pass on split factors: 128-row blocks, each over csplit ranges of 32-column
    // blocks, aiming at kObjectiveWorkgroups work items of at least 8 column blocks (every
    // work item loads its W pieces once: fewer, longer column ranges read them fewer times)
    pl->ob_rp = split_rank(r);
    pl->ob_nrows = round_up(n, 128);
    pl->ob_ncols = round_up(m, 32);
    pl->ob_rblocks = pl->ob_nrows / 128;
    pl->ob_cblocks = pl->ob_ncols / 32;
    {
        const int want = (kObjectiveWorkgroups + pl->ob_rblocks - 1) / pl->ob_rblocks;
        const int most = pl->ob_cblocks / 8 > 1 ? pl->ob_cblocks / 8 : 1;
        pl->ob_csplit = want < 1 ? 1 : (want > most ? most : want);
    }
    // one objective partial per workgroup of the KL H pass or per work item of the objective pass
    pl->nobjp = pl->tiles1 * pl->S1 > pl->ob_rblocks * pl->ob_csplit ? pl->tiles1 * pl->S1
                                                                     : pl->ob_rblocks * pl->ob_csplit;
    const size_t obytes = align256(sizeof(double) * (size_t)pl->nobjp);
    size_t off = 0;
    pl->off_P = off;
    off += pbytes;
    pl->off_Psum = off;
    off += sbytes;
    pl->off_objp = off;
    off += obytes;
    pl->off_G = off;  // reduced grams of the two passes
    off += align256(sizeof(float) * 2 * (size_t)r * pl->rpad);
    pl->off_cnt = off;  // their arrival counters
    off += 256;
    pl->off_Ws = off;  // split pieces of W^T and of H
    off += align256((size_t)3 * pl->ob_nrows * pl->ob_rp * 2);
    pl->off_Hs = off;
    off += align256((size_t)3 * pl->ob_ncols * pl->ob_rp * 2);
    if (precision == GPUMLIB_NMF_BF16) {
        pl->off_Wb = off;
        off += align256((size_t)r * pl->ldv8 * 2);
        pl->off_Hb = off;
        off += align256((size_t)r * pl->ldt8 * 2);
        if (kl_bf16) {
            pl->rps = kl_rank_pitch(r);
            pl->off_Wr = off;
            off += align256((size_t)pl->ldv8 * pl->rps * 2);
            pl->off_Hr = off;
            off += align256((size_t)pl->ldt8 * pl->rps * 2);
        }
    } else if (x3) {  // three BF16 piece planes of W and of H^T
        pl->off_Wb = off;
        off += align256((size_t)3 * r * pl->ldx_n * 2);
        pl->off_Hb = off;
        off += align256((size_t)3 * r * pl->ldx_m * 2);
    } else if (precision == GPUMLIB_NMF_FP8) {
        pl->off_Wb = off;
        off += align256((size_t)r * pl->ld16n);
        pl->off_Hb = off;
        off += align256((size_t)r * pl->ld16m);
        pl->off_Wsc = off;
        off += align256((size_t)r * pl->nbi);
        pl->off_Hsc = off;
        off += align256((size_t)r * pl->nbu);
    }
    pl->bytes = off;
    stream_policy_k0(precision, n, m, pl->nt_k0);
    return true;
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

void pass_args(const Plan& pl, const float* V, const float* W, const float* H, void* ws, float eps, int which,
               PassArgs* a) {
    char* base = (char*)ws;
    memset(a, 0, sizeof(*a));
    a->V = V;
    a->W = W;
    a->H = H;
    a->P = (float*)(base + pl.off_P);
    a->Psum = (float*)(base + pl.off_Psum);
    a->objp = (double*)(base + pl.off_objp);
    a->n = pl.n;
    a->m = pl.m;
    a->r = pl.r;
    a->ldr = pl.ldr;
    a->ldp = pl.ldp;
    a->npad = pl.npad;
    a->eps = eps;
    a->vec = (pl.n % 4 == 0) && (pl.r % 4 == 0) && aligned16(V) && aligned16(W) && aligned16(H);
    a->ldv = (pl.precision == GPUMLIB_NMF_FP32) ? pl.n : (pl.precision == GPUMLIB_NMF_BF16 ? pl.ldv8 : pl.ld16n);
    a->vtiles = lp_tiles(pl.m);
    if (pl.precision == GPUMLIB_NMF_FP8) {
        a->vsc = (const uint8_t*)V + fp8_store_layout(pl.n, pl.m).off_sc;
        a->nbi = pl.nbi;
    }
    a->G = (float*)(base + pl.off_G) + (which == 1 ? 0 : (size_t)pl.r * pl.rpad);
    a->gcount = (int*)(base + pl.off_cnt) + (which == 1 ? 0 : 1);
    if (which == 1) {
        a->kslice = pl.ks1;
        a->S = pl.S1;
        a->tiles = pl.tiles1;
        a->pstride = pl.pstride1;
    } else {
        a->kslice = pl.ks3;
        a->S = pl.S3;
        a->tiles = pl.tiles3;
        a->pstride = pl.pstride3;
    }
}

// FP32 iterations of every rule.  The additive rules run the same passes with the
// additive update; GPUMLIB_NMF_ADDITIVE_EUCLIDEAN leaves H unchanged, as
// NMF_AdditiveEuclidian::DoIteration does (nmf_additive_euclidian.cpp:35-42 hands H to
// the multiplicative launcher as its numerator, so only its scratch deltaH2 changes),
// so its H half is not run at all.
hipError_t run_iterations(const Plan& pl, const float* V, float* W, float* H, float eps, int iterations, int update_w,
                          int objective, double* d_objective, void* ws, hipStream_t st) {
    const int kl = rule_is_kl(pl.rule);
    const int add = rule_is_additive(pl.rule);
    const int h_half = (pl.rule != GPUMLIB_NMF_ADDITIVE_EUCLIDEAN);
    PassArgs a1, a3;
    pass_args(pl, V, W, H, ws, eps, 1, &a1);
    pass_args(pl, V, W, H, ws, eps, 3, &a3);
    CK(launch_zero_i32(a1.gcount, 2, st));
    for (int it = 0; it < iterations; it++) {
        const int obj = (it == 0 && d_objective) ? objective : GPUMLIB_NMF_OBJ_NONE;
        if (h_half) {
            // The KL pass forms W H for its ratio, so the objective rides on it; the Euclidean
            // one would need W H only for the objective: that runs as its own pass.
            const int fobj = kl ? obj : GPUMLIB_NMF_OBJ_NONE;
            if (obj != GPUMLIB_NMF_OBJ_NONE && !kl) CK(run_objective(pl, V, W, H, eps, obj, d_objective, ws, st));
            CK(launch_pass_h_f32(a1, kl, fobj, st));
            if (fobj != GPUMLIB_NMF_OBJ_NONE) CK(launch_reduce_f64(a1.objp, a1.tiles * a1.S, d_objective, st));
            if (kl)
                CK(launch_update_h_kl(H, a1.P, a1.Psum, pl.m, pl.r, pl.ldr, pl.S1, pl.pstride1, eps, st, add));
            else
                CK(launch_update_h_euc(H, a1.P, a1.G, pl.m, pl.r, pl.ldr, pl.S1, pl.pstride1, eps, nullptr, 0, st,
                                       add));
        } else if (obj != GPUMLIB_NMF_OBJ_NONE) {
            CK(run_objective(pl, V, W, H, eps, obj, d_objective, ws, st));
        }
        if (!update_w) continue;
        CK(launch_pass_w_f32(a3, kl, st));
        if (kl)
            CK(launch_update_w_kl(W, a3.P, a3.Psum, pl.n, pl.r, pl.ldp, pl.S3, pl.pstride3, eps, st, add));
        else
            CK(launch_update_w_euc(W, a3.P, a3.G, pl.n, pl.r, pl.ldp, pl.S3, pl.pstride3, eps, st, add));
    }
    return hipSuccess;
}

// Objective of (W, H) against the plan's V store: W and H split into BF16 pieces, then the
// objective pass over V (W H on the BF16 matrix cores to FP32 accuracy), then the FP64
// reduction of its partials.
hipError_t run_objective(const Plan& pl, const void* V, const float* W, const float* H, float eps, int objective,
                         double* d_objective, void* ws, hipStream_t st) {
    char* base = (char*)ws;
    SplitObjArgs a;
    memset(&a, 0, sizeof(a));
    a.V = V;
    a.n = pl.n;
    a.m = pl.m;
    a.r = pl.r;
    a.rp = pl.ob_rp;
    a.nrows = pl.ob_nrows;
    a.ncols = pl.ob_ncols;
    a.rblocks = pl.ob_rblocks;
    a.cblocks = pl.ob_cblocks;
    a.csplit = pl.ob_csplit;
    a.eps = eps;
    a.Ws = (const uint16_t*)(base + pl.off_Ws);
    a.Hs = (const uint16_t*)(base + pl.off_Hs);
    a.objp = (double*)(base + pl.off_objp);
    a.ldv = pl.n;
    a.vec = (pl.n % 4 == 0) && aligned16(V);
    a.vtiles = lp_tiles(pl.m);
    if (pl.precision == GPUMLIB_NMF_FP8) {
        a.vsc = (const uint8_t*)V + fp8_store_layout(pl.n, pl.m).off_sc;
        a.nbi = pl.nbi;
    }
    const int vtype = pl.precision == GPUMLIB_NMF_FP32    ? 0
                      : pl.precision == GPUMLIB_NMF_BF16  ? 1
                      : pl.precision == GPUMLIB_NMF_FP32X ? 3
                                                          : 2;
    CK(launch_split_wh(W, H, pl.n, pl.m, pl.r, pl.ob_rp, pl.ob_nrows, pl.ob_ncols, (uint16_t*)(base + pl.off_Ws),
                       (uint16_t*)(base + pl.off_Hs), st));
    CK(launch_objective_split(a, vtype, objective, st));
    return launch_reduce_f64(a.objp, a.rblocks * a.csplit, d_objective, st);
}

// Ranks above kMaxFusedRank: the reference's own GEMM sequence over materialised operands
// (nmf_multiplicative_euclidian.cpp:29-67, nmf_multiplicative_divergence.cpp:29-50,
// nmf_additive_euclidian.cpp:31-56, nmf_additive_divergence.cpp:30-51), the GEMMs on the
// library's MFMA GEMM, the element-wise rules as in the fused path.
hipError_t run_objective_unfused(const Plan& pl, const float* V, const float* W, const float* H, float eps,
                                 int objective, double* d_objective, void* ws, hipStream_t st) {
    char* base = (char*)ws;
    float* WH = (float*)(base + pl.off_uWH);
    const int n = pl.n, m = pl.m, r = pl.r;
    CK(launch_sgemm('N', 'N', n, m, r, 1.f, W, n, H, r, 0.f, WH, n, st));
    return launch_objective_wh(V, WH, (long)n * m, eps, objective, (double*)(base + pl.off_uobj), d_objective, st);
}

hipError_t run_iterations_unfused(const Plan& pl, const float* V, float* W, float* H, float eps, int iterations,
                                  int update_w, int objective, double* d_objective, void* ws, hipStream_t st) {
    char* base = (char*)ws;
    float* WH = (float*)(base + pl.off_uWH);
    float* A = (float*)(base + pl.off_uA);
    float* B = (float*)(base + pl.off_uB);
    float* G = (float*)(base + pl.off_uG);
    float* sum = (float*)(base + pl.off_usum);
    const int n = pl.n, m = pl.m, r = pl.r;
    const long nm = (long)n * m, rm = (long)r * m, nr = (long)n * r;
    const int add = rule_is_additive(pl.rule);
    for (int it = 0; it < iterations; it++) {
        if (it == 0 && d_objective && objective != GPUMLIB_NMF_OBJ_NONE)
            CK(run_objective_unfused(pl, V, W, H, eps, objective, d_objective, ws, st));
        if (rule_is_kl(pl.rule)) {
            // H from the old W: WH = W H, sumW, Q = V / (WH + eps), H = rule(H, W^T Q, sumW)
            CK(launch_sgemm('N', 'N', n, m, r, 1.f, W, n, H, r, 0.f, WH, n, st));
            CK(launch_sum_w(W, n, r, eps, sum, st));
            CK(launch_kl_ratio(V, WH, nm, eps, add, st));
            CK(launch_sgemm('T', 'N', r, m, n, 1.f, W, n, WH, n, 0.f, A, r, st));
            CK(launch_kl_apply(H, A, sum, rm, r, 1, add, st));
            if (!update_w) continue;
            // W from the new H, WH recomputed
            CK(launch_sgemm('N', 'N', n, m, r, 1.f, W, n, H, r, 0.f, WH, n, st));
            CK(launch_sum_h(H, r, m, eps, sum, st));
            CK(launch_kl_ratio(V, WH, nm, eps, add, st));
            CK(launch_sgemm('N', 'T', n, r, m, 1.f, WH, n, H, r, 0.f, A, n, st));
            CK(launch_kl_apply(W, A, sum, nr, n, 0, add, st));
        } else {
            if (pl.rule != GPUMLIB_NMF_ADDITIVE_EUCLIDEAN) {  // that rule leaves H unchanged (fused path: same)
                CK(launch_sgemm('T', 'N', r, m, n, 1.f, W, n, V, n, 0.f, A, r, st));  // W^T V
                CK(launch_sgemm('T', 'N', r, r, n, 1.f, W, n, W, n, 0.f, G, r, st));  // W^T W
                CK(launch_sgemm('N', 'N', r, m, r, 1.f, G, r, H, r, 0.f, B, r, st));  // W^T W H
                if (add)
                    CK(launch_add_update(H, A, B, rm, st));
                else
                    CK(launch_mu_update(A, B, H, rm, eps, st));
            }
            if (!update_w) continue;
            CK(launch_sgemm('N', 'T', n, r, m, 1.f, V, n, H, r, 0.f, A, n, st));  // V H^T
            CK(launch_sgemm('N', 'T', r, r, m, 1.f, H, r, H, r, 0.f, G, r, st));  // H H^T
            CK(launch_sgemm('N', 'N', n, r, r, 1.f, W, n, G, r, 0.f, B, n, st));  // W H H^T
            if (add)
                CK(launch_add_update(W, A, B, nr, st));
            else
                CK(launch_mu_update(A, B, W, nr, eps, st));
        }
    }
    return hipSuccess;
}

// Slice grams the low-precision updates sum themselves (0: the pass reduces them into G
// at its tail).  Register-staged grams take up to kMaxGramSlices slices of r <= 64.
static int update_gram_slices(int S, int r) { return (S <= kMaxGramSlices && r <= 64) ? S : 0; }

// BF16 storage: V store = [V | V^T] in BF16, each tiled for its pass (bf16_tiled_index).  W and H stay FP32
// (the updates and their ratios are FP32); BF16 shadows feed the MFMA passes.
hipError_t run_iterations_bf16(const Plan& pl, const void* Vstore, float* W, float* H, float eps, int iterations,
                               int update_w, int objective, double* d_objective, void* ws, hipStream_t st) {
    char* base = (char*)ws;
    const uint16_t* Vb = (const uint16_t*)Vstore;
    const uint16_t* Vt = (const uint16_t*)((const char*)Vstore + vstore_bf16_offset_t(pl.n, pl.m));
    uint16_t* Wb = (uint16_t*)(base + pl.off_Wb);
    uint16_t* Hb = (uint16_t*)(base + pl.off_Hb);
    float* P = (float*)(base + pl.off_P);
    CK(launch_shadow_bf16(W, H, pl.n, pl.m, pl.r, Wb, pl.ldv8, Hb, pl.ldt8, st));
    float* G = (float*)(base + pl.off_G);
    int* cnt = (int*)(base + pl.off_cnt);
    CK(launch_zero_i32(cnt, 2, st));
    LpPassArgs ah, aw;
    memset(&ah, 0, sizeof(ah));
    ah.G = G;
    ah.gcount = cnt;
    ah.A = Wb;
    ah.B = Vb;
    ah.P = P;
    ah.K = pl.n;
    ah.N = pl.m;
    ah.r = pl.r;
    ah.lda = pl.ldv8;
    ah.ldr = pl.ldr;
    ah.kslice = pl.ks1;
    ah.S = pl.S1;
    ah.tiles = pl.tiles1;
    ah.nt = pl.lp_nt;
    ah.snt_k0 = pl.nt_k0[0];
    ah.pstride = pl.pstride1;
    aw = ah;
    aw.snt_k0 = pl.nt_k0[1];
    aw.G = G + (size_t)pl.r * pl.rpad;
    aw.gcount = cnt + 1;
    aw.A = Hb;
    aw.B = Vt;
    aw.K = pl.m;
    aw.N = pl.n;
    aw.lda = pl.ldt8;
    aw.kslice = pl.ks3;
    aw.S = pl.S3;
    aw.tiles = pl.tiles3;
    aw.pstride = pl.pstride3;
    const int gs1 = update_gram_slices(pl.S1, pl.r), gs3 = update_gram_slices(pl.S3, pl.r);
    ah.gkeep = gs1 > 0;
    aw.gkeep = gs3 > 0;
    const float* G1 = gs1 ? P + (size_t)pl.m * pl.ldr : ah.G;  // slice 0's gram columns, or the reduced gram
    const float* G3 = gs3 ? P + (size_t)pl.n * pl.ldr : aw.G;
    for (int it = 0; it < iterations; it++) {
        if (it == 0 && d_objective && objective != GPUMLIB_NMF_OBJ_NONE)
            CK(run_objective(pl, Vstore, W, H, eps, objective, d_objective, ws, st));
        CK(launch_pass_tn_bf16(ah, st));
        CK(launch_update_h_euc(H, P, G1, pl.m, pl.r, pl.ldr, pl.S1, pl.pstride1, eps, Hb, pl.ldt8, st, 0, gs1));
        if (!update_w) continue;
        CK(launch_pass_tn_bf16(aw, st));
        CK(launch_update_w_euc_t(W, P, G3, pl.n, pl.r, pl.ldr, pl.S3, pl.pstride3, eps, Wb, pl.ldv8, st, 0, gs3));
    }
    return hipSuccess;
}

// The KL passes over the plan's BF16 V store: H pass (which 0: A = W, C = H, B = V) or W pass
// (which 1: A = H^T, C = W^T, B = V^T), over the shadows in the workspace.
void kl_bf16_args(const Plan& pl, const void* Vstore, void* ws, float eps, int which, KlLpArgs* a) {
    char* base = (char*)ws;
    memset(a, 0, sizeof(*a));
    const uint16_t* Wk = (const uint16_t*)(base + pl.off_Wb);
    const uint16_t* Hk = (const uint16_t*)(base + pl.off_Hb);
    const uint16_t* Wr = (const uint16_t*)(base + pl.off_Wr);
    const uint16_t* Hr = (const uint16_t*)(base + pl.off_Hr);
    a->P = (float*)(base + pl.off_P);
    a->Psum = (float*)(base + pl.off_Psum);
    a->r = pl.r;
    a->rps = pl.rps;
    a->ldr = pl.ldr;
    a->rpad = pl.rpad;
    a->nt = pl.lp_nt;
    a->snt_k0 = pl.nt_k0[which];
    a->eps = eps;
    if (which == 0) {
        a->Ak = Wk, a->Ar = Wr, a->Cr = Hr, a->B = (const uint16_t*)Vstore;
        a->K = pl.n, a->N = pl.m, a->lda = pl.ldv8;
        a->kslice = pl.ks1, a->S = pl.S1, a->tiles = pl.tiles1, a->pstride = pl.pstride1;
    } else {
        a->Ak = Hk, a->Ar = Hr, a->Cr = Wr;
        a->B = (const uint16_t*)((const char*)Vstore + vstore_bf16_offset_t(pl.n, pl.m));
        a->K = pl.m, a->N = pl.n, a->lda = pl.ldt8;
        a->kslice = pl.ks3, a->S = pl.S3, a->tiles = pl.tiles3, a->pstride = pl.pstride3;
    }
}

// KL rule, BF16 storage (nmf_multiplicative_divergence.cpp:29-50): the shadows of W and H in
// both layouts are built once, then per iteration the H pass (W H of the old factors, formed
// in the pass) and the H update (which rewrites the H shadows), the W pass (W H with the new
// H) and the W update.
hipError_t run_iterations_kl_bf16(const Plan& pl, const void* Vstore, float* W, float* H, float eps, int iterations,
                                  int update_w, int objective, double* d_objective, void* ws, hipStream_t st) {
    if (!(eps > 0.f)) return hipErrorInvalidValue;  // padded rows rely on V / (0 + eps) = 0
    char* base = (char*)ws;
    uint16_t* Wk = (uint16_t*)(base + pl.off_Wb);
    uint16_t* Hk = (uint16_t*)(base + pl.off_Hb);
    uint16_t* Wr = (uint16_t*)(base + pl.off_Wr);
    uint16_t* Hr = (uint16_t*)(base + pl.off_Hr);
    const int n = pl.n, m = pl.m, r = pl.r;
    CK(launch_shadow_bf16(W, H, n, m, r, Wk, pl.ldv8, Hk, pl.ldt8, st));
    CK(launch_shadow_rank_bf16(W, 1, n, n, r, pl.ldv8, Wr, pl.rps, st));
    CK(launch_shadow_rank_bf16(H, r, 1, m, r, pl.ldt8, Hr, pl.rps, st));
    KlLpArgs ah, aw;
    kl_bf16_args(pl, Vstore, ws, eps, 0, &ah);
    kl_bf16_args(pl, Vstore, ws, eps, 1, &aw);
    for (int it = 0; it < iterations; it++) {
        if (it == 0 && d_objective && objective != GPUMLIB_NMF_OBJ_NONE)
            CK(run_objective(pl, Vstore, W, H, eps, objective, d_objective, ws, st));
        CK(launch_pass_kl_bf16(ah, st));
        CK(launch_update_kl_rows_bf16(H, r, 1, ah.P, ah.Psum, m, r, pl.ldr, pl.rpad, pl.S1, pl.pstride1, eps, Hk,
                                      pl.ldt8, Hr, pl.rps, st));
        if (!update_w) continue;
        CK(launch_pass_kl_bf16(aw, st));
        CK(launch_update_kl_rows_bf16(W, 1, n, aw.P, aw.Psum, n, r, pl.ldr, pl.rpad, pl.S3, pl.pstride3, eps, Wk,
                                      pl.ldv8, Wr, pl.rps, st));
    }
    return hipSuccess;
}

// The FP32X passes over the plan's V store: H pass (which 0: A = the W pieces, B = V) or W
// pass (which 1: A = the H^T pieces, B = V^T).
static void x3_args(const Plan& pl, const void* Vstore, void* ws, int which, LpPassArgs* a) {
    char* base = (char*)ws;
    memset(a, 0, sizeof(*a));
    a->P = (float*)(base + pl.off_P);
    a->r = pl.r;
    a->ldr = pl.ldr;
    a->nt = pl.lp_nt;
    a->G = (float*)(base + pl.off_G) + (which == 0 ? 0 : (size_t)pl.r * pl.rpad);
    a->gcount = (int*)(base + pl.off_cnt) + which;
    if (which == 0) {
        a->A = (const uint16_t*)(base + pl.off_Wb);
        a->B = (const uint16_t*)Vstore;
        a->K = pl.n, a->N = pl.m, a->lda = pl.ldx_n;
        a->kslice = pl.ks1, a->S = pl.S1, a->tiles = pl.tiles1, a->pstride = pl.pstride1;
    } else {
        a->A = (const uint16_t*)(base + pl.off_Hb);
        a->B = (const uint16_t*)((const char*)Vstore + vstore_x3_offset_t(pl.n, pl.m));
        a->K = pl.m, a->N = pl.n, a->lda = pl.ldx_m;
        a->kslice = pl.ks3, a->S = pl.S3, a->tiles = pl.tiles3, a->pstride = pl.pstride3;
    }
    a->aplane = (long)pl.r * a->lda;
}

// FP32X storage: V store = [V | V^T] in FP32, each tiled for its pass (x3_tiled_index).  W and H
// stay FP32; the passes take them as three BF16 piece planes, which the updates rewrite.
hipError_t run_iterations_x3(const Plan& pl, const void* Vstore, float* W, float* H, float eps, int iterations,
                             int update_w, int objective, double* d_objective, void* ws, hipStream_t st) {
    char* base = (char*)ws;
    uint16_t* Wp = (uint16_t*)(base + pl.off_Wb);
    uint16_t* Hp = (uint16_t*)(base + pl.off_Hb);
    float* P = (float*)(base + pl.off_P);
    CK(launch_shadow_x3(W, H, pl.n, pl.m, pl.r, Wp, pl.ldx_n, Hp, pl.ldx_m, st));
    CK(launch_zero_i32((int*)(base + pl.off_cnt), 2, st));
    LpPassArgs ah, aw;
    x3_args(pl, Vstore, ws, 0, &ah);
    x3_args(pl, Vstore, ws, 1, &aw);
    const int gs1 = update_gram_slices(pl.S1, pl.r), gs3 = update_gram_slices(pl.S3, pl.r);
    ah.gkeep = gs1 > 0;
    aw.gkeep = gs3 > 0;
    const float* G1 = gs1 ? P + (size_t)pl.m * pl.ldr : ah.G;  // slice 0's gram columns, or the reduced gram
    const float* G3 = gs3 ? P + (size_t)pl.n * pl.ldr : aw.G;
    for (int it = 0; it < iterations; it++) {
        if (it == 0 && d_objective && objective != GPUMLIB_NMF_OBJ_NONE)
            CK(run_objective(pl, Vstore, W, H, eps, objective, d_objective, ws, st));
        CK(launch_pass_tn_x3(ah, st));
        CK(launch_update_h_euc(H, P, G1, pl.m, pl.r, pl.ldr, pl.S1, pl.pstride1, eps, Hp, pl.ldx_m, st, 0, gs1,
                               aw.aplane));
        if (!update_w) continue;
        CK(launch_pass_tn_x3(aw, st));
        CK(launch_update_w_euc_t(W, P, G3, pl.n, pl.r, pl.ldr, pl.S3, pl.pstride3, eps, Wp, pl.ldx_n, st, 0, gs3,
                                 ah.aplane));
    }
    return hipSuccess;
}

static void fp8_args(const Plan& pl, const void* Vstore, void* ws, int which, Fp8PassArgs* a) {
    char* base = (char*)ws;
    const Fp8Store L = fp8_store_layout(pl.n, pl.m);
    memset(a, 0, sizeof(*a));
    a->snt_k0 = pl.nt_k0[which];
    a->P = (float*)(base + pl.off_P);
    a->r = pl.r;
    a->ldr = pl.ldr;
    a->Bsc = (const uint8_t*)Vstore + L.off_sc;
    a->bcol_shift = 5;
    a->G = (float*)(base + pl.off_G) + (which == 0 ? 0 : (size_t)pl.r * pl.rpad);
    a->gcount = (int*)(base + pl.off_cnt) + (which == 0 ? 0 : 1);
    if (which == 0) {  // H pass: A = W (blocks along i), B = V (columns u, K = i)
        a->A = (const uint8_t*)(base + pl.off_Wb);
        a->Asc = (const uint8_t*)(base + pl.off_Wsc);
        a->B = (const uint8_t*)Vstore;
        a->K = pl.n, a->N = pl.m, a->lda = pl.ld16n, a->nbk = pl.nbi;
        a->kslice = pl.ks1, a->S = pl.S1, a->tiles = pl.tiles1, a->pstride = pl.pstride1;
        a->bs_col = pl.nbi, a->bs_k = 1;  // tile scale Vsc[ub * nbi + ib], column u -> ub
    } else {  // W pass: A = H^T (blocks along u), B = V^T (columns i, K = u)
        a->A = (const uint8_t*)(base + pl.off_Hb);
        a->Asc = (const uint8_t*)(base + pl.off_Hsc);
        a->B = (const uint8_t*)Vstore + L.off_t;
        a->K = pl.m, a->N = pl.n, a->lda = pl.ld16m, a->nbk = pl.nbu;
        a->kslice = pl.ks3, a->S = pl.S3, a->tiles = pl.tiles3, a->pstride = pl.pstride3;
        a->bs_col = 1, a->bs_k = pl.nbi;  // column i -> ib, K block -> ub
    }
}

hipError_t run_iterations_fp8(const Plan& pl, const void* Vstore, float* W, float* H, float eps, int iterations,
                              int update_w, int objective, double* d_objective, void* ws, hipStream_t st) {
    char* base = (char*)ws;
    uint8_t* Wq = (uint8_t*)(base + pl.off_Wb);
    uint8_t* Hq = (uint8_t*)(base + pl.off_Hb);
    uint8_t* Wsc = (uint8_t*)(base + pl.off_Wsc);
    uint8_t* Hsc = (uint8_t*)(base + pl.off_Hsc);
    float* P = (float*)(base + pl.off_P);
    Fp8PassArgs ah, aw;
    fp8_args(pl, Vstore, ws, 0, &ah);
    fp8_args(pl, Vstore, ws, 1, &aw);
    CK(launch_shadow_fp8(W, pl.n, pl.m, pl.r, 0, Wq, pl.ld16n, Wsc, st));
    CK(launch_shadow_fp8(H, pl.n, pl.m, pl.r, 1, Hq, pl.ld16m, Hsc, st));
    CK(launch_zero_i32(ah.gcount, 2, st));
    const int gs1 = update_gram_slices(pl.S1, pl.r), gs3 = update_gram_slices(pl.S3, pl.r);
    ah.gkeep = gs1 > 0;
    aw.gkeep = gs3 > 0;
    const float* G1 = gs1 ? P + (size_t)pl.m * pl.ldr : ah.G;  // slice 0's gram columns, or the reduced gram
    const float* G3 = gs3 ? P + (size_t)pl.n * pl.ldr : aw.G;
    for (int it = 0; it < iterations; it++) {
        if (it == 0 && d_objective && objective != GPUMLIB_NMF_OBJ_NONE)
            CK(run_objective(pl, Vstore, W, H, eps, objective, d_objective, ws, st));
        CK(launch_pass_tn_fp8(ah, st));
        CK(launch_update_h_euc_q8(H, P, G1, pl.m, pl.r, pl.ldr, pl.S1, pl.pstride1, eps, Hq, pl.ld16m, Hsc, st,
                                  gs1));
        if (!update_w) continue;
        CK(launch_pass_tn_fp8(aw, st));
        CK(launch_update_w_euc_t_q8(W, P, G3, pl.n, pl.r, pl.ldr, pl.S3, pl.pstride3, eps, Wq, pl.ld16n, Wsc, st,
                                    gs3));
    }
    return hipSuccess;
}

}  // namespace gpumlib_nmf

// ---------------------------------------------------------------------------
extern "C" {

hipError_t gpumlib_cuda_nmf_multiplicative_euclidean_update_matrix(unsigned int, unsigned int, float* nm, float* dm,
                                                                   float* m, int elements) {
    return launch_mu_update(nm, dm, m, elements, kReferenceEpsilon, nullptr);
}

hipError_t gpumlib_cuda_nmf_multiplicative_divergence_update_matrix_w(dim3, dim3, float* W, float* H, float* V,
                                                                      float* WH, float* sumH, int n, int m, int r) {
    return launch_kl_update_w_wh(W, H, V, WH, sumH, n, m, r, kReferenceEpsilon, nullptr);
}

hipError_t gpumlib_cuda_nmf_multiplicative_divergence_update_matrix_h(dim3, dim3, float* H, float* W, float* V,
                                                                      float* WH, float* sumW, int n, int m, int r) {
    return launch_kl_update_h_wh(H, W, V, WH, sumW, n, m, r, kReferenceEpsilon, nullptr);
}

hipError_t gpumlib_cuda_nmf_additive_euclidean_update_matrix(unsigned int, unsigned int, float* X, float* deltaX1,
                                                             float* deltaX2, int elements) {
    return launch_add_update(X, deltaX1, deltaX2, elements, nullptr);
}

hipError_t gpumlib_cuda_nmf_additive_divergence_update_matrix_h(dim3, dim3, float* H, float* W, float* V, float* WH,
                                                                float* sumW, int n, int m, int r) {
    return launch_kl_update_h_wh(H, W, V, WH, sumW, n, m, r, kReferenceEpsilon, nullptr, 1);
}

hipError_t gpumlib_cuda_nmf_additive_divergence_update_matrix_w(dim3, dim3, float* W, float* H, float* V, float* WH,
                                                                float* sumH, int n, int m, int r) {
    return launch_kl_update_w_wh(W, H, V, WH, sumH, n, m, r, kReferenceEpsilon, nullptr, 1);
}

hipError_t gpumlib_nmf_additive_update(float* X, const float* d1, const float* d2, long elements,
                                       hipStream_t stream) {
    if (elements < 0) return hipErrorInvalidValue;
    return launch_add_update(X, d1, d2, elements, stream);
}

hipError_t gpumlib_nmf_additive_kl_update_h(float* H, const float* W, const float* V, const float* WH,
                                            const float* sumW, int n, int m, int r, float eps, hipStream_t stream) {
    if (n < 0 || m < 0 || r < 0) return hipErrorInvalidValue;
    return launch_kl_update_h_wh(H, W, V, WH, sumW, n, m, r, eps, stream, 1);
}

hipError_t gpumlib_nmf_additive_kl_update_w(float* W, const float* H, const float* V, const float* WH,
                                            const float* sumH, int n, int m, int r, float eps, hipStream_t stream) {
    if (n < 0 || m < 0 || r < 0) return hipErrorInvalidValue;
    return launch_kl_update_w_wh(W, H, V, WH, sumH, n, m, r, eps, stream, 1);
}

hipError_t gpumlib_nmf_mu_update(const float* nm, const float* dm, float* m, long elements, float eps,
                                 hipStream_t stream) {
    if (elements < 0) return hipErrorInvalidValue;
    return launch_mu_update(nm, dm, m, elements, eps, stream);
}

hipError_t gpumlib_nmf_sum_w(const float* W, int n, int r, float eps, float* sumW, hipStream_t stream) {
    if (n < 0 || r < 0) return hipErrorInvalidValue;
    return launch_sum_w(W, n, r, eps, sumW, stream);
}

hipError_t gpumlib_nmf_sum_h(const float* H, int r, int m, float eps, float* sumH, hipStream_t stream) {
    if (m < 0 || r < 0) return hipErrorInvalidValue;
    return launch_sum_h(H, r, m, eps, sumH, stream);
}

hipError_t gpumlib_nmf_kl_update_h(float* H, const float* W, const float* V, const float* WH, const float* sumW, int n,
                                   int m, int r, float eps, hipStream_t stream) {
    if (n < 0 || m < 0 || r < 0) return hipErrorInvalidValue;
    return launch_kl_update_h_wh(H, W, V, WH, sumW, n, m, r, eps, stream);
}

hipError_t gpumlib_nmf_kl_update_w(float* W, const float* H, const float* V, const float* WH, const float* sumH, int n,
                                   int m, int r, float eps, hipStream_t stream) {
    if (n < 0 || m < 0 || r < 0) return hipErrorInvalidValue;
    return launch_kl_update_w_wh(W, H, V, WH, sumH, n, m, r, eps, stream);
}

static int quality_partials(long count) {
    long p = (count + (long)NTHREADS * 16 - 1) / ((long)NTHREADS * 16);
    if (p < 1) p = 1;
    if (p > 1024) p = 1024;
    return (int)p;
}

size_t gpumlib_nmf_quality_workspace_bytes(long count) {
    return sizeof(double) * (size_t)quality_partials(count < 0 ? 0 : count);
}

hipError_t gpumlib_nmf_quality(const float* V, const float* WH, long count, float eps, void* workspace,
                               size_t workspace_bytes, double* d_quality, hipStream_t stream) {
    if (count < 0 || !d_quality) return hipErrorInvalidValue;
    const int np = quality_partials(count);
    if (!workspace || workspace_bytes < sizeof(double) * (size_t)np) return hipErrorInvalidValue;
    return launch_quality(V, WH, count, eps, (double*)workspace, np, d_quality, stream);
}

hipError_t gpumlib_nmf_quality_f32(const float* V, const float* WH, long count, float eps, void* workspace,
                                   size_t workspace_bytes, float* d_quality, hipStream_t stream) {
    if (count < 0 || !d_quality) return hipErrorInvalidValue;
    const int np = quality_partials(count);
    if (!workspace || workspace_bytes < sizeof(double) * (size_t)np) return hipErrorInvalidValue;
    return launch_quality(V, WH, count, eps, (double*)workspace, np, nullptr, stream, d_quality);
}

hipError_t gpumlib_nmf_sgemm(char transa, char transb, int M, int N, int K, float alpha, const float* A, int lda,
                             const float* B, int ldb, float beta, float* C, int ldc, hipStream_t stream) {
    return launch_sgemm(transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, stream);
}

hipError_t gpumlib_nmf_gemm_bf16(char transa, char transb, int M, int N, int K, float alpha, const uint16_t* A,
                                 int lda, const uint16_t* B, int ldb, float beta, float* C, int ldc, hipStream_t stream) {
    return launch_gemm(1, transa, transb, M, N, K, alpha, A, lda, 127, B, ldb, 127, beta, C, ldc, stream);
}

hipError_t gpumlib_nmf_gemm_fp8(char transa, char transb, int M, int N, int K, float alpha, const uint8_t* A, int lda,
                                int scale_a, const uint8_t* B, int ldb, int scale_b, float beta, float* C, int ldc,
                                hipStream_t stream) {
    if (scale_a < 0 || scale_a > 254 || scale_b < 0 || scale_b > 254) return hipErrorInvalidValue;
    return launch_gemm(2, transa, transb, M, N, K, alpha, A, lda, scale_a, B, ldb, scale_b, beta, C, ldc, stream);
}

hipError_t gpumlib_nmf_set_stream_policy(int policy) {
    if (policy != GPUMLIB_NMF_STREAM_AUTO && policy != GPUMLIB_NMF_STREAM_NONTEMPORAL && policy != GPUMLIB_NMF_STREAM_CACHED)
        return hipErrorInvalidValue;
    g_stream_policy.store(policy, std::memory_order_relaxed);
    return hipSuccess;
}

size_t gpumlib_nmf_workspace_bytes(int rule, int precision, int n, int m, int r) {
    Plan pl;
    if (!make_plan(rule, precision, n, m, r, &pl)) return 0;
    return pl.bytes;
}

hipError_t gpumlib_nmf_iterate(int rule, int precision, int n, int m, int r, const void* V, float* W, float* H,
                               float eps, int iterations, int update_w, int objective, double* d_objective,
                               void* workspace, size_t workspace_bytes, hipStream_t stream) {
    Plan pl;
    if (!make_plan(rule, precision, n, m, r, &pl)) return hipErrorInvalidValue;
    if (!V || !W || !H || !workspace || workspace_bytes < pl.bytes || iterations < 0) return hipErrorInvalidValue;
    if (objective < GPUMLIB_NMF_OBJ_NONE || objective > GPUMLIB_NMF_OBJ_QUALITY) return hipErrorInvalidValue;
    if (objective != GPUMLIB_NMF_OBJ_NONE && !d_objective) return hipErrorInvalidValue;
    if (iterations == 0) {
        if (objective == GPUMLIB_NMF_OBJ_NONE) return hipSuccess;
        if (pl.unfused)
            return run_objective_unfused(pl, (const float*)V, W, H, eps, objective, d_objective, workspace, stream);
        return run_objective(pl, V, W, H, eps, objective, d_objective, workspace, stream);
    }
    if (pl.unfused)
        return run_iterations_unfused(pl, (const float*)V, W, H, eps, iterations, update_w, objective, d_objective,
                                      workspace, stream);
    if (precision == GPUMLIB_NMF_BF16 && rule_is_kl(rule))
        return run_iterations_kl_bf16(pl, V, W, H, eps, iterations, update_w, objective, d_objective, workspace,
                                      stream);
    if (precision == GPUMLIB_NMF_BF16)
        return run_iterations_bf16(pl, V, W, H, eps, iterations, update_w, objective, d_objective, workspace, stream);
    if (precision == GPUMLIB_NMF_FP8)
        return run_iterations_fp8(pl, V, W, H, eps, iterations, update_w, objective, d_objective, workspace, stream);
    if (precision == GPUMLIB_NMF_FP32X)
        return run_iterations_x3(pl, V, W, H, eps, iterations, update_w, objective, d_objective, workspace, stream);
    return run_iterations(pl, (const float*)V, W, H, eps, iterations, update_w, objective, d_objective, workspace,
                          stream);
}

hipError_t gpumlib_nmf_objective(int rule, int precision, int n, int m, int r, const void* V, const float* W,
                                 const float* H, float eps, int objective, double* d_objective, void* workspace,
                                 size_t workspace_bytes, hipStream_t stream) {
    Plan pl;
    if (!make_plan(rule, precision, n, m, r, &pl)) return hipErrorInvalidValue;
    if (!V || !W || !H || !workspace || workspace_bytes < pl.bytes || !d_objective) return hipErrorInvalidValue;
    if (objective < GPUMLIB_NMF_OBJ_FROBENIUS || objective > GPUMLIB_NMF_OBJ_QUALITY) return hipErrorInvalidValue;
    if (pl.unfused)
        return run_objective_unfused(pl, (const float*)V, W, H, eps, objective, d_objective, workspace, stream);
    return run_objective(pl, V, W, H, eps, objective, d_objective, workspace, stream);
}

size_t gpumlib_nmf_vstore_bytes(int precision, int n, int m) { return vstore_bytes(precision, n, m); }

hipError_t gpumlib_nmf_vstore_pack(int precision, int n, int m, const float* V, void* vstore, size_t vstore_bytes_,
                                   hipStream_t stream) {
    const size_t need = vstore_bytes(precision, n, m);
    if (!V || !vstore || need == 0 || vstore_bytes_ < need) return hipErrorInvalidValue;
    if (precision == GPUMLIB_NMF_FP32)
        return hipMemcpyAsync(vstore, V, need, hipMemcpyDeviceToDevice, stream);
    if (precision == GPUMLIB_NMF_FP8) {
        const Fp8Store L = fp8_store_layout(n, m);
        uint8_t* base = (uint8_t*)vstore;
        return launch_pack_v_fp8(V, n, m, base, fp8_ld(n), base + L.off_t, fp8_ld(m), base + L.off_sc, stream);
    }
    if (precision == GPUMLIB_NMF_FP32X)
        return launch_pack_v_x3(V, n, m, (float*)vstore, x3_ld(n), (float*)((char*)vstore + vstore_x3_offset_t(n, m)),
                                x3_ld(m), stream);
    uint16_t* Vb = (uint16_t*)vstore;
    uint16_t* Vt = (uint16_t*)((char*)vstore + vstore_bf16_offset_t(n, m));
    return launch_pack_v_bf16(V, n, m, Vb, bf16_ld(n), Vt, bf16_ld(m), stream);
}

hipError_t gpumlib_nmf_pass(int rule, int precision, int which, int n, int m, int r, const void* V, const float* W,
                            const float* H, float eps, void* workspace, size_t workspace_bytes, hipStream_t stream) {
    Plan pl;
    if (!make_plan(rule, precision, n, m, r, &pl)) return hipErrorInvalidValue;
    if (!V || !W || !H || !workspace || workspace_bytes < pl.bytes || (which != 0 && which != 1) || pl.unfused)
        return hipErrorInvalidValue;
    const int kl = rule_is_kl(rule);
    // The arrival counters were zeroed by gpumlib_nmf_iterate and every pass leaves them at
    // zero, so the pass runs alone: its launch is the one being timed or profiled, configured
    // as in the iteration (the narrow passes leave the slice grams to the update when the
    // iteration does).
    const int gslices = update_gram_slices(which == 0 ? pl.S1 : pl.S3, r);
    if (precision == GPUMLIB_NMF_FP8) {
        Fp8PassArgs a;
        fp8_args(pl, V, workspace, which, &a);
        a.gkeep = gslices > 0;
        return launch_pass_tn_fp8(a, stream);
    }
    if (precision == GPUMLIB_NMF_FP32X) {
        // one FP32X pass over the W/H pieces the last iterate call left in this workspace
        LpPassArgs a;
        x3_args(pl, V, workspace, which, &a);
        a.gkeep = gslices > 0;
        return launch_pass_tn_x3(a, stream);
    }
    if (precision == GPUMLIB_NMF_BF16 && kl) {
        // one KL pass over the W/H shadows the last iterate call left in this workspace
        KlLpArgs a;
        kl_bf16_args(pl, V, workspace, eps, which, &a);
        return launch_pass_kl_bf16(a, stream);
    }
    if (precision == GPUMLIB_NMF_BF16) {
        // one BF16 pass over the W/H shadows the last iterate call left in this workspace
        char* base = (char*)workspace;
        uint16_t* Wb = (uint16_t*)(base + pl.off_Wb);
        uint16_t* Hb = (uint16_t*)(base + pl.off_Hb);
        LpPassArgs a;
        memset(&a, 0, sizeof(a));
        a.P = (float*)(base + pl.off_P);
        a.r = r;
        a.ldr = pl.ldr;
        a.G = (float*)(base + pl.off_G) + (which == 0 ? 0 : (size_t)r * pl.rpad);
        a.gcount = (int*)(base + pl.off_cnt) + which;
        a.nt = pl.lp_nt;
        a.snt_k0 = pl.nt_k0[which];
        a.gkeep = gslices > 0;
        if (which == 0) {
            a.A = Wb, a.B = (const uint16_t*)V, a.K = n, a.N = m, a.lda = pl.ldv8;
            a.kslice = pl.ks1, a.S = pl.S1, a.tiles = pl.tiles1, a.pstride = pl.pstride1;
        } else {
            a.A = Hb, a.B = (const uint16_t*)((const char*)V + vstore_bf16_offset_t(n, m)), a.K = m, a.N = n;
            a.lda = pl.ldt8;
            a.kslice = pl.ks3, a.S = pl.S3, a.tiles = pl.tiles3, a.pstride = pl.pstride3;
        }
        return launch_pass_tn_bf16(a, stream);
    }
    PassArgs a;
    pass_args(pl, (const float*)V, W, H, workspace, eps, which == 0 ? 1 : 3, &a);
    if (which == 0) return launch_pass_h_f32(a, kl, GPUMLIB_NMF_OBJ_NONE, stream);
    return launch_pass_w_f32(a, kl, stream);
}

}  // extern "C"
