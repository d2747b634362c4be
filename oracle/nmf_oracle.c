/*
 * CPU oracle for the GPUMLib NMF multiplicative-update path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  It is used by tests/ (as the checker), by
 * __graft_entry__.smoke() (as the checker) and by bench.py's cpu_baseline leg
 * (as the timed CPU restatement of the reference rules).
 *
 * Every routine restates the reference algorithm from its own source; the
 * reference file and line each one follows is cited beside it (paths are
 * relative to the GPUMLib source tree).  Storage is column-major throughout,
 * as asserted by the reference (src/algorithms/nmf/nmf_base.h:109-111):
 *   V is n x m, V(i,u) = V[u*n + i]
 *   W is n x r, W(i,a) = W[a*n + i]
 *   H is r x m, H(a,u) = H[u*r + a]
 *
 * Two arithmetic flavours exist:
 *   *_f64  double precision: the expected values of every parity test;
 *   *_f32  single precision, reference-faithful operation order (FP32 is the
 *          reference's cudafloat, src/cuda/definitions.h:68);
 *   *_f32_omp  the same FP32 rules, cache-blocked and OpenMP-parallel: the
 *          host-cores CPU baseline.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Textbook GEMM with the reference container semantics.                     */
/* C (M x N, column-major, ldc) = alpha * op(A) * op(B) + beta * C           */
/* op(X) = X ('N') or X^T ('T').  Follows DeviceMatrix<float>::Multiply      */
/* (src/memory/DeviceMatrix.h:217-227: transposition is a flag, C is         */
/* column-major) and the triple loop of src/examples/FacesNMF/Matrix.h:48-58.*/
/* ------------------------------------------------------------------------ */
#define GEMM_BODY(T)                                                              \
    for (int j = 0; j < N; j++) {                                                 \
        for (int i = 0; i < M; i++) {                                             \
            T acc = 0;                                                            \
            for (int l = 0; l < K; l++) {                                         \
                T a = (ta == 'N') ? A[(size_t)l * lda + i] : A[(size_t)i * lda + l]; \
                T b = (tb == 'N') ? B[(size_t)j * ldb + l] : B[(size_t)l * ldb + j]; \
                acc += a * b;                                                     \
            }                                                                     \
            T *c = &C[(size_t)j * ldc + i];                                       \
            *c = (beta == 0) ? alpha * acc : alpha * acc + beta * *c;             \
        }                                                                         \
    }

void orc_gemm_f64(char ta, char tb, int M, int N, int K, double alpha, const double *A, int lda,
                  const double *B, int ldb, double beta, double *C, int ldc) {
    GEMM_BODY(double)
}

void orc_gemm_f32(char ta, char tb, int M, int N, int K, float alpha, const float *A, int lda,
                  const float *B, int ldb, float beta, float *C, int ldc) {
    GEMM_BODY(float)
}

/* ------------------------------------------------------------------------ */
/* Element-wise multiplicative update  m[k] *= nm[k] / (dm[k] + eps)          */
/* src/cuda/nmf/nmf_kernels.cu:43-47 (same rule: HostNMF.h:83,110).           */
/* ------------------------------------------------------------------------ */
void orc_mu_update_f32(const float *nm, const float *dm, float *mm, long elements, float eps) {
    for (long k = 0; k < elements; k++) mm[k] *= nm[k] / (dm[k] + eps);
}

void orc_mu_update_f64(const double *nm, const double *dm, double *mm, long elements, double eps) {
    for (long k = 0; k < elements; k++) mm[k] *= nm[k] / (dm[k] + eps);
}

/* ------------------------------------------------------------------------ */
/* Clamped column sums of W and row sums of H.                                */
/* SumW: src/cuda/nmf/nmf_kernels.cu:63-109 (clamp at :104).                  */
/* SumH: src/cuda/nmf/nmf_kernels.cu:149-195 (clamp at :190).                 */
/* ------------------------------------------------------------------------ */
void orc_sum_w_f64(int n, int r, const double *W, double eps, double *out) {
    for (int a = 0; a < r; a++) {
        double s = 0;
        for (int i = 0; i < n; i++) s += W[(size_t)a * n + i];
        out[a] = (s < eps) ? eps : s;
    }
}

void orc_sum_h_f64(int r, int m, const double *H, double eps, double *out) {
    for (int a = 0; a < r; a++) {
        double s = 0;
        for (int u = 0; u < m; u++) s += H[(size_t)u * r + a];
        out[a] = (s < eps) ? eps : s;
    }
}

void orc_sum_w_f32(int n, int r, const float *W, float eps, float *out) {
    for (int a = 0; a < r; a++) {
        float s = 0;
        for (int i = 0; i < n; i++) s += W[(size_t)a * n + i];
        out[a] = (s < eps) ? eps : s;
    }
}

void orc_sum_h_f32(int r, int m, const float *H, float eps, float *out) {
    for (int a = 0; a < r; a++) {
        float s = 0;
        for (int u = 0; u < m; u++) s += H[(size_t)u * r + a];
        out[a] = (s < eps) ? eps : s;
    }
}

/* ------------------------------------------------------------------------ */
/* KL kernels with a materialised WH, exactly the reference kernel rules:     */
/*  H(a,u) *= [sum_i W(i,a) * (V(i,u) / (WH(i,u) + eps))] / sumW[a]          */
/*      src/cuda/nmf/nmf_kernels.cu:295-341 (ratio :319,322; update :337-339) */
/*  W(i,a) *= [sum_u H(a,u) * (V(i,u) / (WH(i,u) + eps))] / sumH[a]          */
/*      src/cuda/nmf/nmf_kernels.cu:247-293 (ratio :271,274; update :289-291) */
/* ------------------------------------------------------------------------ */
#define KL_H_BODY(T)                                                          \
    for (int u = 0; u < m; u++) {                                             \
        for (int a = 0; a < r; a++) {                                         \
            T s = 0;                                                          \
            for (int i = 0; i < n; i++) {                                     \
                size_t k = (size_t)u * n + i;                                 \
                s += W[(size_t)a * n + i] * (V[k] / (WH[k] + eps));           \
            }                                                                 \
            H[(size_t)u * r + a] *= (s / sumW[a]);                            \
        }                                                                     \
    }
#define KL_W_BODY(T)                                                          \
    for (int a = 0; a < r; a++) {                                             \
        for (int i = 0; i < n; i++) {                                         \
            T s = 0;                                                          \
            for (int u = 0; u < m; u++) {                                     \
                size_t k = (size_t)u * n + i;                                 \
                s += H[(size_t)u * r + a] * (V[k] / (WH[k] + eps));           \
            }                                                                 \
            W[(size_t)a * n + i] *= (s / sumH[a]);                            \
        }                                                                     \
    }

void orc_kl_update_h_f64(double *H, const double *W, const double *V, const double *WH,
                         const double *sumW, int n, int m, int r, double eps) {
    KL_H_BODY(double)
}
void orc_kl_update_w_f64(double *W, const double *H, const double *V, const double *WH,
                         const double *sumH, int n, int m, int r, double eps) {
    KL_W_BODY(double)
}
void orc_kl_update_h_f32(float *H, const float *W, const float *V, const float *WH,
                         const float *sumW, int n, int m, int r, float eps) {
    KL_H_BODY(float)
}
void orc_kl_update_w_f32(float *W, const float *H, const float *V, const float *WH,
                         const float *sumH, int n, int m, int r, float eps) {
    KL_W_BODY(float)
}

/* ------------------------------------------------------------------------ */
/* One Euclidean multiplicative-update iteration.                             */
/* Operation order: src/algorithms/nmf/nmf_multiplicative_euclidian.cpp:29-67 */
/*   WtV = W^T V ; WtW = W^T W ; WtWH = WtW H ; H *= WtV / (WtWH + eps)       */
/*   if (!updateW) return                                                     */
/*   VHt = V H^T ; HHt = H H^T ; WHHt = W HHt ; W *= VHt / (WHHt + eps)       */
/* The H half uses the old W; the W half uses the new H.                      */
/* ------------------------------------------------------------------------ */
#define EUC_ITER_BODY(T, GEMM, UPD)                                              \
    T *WtV = (T *)malloc(sizeof(T) * (size_t)r * m);                             \
    T *WtW = (T *)malloc(sizeof(T) * (size_t)r * r);                             \
    T *WtWH = (T *)malloc(sizeof(T) * (size_t)r * m);                            \
    GEMM('T', 'N', r, m, n, 1, W, n, V, n, 0, WtV, r);                           \
    GEMM('T', 'N', r, r, n, 1, W, n, W, n, 0, WtW, r);                           \
    GEMM('N', 'N', r, m, r, 1, WtW, r, H, r, 0, WtWH, r);                        \
    UPD(WtV, WtWH, H, (long)r * m, eps);                                         \
    free(WtV);                                                                   \
    free(WtWH);                                                                  \
    if (updateW) {                                                               \
        T *VHt = (T *)malloc(sizeof(T) * (size_t)n * r);                         \
        T *WHHt = (T *)malloc(sizeof(T) * (size_t)n * r);                        \
        GEMM('N', 'T', n, r, m, 1, V, n, H, r, 0, VHt, n);                       \
        GEMM('N', 'T', r, r, m, 1, H, r, H, r, 0, WtW, r);                       \
        GEMM('N', 'N', n, r, r, 1, W, n, WtW, r, 0, WHHt, n);                    \
        UPD(VHt, WHHt, W, (long)n * r, eps);                                     \
        free(VHt);                                                               \
        free(WHHt);                                                              \
    }                                                                            \
    free(WtW);

void orc_euc_iter_f64(int n, int m, int r, const double *V, double *W, double *H, double eps,
                      int updateW) {
    EUC_ITER_BODY(double, orc_gemm_f64, orc_mu_update_f64)
}

void orc_euc_iter_f32(int n, int m, int r, const float *V, float *W, float *H, float eps,
                      int updateW) {
    EUC_ITER_BODY(float, orc_gemm_f32, orc_mu_update_f32)
}

/* ------------------------------------------------------------------------ */
/* One KL-divergence multiplicative-update iteration.                         */
/* Operation order: src/algorithms/nmf/nmf_multiplicative_divergence.cpp:29-50*/
/*   WH = W H ; sumW = clamp(colsum W) ; H update (kernel rules above)        */
/*   if (!updateW) return                                                     */
/*   WH = W H (new H) ; sumH = clamp(rowsum H) ; W update                     */
/* ------------------------------------------------------------------------ */
#define KL_ITER_BODY(T, GEMM, SW, SH, UH, UW)                                    \
    T *WH = (T *)malloc(sizeof(T) * (size_t)n * m);                              \
    T *sum = (T *)malloc(sizeof(T) * (size_t)r);                                 \
    GEMM('N', 'N', n, m, r, 1, W, n, H, r, 0, WH, n);                            \
    SW(n, r, W, eps, sum);                                                       \
    UH(H, W, V, WH, sum, n, m, r, eps);                                          \
    if (updateW) {                                                               \
        GEMM('N', 'N', n, m, r, 1, W, n, H, r, 0, WH, n);                        \
        SH(r, m, H, eps, sum);                                                   \
        UW(W, H, V, WH, sum, n, m, r, eps);                                      \
    }                                                                            \
    free(WH);                                                                    \
    free(sum);

void orc_kl_iter_f64(int n, int m, int r, const double *V, double *W, double *H, double eps,
                     int updateW) {
    KL_ITER_BODY(double, orc_gemm_f64, orc_sum_w_f64, orc_sum_h_f64, orc_kl_update_h_f64,
                 orc_kl_update_w_f64)
}

void orc_kl_iter_f32(int n, int m, int r, const float *V, float *W, float *H, float eps,
                     int updateW) {
    KL_ITER_BODY(float, orc_gemm_f32, orc_sum_w_f32, orc_sum_h_f32, orc_kl_update_h_f32,
                 orc_kl_update_w_f32)
}

/* ------------------------------------------------------------------------ */
/* Additive-update variants.                                                  */
/*                                                                            */
/* Element rule of NMF_AdditiveEuclidian (src/cuda/nmf/nmf_kernels.cu:31-39):  */
/*   X[k] = X[k] + (X[k] / d2[k]) * (d1[k] - d2[k]);  X[k] = 0 if negative     */
/* (no eps; a NaN stays NaN, as `v < 0` is false for it).                     */
/* ------------------------------------------------------------------------ */
#define ADD_UPDATE_BODY(T)                                                    \
    for (long k = 0; k < elements; k++) {                                     \
        T v = X[k] + (X[k] / d2[k]) * (d1[k] - d2[k]);                         \
        if (v < 0) v = 0;                                                     \
        X[k] = v;                                                             \
    }
void orc_add_update_f64(const double *d1, const double *d2, double *X, long elements) { ADD_UPDATE_BODY(double) }
void orc_add_update_f32(const float *d1, const float *d2, float *X, long elements) { ADD_UPDATE_BODY(float) }

/* Additive KL kernels over a materialised WH                                 */
/* (src/cuda/nmf/nmf_kernels_additive_divergence.cu):                         */
/*  H(a,u) = H + (H / sumW[a]) * sum_i W(i,a) * (V(i,u)/(WH(i,u)+eps) - 1)    */
/*      :126-193 (term :169, update and clamp :178-190)                       */
/*  W(i,a) = W + (W / sumH[a]) * sum_u H(a,u) * (V(i,u)/(WH(i,u)+eps) - 1)    */
/*      :57-124 (term :100, update and clamp :109-121)                        */
/* each clamped at 0.  sumW / sumH are the clamped sums of SumW / SumH.       */
#define ADD_KL_H_BODY(T)                                                      \
    for (int u = 0; u < m; u++) {                                             \
        for (int a = 0; a < r; a++) {                                         \
            T s = 0;                                                          \
            for (int i = 0; i < n; i++) {                                     \
                size_t k = (size_t)u * n + i;                                 \
                s += W[(size_t)a * n + i] * (V[k] / (WH[k] + eps) - 1);       \
            }                                                                 \
            T h = H[(size_t)u * r + a];                                       \
            T v = h + (h / sumW[a]) * s;                                      \
            if (v < 0) v = 0;                                                 \
            H[(size_t)u * r + a] = v;                                         \
        }                                                                     \
    }
#define ADD_KL_W_BODY(T)                                                      \
    for (int a = 0; a < r; a++) {                                             \
        for (int i = 0; i < n; i++) {                                         \
            T s = 0;                                                          \
            for (int u = 0; u < m; u++) {                                     \
                size_t k = (size_t)u * n + i;                                 \
                s += H[(size_t)u * r + a] * (V[k] / (WH[k] + eps) - 1);       \
            }                                                                 \
            T w = W[(size_t)a * n + i];                                       \
            T v = w + (w / sumH[a]) * s;                                      \
            if (v < 0) v = 0;                                                 \
            W[(size_t)a * n + i] = v;                                         \
        }                                                                     \
    }
void orc_add_kl_update_h_f64(double *H, const double *W, const double *V, const double *WH,
                             const double *sumW, int n, int m, int r, double eps) {
    ADD_KL_H_BODY(double)
}
void orc_add_kl_update_w_f64(double *W, const double *H, const double *V, const double *WH,
                             const double *sumH, int n, int m, int r, double eps) {
    ADD_KL_W_BODY(double)
}
void orc_add_kl_update_h_f32(float *H, const float *W, const float *V, const float *WH,
                             const float *sumW, int n, int m, int r, float eps) {
    ADD_KL_H_BODY(float)
}
void orc_add_kl_update_w_f32(float *W, const float *H, const float *V, const float *WH,
                             const float *sumH, int n, int m, int r, float eps) {
    ADD_KL_W_BODY(float)
}

/* One NMF_AdditiveEuclidian iteration (src/algorithms/nmf/nmf_additive_euclidian.cpp:31-56):   */
/*   H half (:35-42): deltaH = W^T V ; aux = W^T W ; deltaH2 = aux H ; then the reference calls */
/*     the multiplicative launcher with (nm = H, dm = deltaH, m = deltaH2) (:42), which         */
/*     rewrites the scratch deltaH2 and leaves H unchanged.  fixH != 0 applies the additive     */
/*     element rule to H instead: H = add(H, deltaH, deltaH2).                                  */
/*   if (!updateW) return (:44)                                                                 */
/*   W half (:47-55): deltaW = V H^T ; aux = H H^T ; deltaW2 = W aux ; W = add(W, deltaW, deltaW2) */
#define ADD_EUC_ITER_BODY(T, GEMM, UPD)                                          \
    T *aux = (T *)malloc(sizeof(T) * (size_t)r * r);                             \
    if (fixH) {                                                                  \
        T *dH = (T *)malloc(sizeof(T) * (size_t)r * m);                          \
        T *dH2 = (T *)malloc(sizeof(T) * (size_t)r * m);                         \
        GEMM('T', 'N', r, m, n, 1, W, n, V, n, 0, dH, r);                        \
        GEMM('T', 'N', r, r, n, 1, W, n, W, n, 0, aux, r);                       \
        GEMM('N', 'N', r, m, r, 1, aux, r, H, r, 0, dH2, r);                     \
        UPD(dH, dH2, H, (long)r * m);                                            \
        free(dH);                                                                \
        free(dH2);                                                               \
    }                                                                            \
    if (updateW) {                                                               \
        T *dW = (T *)malloc(sizeof(T) * (size_t)n * r);                          \
        T *dW2 = (T *)malloc(sizeof(T) * (size_t)n * r);                         \
        GEMM('N', 'T', n, r, m, 1, V, n, H, r, 0, dW, n);                        \
        GEMM('N', 'T', r, r, m, 1, H, r, H, r, 0, aux, r);                       \
        GEMM('N', 'N', n, r, r, 1, W, n, aux, r, 0, dW2, n);                     \
        UPD(dW, dW2, W, (long)n * r);                                            \
        free(dW);                                                                \
        free(dW2);                                                               \
    }                                                                            \
    free(aux);

void orc_add_euc_iter_f64(int n, int m, int r, const double *V, double *W, double *H, int updateW,
                          int fixH) {
    ADD_EUC_ITER_BODY(double, orc_gemm_f64, orc_add_update_f64)
}
void orc_add_euc_iter_f32(int n, int m, int r, const float *V, float *W, float *H, int updateW,
                          int fixH) {
    ADD_EUC_ITER_BODY(float, orc_gemm_f32, orc_add_update_f32)
}

/* One NMF_AdditiveDivergence iteration (src/algorithms/nmf/nmf_additive_divergence.cpp:30-51): */
/*   WH = W H ; sumW = clamp(colsum W) ; additive H update ; if (!updateW) return ;            */
/*   WH = W H (new H) ; sumH = clamp(rowsum H) ; additive W update                            */
void orc_add_kl_iter_f64(int n, int m, int r, const double *V, double *W, double *H, double eps,
                         int updateW) {
    KL_ITER_BODY(double, orc_gemm_f64, orc_sum_w_f64, orc_sum_h_f64, orc_add_kl_update_h_f64,
                 orc_add_kl_update_w_f64)
}
void orc_add_kl_iter_f32(int n, int m, int r, const float *V, float *W, float *H, float eps,
                         int updateW) {
    KL_ITER_BODY(float, orc_gemm_f32, orc_sum_w_f32, orc_sum_h_f32, orc_add_kl_update_h_f32,
                 orc_add_kl_update_w_f32)
}

/* ------------------------------------------------------------------------ */
/* Reductions of the Reduction class (src/cuda/reduction/reduction.h:41-157). */
/*  sum:      factor * sum_k x[k]  (Sum :60-71, Average :77-80), in double    */
/*  min/max:  start from FLT_MAX / -FLT_MAX (MAX_CUDAFLOAT / MIN_CUDAFLOAT,   */
/*            src/cuda/definitions.h:78, :91); an element replaces the        */
/*            candidate only when strictly smaller / larger (the comparisons */
/*            of reduction_kernels_min.cu:35-80 and _max.cu), scanned in index */
/*            order, so the first of equal extremes is kept and NaN never     */
/*            replaces; index -1 when nothing replaced the start value.       */
/* ------------------------------------------------------------------------ */
double orc_reduce_sum(const float *x, long count, double factor) {
    double s = 0.0;
    for (long k = 0; k < count; k++) s += (double)x[k];
    return s * factor;
}

void orc_reduce_extreme(const float *x, long count, int is_max, float *value, int *index) {
    float v = is_max ? -FLT_MAX : FLT_MAX;
    int idx = -1;
    for (long k = 0; k < count; k++) {
        const int better = is_max ? (x[k] > v) : (x[k] < v);
        if (better) {
            v = x[k];
            idx = (int)k;
        }
    }
    *value = v;
    *index = idx;
}

/* ------------------------------------------------------------------------ */
/* Objectives, accumulated in double.                                         */
/*  Frobenius  sum (V - WH)^2                                                 */
/*  KL         sum V ln(V / WH) - V + WH, with 0 ln 0 = 0                     */
/*  quality    sum V log10(WH + eps) - WH   (src/cuda/nmf/nmf_kernels_quality.cu:35) */
/* The first two have no reference counterpart; the third is the reference's  */
/* own quality measure over a materialised WH.                                */
/* ------------------------------------------------------------------------ */
static double wh_at(int n, int r, const double *W, const double *H, int i, int u) {
    double s = 0;
    for (int a = 0; a < r; a++) s += W[(size_t)a * n + i] * H[(size_t)u * r + a];
    return s;
}

double orc_frob_f64(int n, int m, int r, const double *V, const double *W, const double *H) {
    double acc = 0;
    for (int u = 0; u < m; u++)
        for (int i = 0; i < n; i++) {
            double d = V[(size_t)u * n + i] - wh_at(n, r, W, H, i, u);
            acc += d * d;
        }
    return acc;
}

double orc_kl_f64(int n, int m, int r, const double *V, const double *W, const double *H) {
    double acc = 0;
    for (int u = 0; u < m; u++)
        for (int i = 0; i < n; i++) {
            double v = V[(size_t)u * n + i], wh = wh_at(n, r, W, H, i, u);
            double t = (v > 0) ? v * log(v / wh) : 0.0;
            acc += t - v + wh;
        }
    return acc;
}

double orc_quality_f64(long count, const double *V, const double *WH, double eps) {
    double acc = 0;
    for (long k = 0; k < count; k++) acc += V[k] * log10(WH[k] + eps) - WH[k];
    return acc;
}

/* ------------------------------------------------------------------------ */
/* Host-cores FP32 baseline: the same FP32 rules as orc_euc_iter_f32 /        */
/* orc_kl_iter_f32 (operation order and per-element expressions identical),  */
/* with the contractions cache-blocked and OpenMP-parallel.  Summation order  */
/* inside each dot product differs from the naive loops (blocked, vectorised).*/
/* ------------------------------------------------------------------------ */
#define OMP_IB 1024

/* C(a,u) = sum_i X(i,a) Y(i,u): X is len x p (ldx), Y is len x q (ldy), both  */
/* contiguous along i; C is p x q with ldc.  Used for W^T V, W^T W.           */
static void omp_tn(int len, int p, int q, const float *X, int ldx, const float *Y, int ldy,
                   float *C, int ldc) {
#pragma omp parallel for schedule(static)
    for (int u = 0; u < q; u++) {
        for (int a = 0; a < p; a++) C[(size_t)u * ldc + a] = 0.f;
        for (int i0 = 0; i0 < len; i0 += OMP_IB) {
            int i1 = (i0 + OMP_IB < len) ? i0 + OMP_IB : len;
            const float *y = Y + (size_t)u * ldy;
            for (int a = 0; a < p; a++) {
                const float *x = X + (size_t)a * ldx;
                float s = 0.f;
#pragma omp simd reduction(+ : s)
                for (int i = i0; i < i1; i++) s += x[i] * y[i];
                C[(size_t)u * ldc + a] += s;
            }
        }
    }
}

/* C(i,a) = sum_u X(i,u) Y(a,u): X is len_i x K (ldx, contiguous along i),   */
/* Y is p x K (ldy, contiguous along a); C is len_i x p (ldc).  V H^T, H H^T.*/
static void omp_nt(int len_i, int p, int K, const float *X, int ldx, const float *Y, int ldy,
                   float *C, int ldc) {
#pragma omp parallel for schedule(static)
    for (int i0 = 0; i0 < len_i; i0 += 256) {
        int i1 = (i0 + 256 < len_i) ? i0 + 256 : len_i;
        for (int a = 0; a < p; a++)
            for (int i = i0; i < i1; i++) C[(size_t)a * ldc + i] = 0.f;
        for (int u = 0; u < K; u++) {
            const float *x = X + (size_t)u * ldx;
            const float *y = Y + (size_t)u * ldy;
            for (int a = 0; a < p; a++) {
                float ya = y[a];
                float *c = C + (size_t)a * ldc;
#pragma omp simd
                for (int i = i0; i < i1; i++) c[i] += x[i] * ya;
            }
        }
    }
}

/* C(i,u) = sum_a X(i,a) Y(a,u): n x r times r x m (W H, WtW H, W HHt).      */
static void omp_nn(int M, int N, int K, const float *X, int ldx, const float *Y, int ldy,
                   float *C, int ldc) {
#pragma omp parallel for schedule(static)
    for (int u = 0; u < N; u++) {
        float *c = C + (size_t)u * ldc;
        for (int i = 0; i < M; i++) c[i] = 0.f;
        for (int a = 0; a < K; a++) {
            float ya = Y[(size_t)u * ldy + a];
            const float *x = X + (size_t)a * ldx;
#pragma omp simd
            for (int i = 0; i < M; i++) c[i] += x[i] * ya;
        }
    }
}

static void omp_mu_update(const float *nm, const float *dm, float *mm, long elements, float eps) {
#pragma omp parallel for schedule(static)
    for (long k = 0; k < elements; k++) mm[k] *= nm[k] / (dm[k] + eps);
}

void orc_euc_iter_f32_omp(int n, int m, int r, const float *V, float *W, float *H, float eps,
                          int updateW) {
    float *WtV = (float *)malloc(sizeof(float) * (size_t)r * m);
    float *WtW = (float *)malloc(sizeof(float) * (size_t)r * r);
    float *WtWH = (float *)malloc(sizeof(float) * (size_t)r * m);
    omp_tn(n, r, m, W, n, V, n, WtV, r);
    omp_tn(n, r, r, W, n, W, n, WtW, r);
    omp_nn(r, m, r, WtW, r, H, r, WtWH, r);
    omp_mu_update(WtV, WtWH, H, (long)r * m, eps);
    if (updateW) {
        float *VHt = (float *)malloc(sizeof(float) * (size_t)n * r);
        float *WHHt = (float *)malloc(sizeof(float) * (size_t)n * r);
        omp_nt(n, r, m, V, n, H, r, VHt, n);
        omp_nt(r, r, m, H, r, H, r, WtW, r);
        omp_nn(n, r, r, W, n, WtW, r, WHHt, n);
        omp_mu_update(VHt, WHHt, W, (long)n * r, eps);
        free(VHt);
        free(WHHt);
    }
    free(WtV);
    free(WtW);
    free(WtWH);
}

void orc_kl_iter_f32_omp(int n, int m, int r, const float *V, float *W, float *H, float eps,
                         int updateW) {
    float *WH = (float *)malloc(sizeof(float) * (size_t)n * m);
    float *sum = (float *)malloc(sizeof(float) * (size_t)r);
    float *Q = (float *)malloc(sizeof(float) * (size_t)n * m);
    /* H half */
    omp_nn(n, m, r, W, n, H, r, WH, n);
    orc_sum_w_f32(n, r, W, eps, sum);
#pragma omp parallel for schedule(static)
    for (long k = 0; k < (long)n * m; k++) Q[k] = V[k] / (WH[k] + eps);
    {
        float *num = (float *)malloc(sizeof(float) * (size_t)r * m);
        omp_tn(n, r, m, W, n, Q, n, num, r);
#pragma omp parallel for schedule(static)
        for (int u = 0; u < m; u++)
            for (int a = 0; a < r; a++) H[(size_t)u * r + a] *= (num[(size_t)u * r + a] / sum[a]);
        free(num);
    }
    if (updateW) {
        omp_nn(n, m, r, W, n, H, r, WH, n);
        orc_sum_h_f32(r, m, H, eps, sum);
#pragma omp parallel for schedule(static)
        for (long k = 0; k < (long)n * m; k++) Q[k] = V[k] / (WH[k] + eps);
        float *num = (float *)malloc(sizeof(float) * (size_t)n * r);
        omp_nt(n, r, m, Q, n, H, r, num, n);
#pragma omp parallel for schedule(static)
        for (int a = 0; a < r; a++)
            for (int i = 0; i < n; i++) W[(size_t)a * n + i] *= (num[(size_t)a * n + i] / sum[a]);
        free(num);
    }
    free(WH);
    free(sum);
    free(Q);
}

/* FP64 KL iteration for full-size parity runs: the same rules as orc_kl_iter_f64
 * (nmf_multiplicative_divergence.cpp:29-50), with W H, the ratio and the
 * numerators fused per column of V and OpenMP over columns.  Summation inside
 * each dot product is sequential, as in the textbook loops. */
/* add != 0: the additive KL iteration (orc_add_kl_iter_f64) with the same loop structure:
 * q(i) = V / (WH + eps) - 1 and X = max(0, X + (X / sum) * num). */
static double kl_apply(double x, double num, double sum, int add) {
    if (!add) return x * (num / sum);
    double v = x + (x / sum) * num;
    return v < 0 ? 0 : v;
}

static void kl_iter_f64_omp_impl(int n, int m, int r, const double *V, double *W, double *H, double eps,
                                 int updateW, int add) {
    const double sub = add ? 1.0 : 0.0;
    double *sumW = (double *)malloc(sizeof(double) * (size_t)r);
    orc_sum_w_f64(n, r, W, eps, sumW);
    /* H half: per column u, q(i) = V(i,u) / ((W H)(i,u) + eps), H(a,u) *= sum_i W(i,a) q(i) / sumW[a] */
#pragma omp parallel
    {
        double *q = (double *)malloc(sizeof(double) * (size_t)n);
#pragma omp for schedule(static)
        for (int u = 0; u < m; u++) {
            const double *h = H + (size_t)u * r;
            for (int i = 0; i < n; i++) q[i] = 0.0;
            for (int a = 0; a < r; a++) {
                const double ha = h[a];
                const double *w = W + (size_t)a * n;
                for (int i = 0; i < n; i++) q[i] += w[i] * ha;
            }
            const double *v = V + (size_t)u * n;
            for (int i = 0; i < n; i++) q[i] = v[i] / (q[i] + eps) - sub;
            double *hn = H + (size_t)u * r;
            double num[512];
            for (int a = 0; a < r; a++) {
                const double *w = W + (size_t)a * n;
                double sacc = 0.0;
                for (int i = 0; i < n; i++) sacc += w[i] * q[i];
                num[a] = sacc;
            }
            for (int a = 0; a < r; a++) hn[a] = kl_apply(hn[a], num[a], sumW[a], add);
        }
        free(q);
    }
    free(sumW);
    if (!updateW) return;
    double *sumH = (double *)malloc(sizeof(double) * (size_t)r);
    orc_sum_h_f64(r, m, H, eps, sumH);
    /* W half: num(i,a) = sum_u H(a,u) V(i,u) / ((W H)(i,u) + eps), per-thread accumulators in u order */
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    double *acc = (double *)calloc((size_t)nt * n * r, sizeof(double));
    /* Each thread owns a contiguous range of columns u; the thread partials are then
     * summed in thread order, so the result depends only on the thread count. */
#pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double *mine = acc + (size_t)t * n * r;
        double *q = (double *)malloc(sizeof(double) * (size_t)n);
        const int per = (m + nt - 1) / nt;
        const int u0 = t * per, u1 = (u0 + per < m) ? u0 + per : m;
        for (int u = u0; u < u1; u++) {
            const double *h = H + (size_t)u * r;
            for (int i = 0; i < n; i++) q[i] = 0.0;
            for (int a = 0; a < r; a++) {
                const double ha = h[a];
                const double *w = W + (size_t)a * n;
                for (int i = 0; i < n; i++) q[i] += w[i] * ha;
            }
            const double *v = V + (size_t)u * n;
            for (int i = 0; i < n; i++) q[i] = v[i] / (q[i] + eps) - sub;
            for (int a = 0; a < r; a++) {
                const double ha = h[a];
                double *dst = mine + (size_t)a * n;
                for (int i = 0; i < n; i++) dst[i] += ha * q[i];
            }
        }
        free(q);
    }
    for (int a = 0; a < r; a++)
        for (int i = 0; i < n; i++) {
            double num = 0.0;
            for (int t = 0; t < nt; t++) num += acc[(size_t)t * n * r + (size_t)a * n + i];
            W[(size_t)a * n + i] = kl_apply(W[(size_t)a * n + i], num, sumH[a], add);
        }
    free(acc);
    free(sumH);
}

void orc_kl_iter_f64_omp(int n, int m, int r, const double *V, double *W, double *H, double eps, int updateW) {
    kl_iter_f64_omp_impl(n, m, r, V, W, H, eps, updateW, 0);
}

void orc_add_kl_iter_f64_omp(int n, int m, int r, const double *V, double *W, double *H, double eps, int updateW) {
    kl_iter_f64_omp_impl(n, m, r, V, W, H, eps, updateW, 1);
}

int orc_omp_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
