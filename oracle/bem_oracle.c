/*
 * CPU ORACLE for the BEM Helmholtz hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch C restatement of the reference's Galerkin quadrature
 * (src/bem/cuda/include/{potential.h,integrand.h,gauss.h}, src/bem/cuda/bind.cu).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and there only as the checker or as the timed CPU baseline; the
 * product path (neuralat_amd) never calls it.
 *
 * Two precisions are built from one body (bem_oracle_impl.h):
 *   *_f64 : everything in double (the FP64 oracle; r-clamp switchable);
 *   *_f32 : mirror of the reference's single-precision arithmetic, including
 *           the thrust::complex<float> / double promotions of potential.h:18-65
 *           (the kernel tail is evaluated in double and narrowed on return).
 *
 * Quadrature rules are arguments (rows of (xi1, xi2, w) with sum(w) = 1 for
 * the triangle, rows of (s, w) on [-1, 1] for the line), so tests can pass the
 * reference's rules or high-order ones.
 *
 * Pinning of each function is recorded in tests/test_oracle_*.py.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { ORACLE_OK = 0, ORACLE_BAD_ARG = 1, ORACLE_BAD_MESH = 2 };
enum { ORACLE_SLP = 0, ORACLE_DLP = 1, ORACLE_ADLP = 2, ORACLE_HYP = 3, ORACLE_BM_LHS = 4, ORACLE_BM_RHS = 5 };
enum { ORACLE_P1 = 0, ORACLE_DP0 = 1 };

#define ORACLE_EPS 1e-3 /* potential.h:16 */
#define FOUR_PI_D (4.0 * 3.14159265358979323846)

/* ---- mesh adjacency: vertex->triangle CSR, neighbour CSR, colouring ---- */

typedef struct {
    int64_t *nb_ptr; /* [nt+1] */
    int32_t *nb_idx; /* neighbours (share >= 1 vertex, self included), ascending */
    int32_t *color;  /* greedy colouring: same colour => no shared vertex */
    int ncolors;
} adj_t;

static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

static int build_adjacency(const int32_t *tris, int64_t nt, int64_t nv, adj_t *A)
{
    int64_t *vptr = (int64_t *)calloc((size_t)nv + 1, sizeof(int64_t));
    for (int64_t t = 0; t < 3 * nt; ++t) vptr[tris[t] + 1]++;
    for (int64_t v = 0; v < nv; ++v) vptr[v + 1] += vptr[v];
    int32_t *vtri = (int32_t *)malloc(sizeof(int32_t) * (size_t)(3 * nt));
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)nv);
    memcpy(fill, vptr, sizeof(int64_t) * (size_t)nv);
    for (int64_t t = 0; t < nt; ++t)
        for (int a = 0; a < 3; ++a) vtri[fill[tris[3 * t + a]]++] = (int32_t)t;
    free(fill);
    A->nb_ptr = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nt + 1));
    int64_t cap = 16 * nt + 16, used = 0;
    A->nb_idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
    int32_t cand[512];
    for (int64_t t = 0; t < nt; ++t) {
        int nc = 0;
        for (int a = 0; a < 3; ++a) {
            int32_t v = tris[3 * t + a];
            for (int64_t e = vptr[v]; e < vptr[v + 1]; ++e) {
                if (nc >= 512) { free(vptr); free(vtri); return -1; }
                cand[nc++] = vtri[e];
            }
        }
        qsort(cand, (size_t)nc, sizeof(int32_t), cmp_i32);
        A->nb_ptr[t] = used;
        for (int c = 0; c < nc; ++c) {
            if (c > 0 && cand[c] == cand[c - 1]) continue;
            if (used >= cap) {
                cap *= 2;
                A->nb_idx = (int32_t *)realloc(A->nb_idx, sizeof(int32_t) * (size_t)cap);
            }
            A->nb_idx[used++] = cand[c];
        }
    }
    A->nb_ptr[nt] = used;
    A->color = (int32_t *)malloc(sizeof(int32_t) * (size_t)nt);
    A->ncolors = 0;
    for (int64_t t = 0; t < nt; ++t) {
        uint64_t usedmask[8] = {0};
        for (int64_t e = A->nb_ptr[t]; e < A->nb_ptr[t + 1]; ++e) {
            int32_t u = A->nb_idx[e];
            if (u < t) {
                int c = A->color[u];
                usedmask[c >> 6] |= 1ull << (c & 63);
            }
        }
        int c = 0;
        while (usedmask[c >> 6] & (1ull << (c & 63))) ++c;
        A->color[t] = c;
        if (c + 1 > A->ncolors) A->ncolors = c + 1;
    }
    free(vptr);
    free(vtri);
    return 0;
}

static void free_adjacency(adj_t *A)
{
    free(A->nb_ptr);
    free(A->nb_idx);
    free(A->color);
}

/* ---- pointwise kernels (potential.h:18-65) ------------------------------ */

/* FP64: G = e^{ikr}/(4 pi r), r = max(|x-y|, 1e-3) when clamp != 0.
 * (vr, vi) is the kernel value; (sr, si) the secondary value of BM_LHS
 * ((i/k) G, potential.h:55-65).  HYP is evaluated as SLP by the callers. */
static void kern_f64(int op, const double *x, const double *y, const double *nx, const double *ny, double k, int clamp,
                     double *vr, double *vi, double *sr, double *si)
{
    double d[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (clamp && r < ORACLE_EPS) r = ORACLE_EPS;
    double th = k * r, c = cos(th), s = sin(th);
    double g = 1.0 / (FOUR_PI_D * r);
    double dny = d[0] * ny[0] + d[1] * ny[1] + d[2] * ny[2];
    double dnx = d[0] * nx[0] + d[1] * nx[1] + d[2] * nx[2];
    double r2 = r * r;
    switch (op) {
    case ORACLE_SLP:
    case ORACLE_HYP:
        *vr = c * g; *vi = s * g;
        break;
    case ORACLE_DLP: { /* -G (1 - ikr) (y-x).n_y / r^2 */
        double f = -g * dny / r2;
        *vr = f * (c + th * s); *vi = f * (s - th * c);
        break;
    }
    case ORACLE_ADLP: { /* -G (1 - ikr) (x-y).n_x / r^2 */
        double f = g * dnx / r2;
        *vr = f * (c + th * s); *vi = f * (s - th * c);
        break;
    }
    case ORACLE_BM_LHS: { /* (-ik nx.ny + (1-ikr)(y-x).n_y/r^2) G ; (i/k) G */
        double nn = nx[0] * ny[0] + nx[1] * ny[1] + nx[2] * ny[2];
        double ar = dny / r2, ai = -k * nn - th * dny / r2;
        *vr = (ar * c - ai * s) * g; *vi = (ar * s + ai * c) * g;
        *sr = -s * g / k; *si = c * g / k;
        break;
    }
    case ORACLE_BM_RHS: { /* -G + (i/k) G (1-ikr) (x-y).n_x / r^2 */
        double q = -dnx / r2;
        double tr = th / k, ti = 1.0 / k; /* (i/k)(1 - i th) */
        double gr = c * g, gi = s * g;
        *vr = -gr + q * (gr * tr - gi * ti);
        *vi = -gi + q * (gr * ti + gi * tr);
        break;
    }
    default:
        *vr = *vi = 0;
    }
}

/* FP32 mirror: float geometry, sinf/cosf of the float phase, and the
 * thrust mixed complex<float>/double tail of potential.h evaluated in double
 * then narrowed to float, as the reference returns complex<float>. */
static void kern_f32(int op, const float *x, const float *y, const float *nx, const float *ny, float k, int clamp,
                     float *vr, float *vi, float *sr, float *si)
{
    float dxy[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
    float r = sqrtf(dxy[0] * dxy[0] + dxy[1] * dxy[1] + dxy[2] * dxy[2]);
    if (clamp && (double)r < ORACLE_EPS) r = (float)ORACLE_EPS;
    float rk = r * k;
    float c = cosf(rk), s = sinf(rk);
    float ymx[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
    float dny = ymx[0] * ny[0] + ymx[1] * ny[1] + ymx[2] * ny[2];
    float dnx = dxy[0] * nx[0] + dxy[1] * nx[1] + dxy[2] * nx[2]; /* (x-y).n_x */
    switch (op) {
    case ORACLE_SLP:
    case ORACLE_HYP: {
        double den = FOUR_PI_D * (double)r;
        *vr = (float)((double)c / den); *vi = (float)((double)s / den);
        break;
    }
    case ORACLE_DLP:
    case ORACLE_ADLP: {
        double den = FOUR_PI_D * (double)r * (double)r * (double)r;
        double a = -(double)c / den, b = -(double)s / den;
        double pr = a + b * (double)rk, pi = b - a * (double)rk; /* * (1 - i rk) */
        double dd = (op == ORACLE_DLP) ? (double)dny : (double)dnx;
        *vr = (float)(pr * dd); *vi = (float)(pi * dd);
        break;
    }
    case ORACLE_BM_LHS: {
        float nn = nx[0] * ny[0] + nx[1] * ny[1] + nx[2] * ny[2];
        float m = -k * nn;                          /* -i k (nx.ny), imaginary part */
        double rr = 1.0 / (double)(r * r);
        double tr = rr * (double)dny, ti = rr * (-(double)rk) * (double)dny; /* (1-ikr) dny / r^2 */
        double ar = tr, ai = (double)m + ti;
        double den = FOUR_PI_D * (double)r;
        double spr = (double)c / den, spi = (double)s / den;
        *vr = (float)(ar * spr - ai * spi); *vi = (float)(ar * spi + ai * spr);
        double ik = 1.0 / (double)k; /* complex(0,1)/k in float then promoted */
        float ikf = 1.0f / k;
        (void)ik;
        *sr = (float)(-(double)ikf * spi); *si = (float)((double)ikf * spr);
        break;
    }
    case ORACLE_BM_RHS: {
        double den = FOUR_PI_D * (double)r;
        double spr = (double)c / den, spi = (double)s / den;
        float ikf = 1.0f / k;
        double qr = -(double)ikf * spi, qi = (double)ikf * spr; /* (i/k) sp */
        double r2 = (double)(r * r);
        qr /= r2; qi /= r2;
        double pr = qr + qi * (double)rk, pi = qi - qr * (double)rk; /* * (1 - i rk) */
        pr *= (double)dnx; pi *= (double)dnx;
        *vr = (float)(-spr + pr); *vi = (float)(-spi + pi);
        break;
    }
    default:
        *vr = *vi = 0;
    }
}

/* ---- FP64 instantiation ---- */
#define R double
#define FN(name) name##_f64
#define KERNEL_EVAL kern_f64
#define IS_F32 0
#include "bem_oracle_impl.h"
#undef R
#undef FN
#undef KERNEL_EVAL
#undef IS_F32

/* ---- FP32 mirror instantiation ---- */
#define R float
#define FN(name) name##_f32
#define KERNEL_EVAL kern_f32
#define IS_F32 1
#include "bem_oracle_impl.h"
#undef R
#undef FN
#undef KERNEL_EVAL
#undef IS_F32

/* Neighbour lists as the oracle sees them (for tests): nb_ptr [nt+1],
 * nb_idx [cap], shared-vertex counts nb_shared [cap]; returns the number of
 * entries, or -1 on overflow / bad mesh. */
int64_t oracle_neighbors(const int32_t *tris, int64_t nt, int64_t nv, int64_t *nb_ptr, int32_t *nb_idx,
                         int32_t *nb_shared, int64_t cap, int32_t *ncolors)
{
    for (int64_t t = 0; t < 3 * nt; ++t)
        if (tris[t] < 0 || tris[t] >= nv) return -1;
    adj_t A;
    if (build_adjacency(tris, nt, nv, &A) != 0) return -1;
    int64_t n = A.nb_ptr[nt];
    if (n > cap) { free_adjacency(&A); return -1; }
    memcpy(nb_ptr, A.nb_ptr, sizeof(int64_t) * (size_t)(nt + 1));
    for (int64_t e = 0; e < n; ++e) nb_idx[e] = A.nb_idx[e];
    for (int64_t t = 0; t < nt; ++t)
        for (int64_t e = A.nb_ptr[t]; e < A.nb_ptr[t + 1]; ++e) {
            int32_t u = A.nb_idx[e], s = 0;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) s += (tris[3 * t + a] == tris[3 * (int64_t)u + b]);
            nb_shared[e] = s;
        }
    if (ncolors) *ncolors = A.ncolors;
    free_adjacency(&A);
    return n;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }
