// Batched solves over many (mesh, k) items of uneven size.
//
//  * bem_batch_cbie: the fused DP0 CBIE systems of all items in one
//    persistent launch (a device-side queue of tiles, largest items first),
//    then the neighbour pairs and the row reduction of each item.
//  * bem_batch_gmres: restarted GMRES on all items in lockstep.  Each item has
//    its own matrix and size; vectors of all items live in one concatenated
//    space (item s at offset off_s), and every kernel of an Arnoldi step covers
//    all items (grid y = item).  Per-item Givens rotations and convergence
//    flags stay on the device; a converged item skips its work.  A restart
//    cycle (residual and set-up, Arnoldi steps in segments, update) is
//    replayed from hipGraphs, each part captured on its third use; the host
//    reads the items' flags after each segment and ends the cycle once every
//    item is done.
//  * bem_batch_fgmres: the mixed-precision variant -- flexible GMRES in FP32
//    on the complex64 matrices whose directions come from the same lockstep
//    GMRES run on 16-bit copies of the matrices (diagonal kept in FP32).
// No reference counterpart: the reference solves one item at a time
// (src/scene.py:270-291, experiments/neuPAT/generate.py:23-33).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/bem_hip.h"
#include "bem_internal.hpp"
#include "bem_linalg.hpp"
#include "bem_lowp.hpp"
#include "bem_lowp_cvt.hpp"

struct bem_mesh {
    bem::MeshDev dev;
};

namespace bem {

__device__ __forceinline__ double bsum_d(double v, double* sh)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += sh[q];
    __syncthreads();
    return t;
}

// y_s = alpha A_s x_s + beta y_s, 4 rows per wavefront; grid (row blocks, items)
__global__ __launch_bounds__(256) void xgemv_kernel(const BSys* __restrict__ sys, const float2* __restrict__ x,
                                                    float2* __restrict__ y, float2 alpha, float2 beta, int beta_zero,
                                                    const int* __restrict__ skip)
{
    const BSys S = sys[blockIdx.y];
    if (skip && skip[blockIdx.y]) return;
    const int lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 4;
    if (r0 >= S.n) return;
    const float2* xs = x + S.off;
    float ar[4] = {0, 0, 0, 0}, ai[4] = {0, 0, 0, 0};
    // 16-byte non-temporal loads (two entries per lane) when the item's rows
    // and x are 16-byte aligned; uniform over the workgroup
    const bool vec = ((S.n | S.lda) & 1) == 0 && (((uintptr_t)S.A | (uintptr_t)xs) & 15) == 0;
    if (vec) {
        typedef float v4f __attribute__((ext_vector_type(4)));
        const int64_t n2 = S.n / 2, ld2 = S.lda / 2;
        const v4f* A4 = reinterpret_cast<const v4f*>(S.A);
        const float4* x4 = reinterpret_cast<const float4*>(xs);
        for (int64_t c = lane; c < n2; c += 64) {
            const float4 xv = x4[c];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t r = min(r0 + q, S.n - 1);
                const v4f a = __builtin_nontemporal_load(&A4[r * ld2 + c]);
                ar[q] = fmaf(a.x, xv.x, fmaf(-a.y, xv.y, ar[q]));
                ai[q] = fmaf(a.x, xv.y, fmaf(a.y, xv.x, ai[q]));
                ar[q] = fmaf(a.z, xv.z, fmaf(-a.w, xv.w, ar[q]));
                ai[q] = fmaf(a.z, xv.w, fmaf(a.w, xv.z, ai[q]));
            }
        }
    } else {
        for (int64_t c = lane; c < S.n; c += 64) {
            const float2 xv = xs[c];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t r = min(r0 + q, S.n - 1);
                const float2 a = S.A[r * S.lda + c];
                ar[q] = fmaf(a.x, xv.x, fmaf(-a.y, xv.y, ar[q]));
                ai[q] = fmaf(a.x, xv.y, fmaf(a.y, xv.x, ai[q]));
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float sr = ar[q], si = ai[q];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            sr += __shfl_xor(sr, o, 64);
            si += __shfl_xor(si, o, 64);
        }
        const int64_t r = r0 + q;
        if (lane == q && r < S.n) {
            float2 out = make_float2(alpha.x * sr - alpha.y * si, alpha.x * si + alpha.y * sr);
            if (!beta_zero) {
                const float2 yo = y[S.off + r];
                out.x += beta.x * yo.x - beta.y * yo.y;
                out.y += beta.x * yo.y + beta.y * yo.x;
            }
            y[S.off + r] = out;
        }
    }
}

// h[s][q] = <V_q, w> on item s; grid (nq, items)
__global__ __launch_bounds__(256) void xmultidot_kernel(const BSys* __restrict__ sys, const float2* __restrict__ V,
                                                        int64_t ntot, const float2* __restrict__ w,
                                                        double2* __restrict__ h, int hs, const int* __restrict__ done)
{
    __shared__ double sh[4];
    const int q = blockIdx.x, s = blockIdx.y;
    if (done[s]) return;
    const BSys S = sys[s];
    const float2* v = V + (int64_t)q * ntot + S.off;
    const float2* ww = w + S.off;
    double sr = 0.0, si = 0.0;
    if ((S.n & 1) == 0 && (((uintptr_t)v | (uintptr_t)ww) & 15) == 0) {
        const float4* v4 = reinterpret_cast<const float4*>(v);
        const float4* w4 = reinterpret_cast<const float4*>(ww);
#pragma unroll 4
        for (int64_t e = threadIdx.x; e < S.n / 2; e += 256) {
            const float4 a = v4[e], b = w4[e];
            sr += (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z + (double)a.w * b.w;
            si += (double)a.x * b.y - (double)a.y * b.x + (double)a.z * b.w - (double)a.w * b.z;
        }
    } else {
        for (int64_t e = threadIdx.x; e < S.n; e += 256) {
            const float2 a = v[e], b = ww[e];
            sr += (double)a.x * b.x + (double)a.y * b.y;
            si += (double)a.x * b.y - (double)a.y * b.x;
        }
    }
    sr = bsum_d(sr, sh);
    si = bsum_d(si, sh);
    if (threadIdx.x == 0) h[(int64_t)s * hs + q] = make_double2(sr, si);
}

// w_s -= sum_q h[s][q] V_q ; grid (element blocks, items)
// out_s[e] += sign * sum_{q < nq_s} c_s[q] V_q,s[e] (Gram-Schmidt update, sign
// -1, skipping converged items; solution update, sign +1, nq_s = nk[s]).  A
// workgroup owns 64 elements of one item; its 8 wavefronts take every 8th
// basis vector and meet in LDS in a fixed order.  grid (ceil(max n / 64), ns)
constexpr int kXWaves = 8;
__global__ __launch_bounds__(64 * kXWaves) void xcombine_kernel(const BSys* __restrict__ sys,
                                                                const float2* __restrict__ V, int64_t ntot,
                                                                int nq_fixed, const int* __restrict__ nq_arr,
                                                                const double2* __restrict__ c, int cstride,
                                                                double sign, float2* __restrict__ out,
                                                                const int* __restrict__ done)
{
    const int s = blockIdx.y;
    if (done && done[s]) return;  // uniform over the workgroup
    const BSys S = sys[s];
    if ((int64_t)blockIdx.x * 64 >= S.n) return;  // uniform: this item is shorter
    __shared__ double2 part[kXWaves][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t e = (int64_t)blockIdx.x * 64 + lane;
    const int nq = nq_arr ? nq_arr[s] : nq_fixed;
    const double2* cs = c + (int64_t)s * cstride;
    double ar = 0.0, ai = 0.0;
    if (e < S.n) {
        for (int q = wv; q < nq; q += kXWaves) {
            const double2 cq = cs[q];
            const float2 v = V[(int64_t)q * ntot + S.off + e];
            ar += cq.x * v.x - cq.y * v.y;
            ai += cq.x * v.y + cq.y * v.x;
        }
    }
    part[wv][lane] = make_double2(ar, ai);
    __syncthreads();
    if (wv == 0 && e < S.n) {
        double sr = 0.0, si = 0.0;
        for (int q = 0; q < kXWaves; ++q) {
            sr += part[q][lane].x;
            si += part[q][lane].y;
        }
        const float2 ov = out[S.off + e];
        out[S.off + e] = make_float2((float)((double)ov.x + sign * sr), (float)((double)ov.y + sign * si));
    }
}
__global__ __launch_bounds__(256) void xnorm2_kernel(const BSys* __restrict__ sys, const float2* __restrict__ w,
                                                     double* __restrict__ out)
{
    __shared__ double sh[4];
    const BSys S = sys[blockIdx.x];
    double a = 0.0;
    for (int64_t e = threadIdx.x; e < S.n; e += 256) {
        const float2 v = w[S.off + e];
        a += (double)v.x * v.x + (double)v.y * v.y;
    }
    a = bsum_d(a, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = a;
}

// dst_s = src_s / sqrt(nrm2[s]) ; grid (element blocks, items)
__global__ __launch_bounds__(256) void xnormalize_kernel(const BSys* __restrict__ sys, const float2* __restrict__ src,
                                                         const double* __restrict__ nrm2, float2* __restrict__ dst,
                                                         const int* __restrict__ done)
{
    const int s = blockIdx.y;
    const BSys S = sys[s];
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= S.n || (done && done[s])) return;
    const double nn = nrm2[s];
    const float sc = nn > 0.0 ? (float)(1.0 / sqrt(nn)) : 0.f;
    const float2 v = src[S.off + e];
    dst[S.off + e] = make_float2(v.x * sc, v.y * sc);
}

// Caller-side vectors of one item: its right-hand side and its solution
// (initial guess on entry)
struct XVec {
    const float2* b;
    float2* x;
};

// B_s = b_s and X_s = x_s into the concatenated space (scatter = false), or
// x_s = X_s back (scatter = true); grid (element blocks, items)
__global__ __launch_bounds__(256) void xgather_kernel(const BSys* __restrict__ sys, const XVec* __restrict__ vec,
                                                      float2* __restrict__ B, float2* __restrict__ X, bool scatter)
{
    const int s = blockIdx.y;
    const BSys S = sys[s];
    const XVec v = vec[s];
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < S.n; e += (int64_t)gridDim.x * 256) {
        if (scatter) {
            v.x[e] = X[S.off + e];
        } else {
            B[S.off + e] = v.b[e];
            X[S.off + e] = v.x[e];
        }
    }
}

__global__ __launch_bounds__(256) void xcopy_kernel(const float2* __restrict__ src, float2* __restrict__ dst, int64_t n)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < n) dst[e] = src[e];
}

// conv[s] = ||r_s|| <= target_s
__global__ void xconv_kernel(const double* __restrict__ rn2, const double* __restrict__ tgt, int* __restrict__ conv,
                             int ns)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < ns) conv[s] = sqrt(rn2[s]) <= tgt[s] ? 1 : 0;
}

// ---- 16-bit copies of the items' matrices (mixed-precision batch) ---------------
// Entry (r, c) of item s is the pair (re, im) of round(scale_s A_s(r, c)) at
// L[r * ldl + c] (real part in the low half); columns [n, ldl) are zero.  The
// diagonal stays in FP32 (diag[r] = A_s(r, r), L(r, r) = 0): it is the
// largest part of a second-kind operator (the M/2 term), and rounding it
// would dominate the difference between the 16-bit operator and A.
struct BLow {
    uint32_t* L;
    int64_t ldl;
    float2* diag;
    float* rinv;  // fp8: per-row 1 / scale (the 16-bit copies scale per item)
};

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int DT>
__device__ __forceinline__ float2 lp_pair(uint32_t u)
{
    if (DT == kDtBF16) return make_float2(__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u));
    return make_float2(f16_to_f32((uint16_t)(u & 0xffffu)), f16_to_f32((uint16_t)(u >> 16)));
}

__device__ __forceinline__ void cfma(float2 a, float xr, float xi, float& ar, float& ai)
{
    ar = fmaf(a.x, xr, fmaf(-a.y, xi, ar));
    ai = fmaf(a.x, xi, fmaf(a.y, xr, ai));
}

// y_s = alpha inv_scale_s Alow_s x_s + beta y_s.  A wavefront owns kLowRows
// rows; per step each lane reads 16 bytes (four entries) of every row, so a
// row streams 1 KB per wavefront step; grid (ceil(max n / (4 kLowRows)), items)
constexpr int kLowRows = 8;
template <int DT>
__global__ __launch_bounds__(256) void xgemv_lowp_kernel(const BSys* __restrict__ sys, const BLow* __restrict__ low,
                                                         const float* __restrict__ inv_scale,
                                                         const float2* __restrict__ x, float2* __restrict__ y,
                                                         float2 alpha, float2 beta, int beta_zero,
                                                         const int* __restrict__ skip)
{
    const int s = blockIdx.y;
    if (skip && skip[s]) return;
    const BSys S = sys[s];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + wave) * kLowRows;
    if (r0 >= S.n) return;
    const BLow Lw = low[s];
    const float2* xs = x + S.off;
    float ar[kLowRows], ai[kLowRows];
#pragma unroll
    for (int q = 0; q < kLowRows; ++q) ar[q] = ai[q] = 0.f;
    const int64_t n4 = S.n >> 2, ld4 = Lw.ldl >> 2;
    const u32x4* L4 = reinterpret_cast<const u32x4*>(Lw.L);
    const float4* x4 = reinterpret_cast<const float4*>(xs);
    // software-pipelined over a wave-uniform step count, as the 8-bit
    // product below: step t covers column words 64 t + lane, a lane past the
    // last word rereads it with x zeroed at use, and the next step's loads are
    // requested before this step's arithmetic
    typedef const u32x4 __attribute__((address_space(1)))* GRow;
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    typedef const f32x4 __attribute__((address_space(1)))* GVec;
    const GRow Lg = (GRow)L4;
    const GVec xg = (GVec)x4;
    int64_t roff[kLowRows];
#pragma unroll
    for (int q = 0; q < kLowRows; ++q) roff[q] = min(r0 + q, S.n - 1) * ld4;
    const int64_t nst = (n4 + 63) >> 6;
    const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};
    u32x4 a[kLowRows], an[kLowRows];
    f32x4 xv[2], xn[2];
    auto fetch = [&](int64_t t, u32x4 (&m)[kLowRows], f32x4 (&v)[2]) {
        const int64_t cc = min(t * 64 + lane, n4 - 1);
#pragma unroll
        for (int q = 0; q < kLowRows; ++q) m[q] = __builtin_nontemporal_load(&Lg[roff[q] + cc]);
        v[0] = xg[2 * cc];
        v[1] = xg[2 * cc + 1];
    };
    auto step = [&](int64_t t, const u32x4 (&m)[kLowRows], const f32x4 (&vr)[2]) {
        const bool in = t * 64 + lane < n4;
        const f32x4 xa = in ? vr[0] : z4, xb = in ? vr[1] : z4;
#pragma unroll
        for (int q = 0; q < kLowRows; ++q) {
            cfma(lp_pair<DT>(m[q].x), xa.x, xa.y, ar[q], ai[q]);
            cfma(lp_pair<DT>(m[q].y), xa.z, xa.w, ar[q], ai[q]);
            cfma(lp_pair<DT>(m[q].z), xb.x, xb.y, ar[q], ai[q]);
            cfma(lp_pair<DT>(m[q].w), xb.z, xb.w, ar[q], ai[q]);
        }
    };
    if (nst > 0) fetch(0, a, xv);
    int64_t t = 0;
    for (; t + 2 <= nst; t += 2) {
        fetch(t + 1, an, xn);
        __builtin_amdgcn_sched_barrier(0);
        step(t, a, xv);
        __builtin_amdgcn_sched_barrier(0);
        fetch(t + 2, a, xv);  // past the end: every lane rereads the last word
        __builtin_amdgcn_sched_barrier(0);
        step(t + 1, an, xn);
        __builtin_amdgcn_sched_barrier(0);
    }
    if (t < nst) step(t, a, xv);
    for (int64_t c = 4 * n4 + lane; c < S.n; c += 64) {  // the n % 4 last columns
        const float2 xv = xs[c];
#pragma unroll
        for (int q = 0; q < kLowRows; ++q) {
            const int64_t r = min(r0 + q, S.n - 1);
            cfma(lp_pair<DT>(Lw.L[r * Lw.ldl + c]), xv.x, xv.y, ar[q], ai[q]);
        }
    }
    const float isc = inv_scale[s];
#pragma unroll
    for (int q = 0; q < kLowRows; ++q) {
        float sr = ar[q], si = ai[q];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            sr += __shfl_xor(sr, o, 64);
            si += __shfl_xor(si, o, 64);
        }
        const int64_t r = r0 + q;
        if (lane == q && r < S.n) {
            const float2 d = Lw.diag[r], xv = xs[r];
            sr = fmaf(d.x, xv.x, fmaf(-d.y, xv.y, sr * isc));
            si = fmaf(d.x, xv.y, fmaf(d.y, xv.x, si * isc));
            float2 out = make_float2(alpha.x * sr - alpha.y * si, alpha.x * si + alpha.y * sr);
            if (!beta_zero) {
                const float2 yo = y[S.off + r];
                out.x += beta.x * yo.x - beta.y * yo.y;
                out.y += beta.x * yo.y + beta.y * yo.x;
            }
            y[S.off + r] = out;
        }
    }
}

// 8-bit copies (OCP e4m3, two entries per 32-bit word, real part in the low
// byte of each 16-bit half): entry (r, c) is round(scale_r A_s(r, c)) with a
// power-of-two scale per row (rows of meshes with unequal triangle sizes
// differ in magnitude; e4m3 keeps three mantissa bits and a narrow exponent
// range); L holds padded(n) / 2 words per row (ldl in words), diag the FP32
// diagonal and rinv[r] = 1 / scale_r.  The product: a wavefront owns kLowRows
// rows, each lane reads 16 bytes (eight entries) of every row per step.
typedef float f32x2 __attribute__((ext_vector_type(2)));

// rows per wavefront of the 8-bit product (4 and 16 were 26 % and 34 % slower
// on the batching solve: fewer bytes in flight, or fewer waves per SIMD; with
// the pipelined loop below, 4 rows still take 16 % longer per full call)
constexpr int kFp8Rows = 8;
__global__ __launch_bounds__(256) void xgemv_fp8_kernel(const BSys* __restrict__ sys, const BLow* __restrict__ low,
                                                        const float2* __restrict__ x, float2* __restrict__ y,
                                                        float2 alpha, float2 beta, int beta_zero,
                                                        const int* __restrict__ skip)
{
    const int s = blockIdx.y;
    if (skip && skip[s]) return;
    const BSys S = sys[s];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + wave) * kFp8Rows;
    if (r0 >= S.n) return;
    const BLow Lw = low[s];
    const float2* xs = x + S.off;
    float ar[kFp8Rows], ai[kFp8Rows];
#pragma unroll
    for (int q = 0; q < kFp8Rows; ++q) ar[q] = ai[q] = 0.f;
    const int64_t n8 = S.n >> 3, ld4 = Lw.ldl >> 2;
    const u32x4* L4 = reinterpret_cast<const u32x4*>(Lw.L);
    const float4* x4 = reinterpret_cast<const float4*>(xs);
    // Software-pipelined over a step count that is uniform across the
    // wavefront: step t covers column words 64 t + lane; a lane past the last
    // word reads the last word again with x taken as zero, so no lane leaves
    // the loop early and the next step's row words and x are requested before
    // this step's arithmetic (8 KB of matrix reads in flight per wavefront
    // while it decodes).  Global-address-space pointers keep the loads off
    // the flat path.  Per lane the summation order is the plain loop's.
    typedef const u32x4 __attribute__((address_space(1)))* GRow;
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    typedef const f32x4 __attribute__((address_space(1)))* GVec;
    const GRow Lg = (GRow)L4;
    const GVec xg = (GVec)x4;
    int64_t roff[kFp8Rows];
#pragma unroll
    for (int q = 0; q < kFp8Rows; ++q) roff[q] = min(r0 + q, S.n - 1) * ld4;
    const int64_t nst = (n8 + 63) >> 6;
    const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};
    u32x4 a[kFp8Rows], an[kFp8Rows];
    f32x4 xv[4], xn[4];
    // x is zeroed past the last word at use, not at the request, so that
    // nothing waits on the loads before the arithmetic of the step before
    auto fetch = [&](int64_t t, u32x4 (&m)[kFp8Rows], f32x4 (&v)[4]) {
        const int64_t cc = min(t * 64 + lane, n8 - 1);
#pragma unroll
        for (int q = 0; q < kFp8Rows; ++q) m[q] = __builtin_nontemporal_load(&Lg[roff[q] + cc]);
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = xg[4 * cc + i];
    };
    auto step = [&](int64_t t, const u32x4 (&m)[kFp8Rows], const f32x4 (&vr)[4]) {
        const bool in = t * 64 + lane < n8;
        f32x4 v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = in ? vr[i] : z4;
#pragma unroll
        for (int q = 0; q < kFp8Rows; ++q) {
            const uint32_t w[4] = {m[q].x, m[q].y, m[q].z, m[q].w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                f32x2 e = __builtin_amdgcn_cvt_pk_f32_fp8((int)w[j], false);
                cfma(make_float2(e.x, e.y), v[j].x, v[j].y, ar[q], ai[q]);
                e = __builtin_amdgcn_cvt_pk_f32_fp8((int)w[j], true);
                cfma(make_float2(e.x, e.y), v[j].z, v[j].w, ar[q], ai[q]);
            }
        }
    };
    if (nst > 0) fetch(0, a, xv);
    int64_t t = 0;
    for (; t + 2 <= nst; t += 2) {  // ping-pong: no register copies
        fetch(t + 1, an, xn);
        __builtin_amdgcn_sched_barrier(0);  // keep the requests ahead of the arithmetic
        step(t, a, xv);
        __builtin_amdgcn_sched_barrier(0);
        fetch(t + 2, a, xv);  // past the end: every lane rereads the last word
        __builtin_amdgcn_sched_barrier(0);
        step(t + 1, an, xn);
        __builtin_amdgcn_sched_barrier(0);
    }
    if (t < nst) step(t, a, xv);
    for (int64_t c = 8 * n8 + lane; c < S.n; c += 64) {  // the n % 8 last columns
        const float2 xv = xs[c];
#pragma unroll
        for (int q = 0; q < kFp8Rows; ++q) {
            const int64_t r = min(r0 + q, S.n - 1);
            const int w = (int)Lw.L[r * Lw.ldl + (c >> 1)];
            const f32x2 e = (c & 1) ? __builtin_amdgcn_cvt_pk_f32_fp8(w, true) : __builtin_amdgcn_cvt_pk_f32_fp8(w, false);
            cfma(make_float2(e.x, e.y), xv.x, xv.y, ar[q], ai[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < kFp8Rows; ++q) {
        float sr = ar[q], si = ai[q];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            sr += __shfl_xor(sr, o, 64);
            si += __shfl_xor(si, o, 64);
        }
        const int64_t r = r0 + q;
        if (lane == q && r < S.n) {
            const float isc = Lw.rinv[r];
            const float2 d = Lw.diag[r], xv = xs[r];
            sr = fmaf(d.x, xv.x, fmaf(-d.y, xv.y, sr * isc));
            si = fmaf(d.x, xv.y, fmaf(d.y, xv.x, si * isc));
            float2 out = make_float2(alpha.x * sr - alpha.y * si, alpha.x * si + alpha.y * sr);
            if (!beta_zero) {
                const float2 yo = y[S.off + r];
                out.x += beta.x * yo.x - beta.y * yo.y;
                out.y += beta.x * yo.y + beta.y * yo.x;
            }
            y[S.off + r] = out;
        }
    }
}

// The 8-bit copies: per row, the largest |Re|, |Im| off the diagonal m (a
// block reduction), scale_r = 2^(8 - e) with m < 2^e (so |scale_r A| < 256,
// inside e4m3's 448), then round to nearest even; pad columns and the
// diagonal zero.  grid (row blocks, items)
__global__ __launch_bounds__(256) void xto_fp8_kernel(const BSys* __restrict__ sys, const BLow* __restrict__ low)
{
    __shared__ float wmax[4];
    const int s = blockIdx.y;
    const BSys S = sys[s];
    const BLow Lw = low[s];
    for (int64_t r = blockIdx.x; r < S.n; r += gridDim.x) {
        const float2* row = S.A + r * S.lda;
        float m = 0.f;
        for (int64_t c = threadIdx.x; c < S.n; c += 256)
            if (c != r) m = fmaxf(m, fmaxf(fabsf(row[c].x), fabsf(row[c].y)));
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
        __syncthreads();
        m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
        __syncthreads();  // wmax is rewritten for the next row
        float sc = 1.f;
        if (m > 0.f) {
            int e = 0;
            (void)frexpf(m, &e);  // m < 2^e
            sc = ldexpf(1.f, 8 - e);
        }
        if (threadIdx.x == 0) {
            Lw.rinv[r] = 1.f / sc;
            Lw.diag[r] = row[r];
        }
        for (int64_t w = threadIdx.x; w < Lw.ldl; w += 256) {
            const int64_t c0 = 2 * w, c1 = 2 * w + 1;
            const float2 v0 = (c0 < S.n && c0 != r) ? row[c0] : make_float2(0.f, 0.f);
            const float2 v1 = (c1 < S.n && c1 != r) ? row[c1] : make_float2(0.f, 0.f);
            int u = __builtin_amdgcn_cvt_pk_fp8_f32(v0.x * sc, v0.y * sc, 0, false);
            u = __builtin_amdgcn_cvt_pk_fp8_f32(v1.x * sc, v1.y * sc, u, true);
            Lw.L[r * Lw.ldl + w] = (uint32_t)u;
        }
    }
}

// amax_bits[s] = bits of max |Re|, |Im| over A_s (non-negative floats order as
// their bit patterns); amax_bits zeroed beforehand; grid (row blocks, items)
__global__ __launch_bounds__(256) void xamax_kernel(const BSys* __restrict__ sys, unsigned int* __restrict__ amax_bits)
{
    const int s = blockIdx.y;
    const BSys S = sys[s];
    float m = 0.f;
    for (int64_t r = blockIdx.x; r < S.n; r += gridDim.x)
        for (int64_t c = threadIdx.x; c < S.n; c += 256) {
            const float2 v = S.A[r * S.lda + c];
            m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
        }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(&amax_bits[s], __float_as_uint(m));
}

// scale_s: 1 for bf16; for fp16 the largest power of two with
// max |A_s| scale_s <= 2^14 (fp16's largest finite value is 65504)
__global__ void xscale_kernel(const unsigned int* __restrict__ amax_bits, int dt, float* __restrict__ scale,
                              float* __restrict__ inv_scale, int ns)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    float sc = 1.f;
    const float m = __uint_as_float(amax_bits[s]);
    if (dt == kDtFP16 && m > 0.f) {
        int e = 0;
        const float f = frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
        sc = ldexpf(1.f, (f == 0.5f ? 15 : 14) - e);
    }
    scale[s] = sc;
    inv_scale[s] = 1.f / sc;
}

// Alow_s = round(scale_s A_s) off the diagonal, pad columns and the diagonal
// zero, diag = the FP32 diagonal; grid (row blocks, items)
template <int DT>
__global__ __launch_bounds__(256) void xto_lowp_kernel(const BSys* __restrict__ sys, const BLow* __restrict__ low,
                                                       const float* __restrict__ scale)
{
    const int s = blockIdx.y;
    const BSys S = sys[s];
    const BLow Lw = low[s];
    const float sc = scale[s];
    for (int64_t r = blockIdx.x; r < S.n; r += gridDim.x)
        for (int64_t c = threadIdx.x; c < Lw.ldl; c += 256) {
            float2 v = c < S.n ? S.A[r * S.lda + c] : make_float2(0.f, 0.f);
            if (c == r) {
                Lw.diag[r] = v;
                v = make_float2(0.f, 0.f);
            }
            Lw.L[r * Lw.ldl + c] = (uint32_t)to_lp<DT>(v.x * sc) | ((uint32_t)to_lp<DT>(v.y * sc) << 16);
        }
}

__global__ void xfill_kernel(unsigned int* __restrict__ p, int64_t n, unsigned int v)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < n) p[e] = v;
}

// Inner-solve targets for the unit-norm outer direction v_j: relaxed as the
// outer residual falls (inexact Krylov: the outer step only needs the inner
// error below target / |r_j|), never tighter than tol_min; items whose outer
// solve is done get a target every residual meets.  cur: |r_0|^2 (first step)
// or |r_j| (the outer Givens estimate) per item.
__global__ void xinner_tgt_kernel(const int* __restrict__ outer_done, const double* __restrict__ outer_tgt,
                                  const double* __restrict__ cur, int cur_is_square, double tol_min,
                                  double* __restrict__ tgt, int ns)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const double r = cur_is_square ? sqrt(cur[s]) : cur[s];
    const double rel = r > 0.0 ? outer_tgt[s] / r : 1.0;
    tgt[s] = outer_done[s] ? 1e300 : fmin(0.5, fmax(tol_min, rel));
}

namespace {

size_t al(size_t b) { return (b + 255) / 256 * 256; }

struct XLayout {
    size_t sys, vec, V, w, r, b, x, H, h, h2, cs, sn, g, y, res, rn2, wn2, tgt, done, nk, conv, total;
};

XLayout xlayout(int ns, int64_t ntot, int m)
{
    XLayout L{};
    size_t o = 0;
    auto take = [&](size_t b) {
        size_t at = o;
        o += al(b);
        return at;
    };
    L.sys = take(sizeof(BSys) * ns);
    L.vec = take(sizeof(XVec) * ns);
    L.V = take(sizeof(float2) * (size_t)ntot * (m + 1));
    L.w = take(sizeof(float2) * ntot);
    L.r = take(sizeof(float2) * ntot);
    L.b = take(sizeof(float2) * ntot);
    L.x = take(sizeof(float2) * ntot);
    L.H = take(sizeof(double2) * (size_t)ns * (m + 1) * m);
    L.h = take(sizeof(double2) * (size_t)ns * (m + 1));
    L.h2 = take(sizeof(double2) * (size_t)ns * (m + 1));
    L.cs = take(sizeof(double2) * (size_t)ns * m);
    L.sn = take(sizeof(double2) * (size_t)ns * m);
    L.g = take(sizeof(double2) * (size_t)ns * (m + 1));
    L.y = take(sizeof(double2) * (size_t)ns * m);
    L.res = take(sizeof(double) * ns);
    L.rn2 = take(sizeof(double) * ns);
    L.wn2 = take(sizeof(double) * ns);
    L.tgt = take(sizeof(double) * ns);
    L.done = take(sizeof(int) * ns);
    L.nk = take(sizeof(int) * ns);
    L.conv = take(sizeof(int) * ns);
    L.total = o;
    return L;
}

int64_t padded(int64_t n) { return (n + 63) / 64 * 64; }

// One lockstep GMRES instance over the concatenated space of ns items.
struct XEng {
    int ns = 0, m = 0, dt = 0;
    int64_t ntot = 0, nmax = 0;
    const BSys* dsys = nullptr;
    const BLow* dlow = nullptr;  // 16-bit copies of the matrices (null: the complex64 matrices)
    const float* inv_scale = nullptr;
    const int* skip = nullptr;   // items left out of the residual products (null: none)
    bool x0_zero = false;        // cycles start from x = 0 (r = b, no product)
    bool final_residual = true;  // cycles end with the true residual and the flags conv
    bool one_pass = false;       // one Gram-Schmidt pass per step (h2 kept zero)
    float2 *V = nullptr, *w = nullptr, *r = nullptr, *B = nullptr, *X = nullptr;
    double2 *H = nullptr, *h = nullptr, *h2 = nullptr, *cs = nullptr, *sn = nullptr, *g = nullptr, *y = nullptr;
    double *res = nullptr, *rn2 = nullptr, *wn2 = nullptr, *tgt = nullptr;
    int *done = nullptr, *nk = nullptr, *conv = nullptr;
};

XEng xeng_at(char* W, const XLayout& L, const BSys* dsys, int ns, int64_t ntot, int64_t nmax, int m)
{
    XEng E;
    E.ns = ns;
    E.m = m;
    E.ntot = ntot;
    E.nmax = nmax;
    E.dsys = dsys;
    E.V = (float2*)(W + L.V);
    E.w = (float2*)(W + L.w);
    E.r = (float2*)(W + L.r);
    E.B = (float2*)(W + L.b);
    E.X = (float2*)(W + L.x);
    E.H = (double2*)(W + L.H);
    E.h = (double2*)(W + L.h);
    E.h2 = (double2*)(W + L.h2);
    E.cs = (double2*)(W + L.cs);
    E.sn = (double2*)(W + L.sn);
    E.g = (double2*)(W + L.g);
    E.y = (double2*)(W + L.y);
    E.res = (double*)(W + L.res);
    E.rn2 = (double*)(W + L.rn2);
    E.wn2 = (double*)(W + L.wn2);
    E.tgt = (double*)(W + L.tgt);
    E.done = (int*)(W + L.done);
    E.nk = (int*)(W + L.nk);
    E.conv = (int*)(W + L.conv);
    return E;
}

const float2 kOne = {1.f, 0.f}, kZero = {0.f, 0.f}, kMinusOne = {-1.f, 0.f};

hipError_t xmatvec(const XEng& E, const float2* x, float2* y, float2 alpha, float2 beta, int beta_zero,
                   const int* skip, hipStream_t st)
{
    if (!E.dlow) {
        hipLaunchKernelGGL(xgemv_kernel, dim3((unsigned)((E.nmax + 15) / 16), (unsigned)E.ns), dim3(256), 0, st,
                           E.dsys, x, y, alpha, beta, beta_zero, skip);
    } else {
        const dim3 g((unsigned)((E.nmax + 4 * kLowRows - 1) / (4 * kLowRows)), (unsigned)E.ns);
        if (E.dt == kDtFP8)
            hipLaunchKernelGGL(xgemv_fp8_kernel, dim3((unsigned)((E.nmax + 4 * kFp8Rows - 1) / (4 * kFp8Rows)), g.y),
                               dim3(256), 0, st, E.dsys, E.dlow, x, y, alpha, beta, beta_zero, skip);
        else if (E.dt == kDtBF16)
            hipLaunchKernelGGL(xgemv_lowp_kernel<kDtBF16>, g, dim3(256), 0, st, E.dsys, E.dlow, E.inv_scale, x, y,
                               alpha, beta, beta_zero, skip);
        else
            hipLaunchKernelGGL(xgemv_lowp_kernel<kDtFP16>, g, dim3(256), 0, st, E.dsys, E.dlow, E.inv_scale, x, y,
                               alpha, beta, beta_zero, skip);
    }
    return hipGetLastError();
}

// r = B - A X, rn2 = |r|^2 per item
hipError_t xresidual(const XEng& E, hipStream_t st)
{
    hipLaunchKernelGGL(xcopy_kernel, dim3((unsigned)((E.ntot + 255) / 256)), dim3(256), 0, st, E.B, E.r, E.ntot);
    hipError_t e = xmatvec(E, E.X, E.r, kMinusOne, kOne, 0, E.skip, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(xnorm2_kernel, dim3(E.ns), dim3(256), 0, st, E.dsys, E.r, E.rn2);
    return hipGetLastError();
}

// cycle start: residual, Givens right-hand side, flags, first basis vector
hipError_t xcycle_init(const XEng& E, hipStream_t st)
{
    hipError_t e;
    if (E.x0_zero) {
        hipLaunchKernelGGL(xcopy_kernel, dim3((unsigned)((E.ntot + 255) / 256)), dim3(256), 0, st, E.B, E.r, E.ntot);
        hipLaunchKernelGGL(xnorm2_kernel, dim3(E.ns), dim3(256), 0, st, E.dsys, E.r, E.rn2);
        e = hipGetLastError();
    } else {
        e = xresidual(E, st);
    }
    if (e != hipSuccess) return e;
    e = launch_binit(E.g, E.m, E.rn2, E.tgt, E.done, E.nk, E.ns, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(xnormalize_kernel, dim3((unsigned)((E.nmax + 255) / 256), (unsigned)E.ns), dim3(256), 0, st,
                       E.dsys, E.r, E.rn2, E.V, (const int*)nullptr);
    return hipGetLastError();
}

// Steps whose Gram-Schmidt reads at most this many basis entries per system
// ((j + 1) n) run after the product as one launch, a workgroup per system
// (launch_barnoldi_small); larger ones spread each system's dot products and
// combination over many workgroups in separate launches.
constexpr int64_t kXFusedStep = (int64_t)1 << 16;

// Arnoldi steps [j0, j1): product, classical Gram-Schmidt with one
// re-orthogonalisation, normalisation, Givens rotation and convergence flag
hipError_t xcycle_steps(const XEng& E, int j0, int j1, hipStream_t st)
{
    const int m = E.m;
    const dim3 gx((unsigned)((E.nmax + 255) / 256), (unsigned)E.ns);
    const dim3 gc((unsigned)((E.nmax + 63) / 64), (unsigned)E.ns);
    for (int j = j0; j < j1; ++j) {
        hipError_t e = xmatvec(E, E.V + (int64_t)j * E.ntot, E.w, kOne, kZero, 1, E.done, st);
        if (e != hipSuccess) return e;
        if ((int64_t)(j + 1) * E.nmax <= kXFusedStep) {  // small Gram-Schmidt: the rest of the step in one launch
            e = launch_barnoldi_small(E.dsys, E.ns, E.V, E.ntot, j, E.w, E.h, E.h2, m + 1, E.H, m, E.cs, E.sn, E.g,
                                      E.res, E.tgt, E.done, E.nk, E.wn2, E.one_pass ? 1 : 2, st);
            if (e != hipSuccess) return e;
            continue;
        }
        hipLaunchKernelGGL(xmultidot_kernel, dim3(j + 1, E.ns), dim3(256), 0, st, E.dsys, E.V, E.ntot, E.w, E.h,
                           m + 1, (const int*)E.done);
        hipLaunchKernelGGL(xcombine_kernel, gc, dim3(64 * kXWaves), 0, st, E.dsys, E.V, E.ntot, j + 1,
                           (const int*)nullptr, E.h, m + 1, -1.0, E.w, (const int*)E.done);
        if (!E.one_pass) {
            hipLaunchKernelGGL(xmultidot_kernel, dim3(j + 1, E.ns), dim3(256), 0, st, E.dsys, E.V, E.ntot, E.w,
                               E.h2, m + 1, (const int*)E.done);
            hipLaunchKernelGGL(xcombine_kernel, gc, dim3(64 * kXWaves), 0, st, E.dsys, E.V, E.ntot, j + 1,
                               (const int*)nullptr, E.h2, m + 1, -1.0, E.w, (const int*)E.done);
        }
        hipLaunchKernelGGL(xnorm2_kernel, dim3(E.ns), dim3(256), 0, st, E.dsys, E.w, E.wn2);
        hipLaunchKernelGGL(xnormalize_kernel, gx, dim3(256), 0, st, E.dsys, E.w, E.wn2,
                           E.V + (int64_t)(j + 1) * E.ntot, (const int*)E.done);
        e = launch_bgivens(E.H, m, j, E.h, E.h2, m + 1, E.wn2, E.cs, E.sn, E.g, E.res, E.tgt, E.done, E.nk, E.ns,
                           st);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

// cycle end: back-solve, x += V y, then (final_residual) the true residual
// and the flags conv
hipError_t xcycle_fin(const XEng& E, hipStream_t st)
{
    hipError_t e = launch_bbacksolve(E.H, E.m, E.nk, E.g, E.y, E.ns, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(xcombine_kernel, dim3((unsigned)((E.nmax + 63) / 64), (unsigned)E.ns), dim3(64 * kXWaves), 0,
                       st, E.dsys, E.V, E.ntot, 0, (const int*)E.nk, E.y, E.m, 1.0, E.X, (const int*)nullptr);
    if (E.final_residual) {
        e = xresidual(E, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(xconv_kernel, dim3((E.ns + 63) / 64), dim3(64), 0, st, E.rn2, E.tgt, E.conv, E.ns);
    }
    return hipGetLastError();
}

// Arnoldi steps per graph segment; the host reads the flags between segments.
// The first segment is short: a restart cycle that follows a nearly converged
// one (the implicit residual reached the target, the true one not quite)
// usually needs a step or two.
constexpr int kSeg0 = 4, kSeg = 20;
inline int seg_count(int m) { return m <= kSeg0 ? 1 : 1 + (m - kSeg0 + kSeg - 1) / kSeg; }
inline int seg_begin(int k) { return k == 0 ? 0 : kSeg0 + (k - 1) * kSeg; }

// A restart cycle as graphs: slot 0 the start, 1 .. nseg the step segments,
// nseg + 1 the end.  A slot runs as direct launches on its first two uses, is
// captured on its third and replayed from then on: a capture pays only when
// the graph is replayed (later restart cycles, the inner solves of later
// outer steps), and solves done in one or two cycles capture nothing.
struct XGraphs {
    std::vector<hipGraph_t> g;
    std::vector<hipGraphExec_t> x;
    std::vector<int> uses;
    hipStream_t cap = nullptr;
    XGraphs() = default;
    XGraphs(const XGraphs&) = delete;
    XGraphs& operator=(const XGraphs&) = delete;
    ~XGraphs()
    {
        for (auto e : x)
            if (e) (void)hipGraphExecDestroy(e);
        for (auto q : g)
            if (q) (void)hipGraphDestroy(q);
        if (cap) (void)hipStreamDestroy(cap);
    }
};

// slot of G through f (a function of the stream that launches the slot's work)
template <class F>
hipError_t xrun_slot(XGraphs* G, int nslots, int slot, hipStream_t s, F&& f)
{
    if (!G) return f(s);
    if (G->x.empty()) {
        G->g.assign(nslots, nullptr);
        G->x.assign(nslots, nullptr);
        G->uses.assign(nslots, 0);
    }
    if (G->x[slot]) return hipGraphLaunch(G->x[slot], s);
    if (++G->uses[slot] < 3) return f(s);
    hipError_t e = G->cap ? hipSuccess : hipStreamCreateWithFlags(&G->cap, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    e = hipStreamBeginCapture(G->cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    const hipError_t e1 = f(G->cap);
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(G->cap, &graph);
    G->g[slot] = graph;
    if (e1 != hipSuccess) return e1;
    if (e2 != hipSuccess) return e2;
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return e;
    G->x[slot] = ex;
    return hipGraphLaunch(ex, s);
}

// One restart cycle on stream s (through the graphs of G when given).  The
// host reads the items' flags after each segment of Arnoldi steps and ends
// the cycle early once every item is done.
hipError_t xrun_cycle(const XEng& E, XGraphs* G, hipStream_t s, std::vector<int>& hd)
{
    const int nseg = seg_count(E.m), nslots = nseg + 2;
    hipError_t e = xrun_slot(G, nslots, 0, s, [&](hipStream_t st) { return xcycle_init(E, st); });
    if (e != hipSuccess) return e;
    for (int k = 0; k < nseg; ++k) {
        e = xrun_slot(G, nslots, 1 + k, s,
                      [&](hipStream_t st) { return xcycle_steps(E, seg_begin(k), std::min(E.m, seg_begin(k + 1)), st); });
        if (e != hipSuccess) return e;
        e = hipMemcpyAsync(hd.data(), E.done, sizeof(int) * E.ns, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        bool all = true;
        for (int q = 0; q < E.ns && all; ++q) all = hd[q] != 0;
        if (all) break;
    }
    return xrun_slot(G, nslots, nseg + 1, s, [&](hipStream_t st) { return xcycle_fin(E, st); });
}

// Workspace of the mixed-precision batch solve
struct FLayout {
    XLayout o, in;
    size_t o_at, z_at, in_at, low_at, scale_at, inv_at, amax_at, diag_at, rinv_at, planes_at, total;
    std::vector<size_t> plane_at;  // per item, from planes_at
};

FLayout flayout(const bem_batch_system* sys, int ns, int64_t ntot, int mo, int mi)
{
    FLayout F;
    size_t o = 0;
    F.o = xlayout(ns, ntot, mo);
    F.o_at = o;
    o += F.o.total;
    F.z_at = o;
    o += al(sizeof(float2) * (size_t)ntot * mo);
    F.in = xlayout(ns, ntot, mi);
    F.in_at = o;
    o += F.in.total;
    F.low_at = o;
    o += al(sizeof(BLow) * ns);
    F.scale_at = o;
    o += al(sizeof(float) * ns);
    F.inv_at = o;
    o += al(sizeof(float) * ns);
    F.amax_at = o;
    o += al(sizeof(unsigned int) * ns);
    F.diag_at = o;
    o += al(sizeof(float2) * (size_t)ntot);
    F.rinv_at = o;
    o += al(sizeof(float) * (size_t)ntot);
    F.planes_at = o;
    F.plane_at.resize(ns);
    for (int q = 0; q < ns; ++q) {
        F.plane_at[q] = o;
        o += al(sizeof(uint32_t) * (size_t)sys[q].n * padded(sys[q].n));
    }
    F.total = o;
    return F;
}

thread_local std::string g_batch_err;

std::mutex g_tune_mu;
std::map<std::pair<int, int>, int> g_tune;  // (device, log2 of largest nt) -> rows per task

}  // namespace

}  // namespace bem

using namespace bem;

#define XTRY(expr)                                                                       \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) {                                                          \
            g_batch_err = std::string(#expr) + ": " + hipGetErrorString(_e);            \
            return BEM_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)

extern "C" {

const char* bem_batch_last_error(void) { return g_batch_err.c_str(); }

// ---- batched assembly ----------------------------------------------------------
int64_t bem_batch_cbie_work_size(const bem_batch_item* items, int n_items)
{
    if (!items || n_items <= 0) return 0;
    size_t b = al(sizeof(BatchItemDev) * n_items) + 256;
    for (int q = 0; q < n_items; ++q) {
        if (!items[q].mesh) return 0;
        b += al((size_t)cbie_work_bytes(items[q].mesh->dev, 16));  // the largest row-tile layout
    }
    return (int64_t)b;
}

static int batch_run(const bem_batch_item* items, int n_items, int clamp, int rt, void* work, hipStream_t s,
                     bool singular)
{
    std::vector<int> order(n_items);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return items[a].mesh->dev.nt > items[b].mesh->dev.nt; });
    char* W = (char*)work;
    BatchItemDev* table = (BatchItemDev*)W;
    unsigned long long* counter = (unsigned long long*)(W + al(sizeof(BatchItemDev) * n_items));
    char* cur = W + al(sizeof(BatchItemDev) * n_items) + 256;
    std::vector<char*> wk(n_items);
    for (int q = 0; q < n_items; ++q) {
        wk[q] = cur;
        cur += al((size_t)cbie_work_bytes(items[q].mesh->dev, 16));
    }
    std::vector<BatchItemDev> h(n_items);
    int64_t task = 0;
    for (int e = 0; e < n_items; ++e) {
        const int q = order[e];
        const MeshDev& m = items[q].mesh->dev;
        BatchItemDev d{};
        d.rec = m.rec6;
        d.tris = m.tris;
        d.nt = m.nt;
        d.k = items[q].k;
        d.psi = (const float2*)items[q].psi;
        d.lhs = (float2*)items[q].lhs;
        d.ld = items[q].ld;
        const CbieWork w = cbie_work_split(m, rt, wk[q]);
        d.prow = w.prow;
        d.pcol = w.pcol;
        d.task0 = task;
        task += batch_cbie_tasks(m.nt, rt);
        h[e] = d;
    }
    XTRY(hipMemcpyAsync(table, h.data(), sizeof(BatchItemDev) * n_items, hipMemcpyHostToDevice, s));
    XTRY(hipStreamSynchronize(s));
    XTRY(launch_batch_cbie_persistent(table, n_items, task, rt, counter, clamp, s));
    if (singular)
        for (int q = 0; q < n_items; ++q)
            XTRY(launch_cbie_dp0(items[q].mesh->dev, items[q].k, 0, clamp, 2, (const float2*)items[q].psi,
                                 (float2*)items[q].lhs, items[q].ld, (float2*)items[q].rhs,
                                 cbie_work_split(items[q].mesh->dev, rt, wk[q]), s));
    return BEM_OK;
}

int bem_batch_autotune(const bem_batch_item* items, int n_items, int clamp, void* work, int* best_rt, float* ms,
                       void* stream)
{
    if (!items || n_items <= 0 || !work) return BEM_ERR_BAD_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // tune on the largest item alone
    int big = 0;
    for (int q = 1; q < n_items; ++q)
        if (items[q].mesh->dev.nt > items[big].mesh->dev.nt) big = q;
    const int cands[3] = {16, 32, 64};
    hipEvent_t e0, e1;
    XTRY(hipEventCreate(&e0));
    XTRY(hipEventCreate(&e1));
    float bestms = 1e30f;
    int bestv = 64;
    for (int c = 0; c < 3; ++c) {
        for (int rep = 0; rep < 2; ++rep) {  // first run warms up
            XTRY(hipEventRecord(e0, s));
            if (int rc = batch_run(&items[big], 1, clamp, cands[c], work, s, false)) return rc;
            XTRY(hipEventRecord(e1, s));
            XTRY(hipEventSynchronize(e1));
        }
        float t = 0.f;
        XTRY(hipEventElapsedTime(&t, e0, e1));
        if (ms) ms[c] = t;
        if (t < bestms) {
            bestms = t;
            bestv = cands[c];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int dev = 0;
    (void)hipGetDevice(&dev);
    const int bucket = (int)std::floor(std::log2((double)items[big].mesh->dev.nt));
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[{dev, bucket}] = bestv;
    }
    if (best_rt) *best_rt = bestv;
    return BEM_OK;
}

int bem_batch_cbie(const bem_batch_item* items, int n_items, int clamp, int row_tile, void* work, void* stream)
{
    if (!items || n_items <= 0 || !work) return BEM_ERR_BAD_ARG;
    for (int q = 0; q < n_items; ++q) {
        const auto& it = items[q];
        if (!it.mesh || !it.psi || !it.lhs || !it.rhs || it.ld < it.mesh->dev.nt || !std::isfinite(it.k)) {
            g_batch_err = "bad item " + std::to_string(q);
            return BEM_ERR_BAD_ARG;
        }
    }
    if (row_tile == 0) {
        int big = 0;
        for (int q = 1; q < n_items; ++q)
            if (items[q].mesh->dev.nt > items[big].mesh->dev.nt) big = q;
        int dev = 0;
        (void)hipGetDevice(&dev);
        const int bucket = (int)std::floor(std::log2((double)items[big].mesh->dev.nt));
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto f = g_tune.find({dev, bucket});
        row_tile = f != g_tune.end() ? f->second : 64;
    }
    if (row_tile != 16 && row_tile != 32 && row_tile != 64) {
        g_batch_err = "row_tile must be 0 (tuned), 16, 32 or 64";
        return BEM_ERR_BAD_ARG;
    }
    return batch_run(items, n_items, clamp, row_tile, work, reinterpret_cast<hipStream_t>(stream), true);
}

// ---- batched GMRES -----------------------------------------------------------
int64_t bem_batch_gmres_work_size(const bem_batch_system* sys, int ns, int restart)
{
    if (!sys || ns <= 0 || restart <= 0) return 0;
    int64_t ntot = 0;
    for (int q = 0; q < ns; ++q) ntot += padded(sys[q].n);
    return (int64_t)xlayout(ns, ntot, restart).total;
}

static int check_systems(const bem_batch_system* sys, int ns, std::vector<BSys>& hs, int64_t& ntot, int64_t& nmax)
{
    hs.resize(ns);
    ntot = 0;
    nmax = 0;
    for (int q = 0; q < ns; ++q) {
        if (!sys[q].A || !sys[q].b || !sys[q].x || sys[q].n <= 0 || sys[q].lda < sys[q].n) {
            g_batch_err = "bad system " + std::to_string(q);
            return BEM_ERR_BAD_ARG;
        }
        hs[q] = BSys{(const float2*)sys[q].A, sys[q].lda, sys[q].n, ntot};
        ntot += padded(sys[q].n);
        nmax = std::max(nmax, sys[q].n);
    }
    return BEM_OK;
}

static dim3 xgather_grid(const XEng& E)
{
    return dim3((unsigned)std::min<int64_t>((E.nmax + 255) / 256, 1024), (unsigned)E.ns);
}

// b and the initial x into the concatenated space (one launch for all items;
// dvec: the items' vector table, hv its host copy, alive until the stream
// has run the copy); targets tol |b_s|
static int gather_systems(const bem_batch_system* sys, const std::vector<BSys>& hs, const XEng& E, float tol,
                          std::vector<double>& hb, XVec* dvec, std::vector<XVec>& hv, hipStream_t s)
{
    const int ns = E.ns;
    hv.resize(ns);
    for (int q = 0; q < ns; ++q) hv[q] = XVec{(const float2*)sys[q].b, (float2*)sys[q].x};
    XTRY(hipMemcpyAsync((void*)E.dsys, hs.data(), sizeof(BSys) * ns, hipMemcpyHostToDevice, s));
    XTRY(hipMemcpyAsync(dvec, hv.data(), sizeof(XVec) * ns, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(xgather_kernel, xgather_grid(E), dim3(256), 0, s, E.dsys, (const XVec*)dvec, E.B, E.X, false);
    XTRY(hipGetLastError());
    hb.assign(ns, 0.0);
    std::vector<double> ht(ns);
    hipLaunchKernelGGL(xnorm2_kernel, dim3(ns), dim3(256), 0, s, E.dsys, E.B, E.rn2);
    XTRY(hipGetLastError());
    XTRY(hipMemcpyAsync(hb.data(), E.rn2, sizeof(double) * ns, hipMemcpyDeviceToHost, s));
    XTRY(hipStreamSynchronize(s));
    for (int q = 0; q < ns; ++q) ht[q] = (double)tol * std::sqrt(hb[q]);
    XTRY(hipMemcpyAsync(E.tgt, ht.data(), sizeof(double) * ns, hipMemcpyHostToDevice, s));
    return BEM_OK;
}

int bem_batch_gmres(const bem_batch_system* sys, int ns, float tol, int restart, int max_cycles, int use_graph,
                    void* work, bem_solve_info* info, void* stream)
{
    if (!sys || ns <= 0 || restart <= 0 || max_cycles < 0 || !work || !(tol > 0.f)) return BEM_ERR_BAD_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    std::vector<BSys> hs;
    int64_t ntot = 0, nmax = 0;
    if (int rc = check_systems(sys, ns, hs, ntot, nmax)) return rc;
    const XLayout L = xlayout(ns, ntot, restart);
    char* W = (char*)work;
    const XEng E = xeng_at(W, L, (const BSys*)(W + L.sys), ns, ntot, nmax, restart);
    std::vector<double> hb, hr(ns);
    std::vector<XVec> hv;
    XVec* dvec = (XVec*)(W + L.vec);
    if (int rc = gather_systems(sys, hs, E, tol, hb, dvec, hv, s)) return rc;
    XGraphs G;  // each part of a cycle captured on its third use (xrun_slot)
    std::vector<int> hc(ns, 0), hk(ns, 0), hd(ns, 0), its(ns, 0);
    bool all = false;
    for (int cycles = 0; cycles < max_cycles && !all; ++cycles) {
        XTRY(xrun_cycle(E, use_graph ? &G : nullptr, s, hd));
        XTRY(hipMemcpyAsync(hc.data(), E.conv, sizeof(int) * ns, hipMemcpyDeviceToHost, s));
        XTRY(hipMemcpyAsync(hr.data(), E.rn2, sizeof(double) * ns, hipMemcpyDeviceToHost, s));
        XTRY(hipMemcpyAsync(hk.data(), E.nk, sizeof(int) * ns, hipMemcpyDeviceToHost, s));
        XTRY(hipStreamSynchronize(s));
        all = true;
        for (int q = 0; q < ns; ++q) {
            all = all && hc[q];
            its[q] += hk[q];  // Arnoldi steps of this item in the cycle (0 once converged)
        }
    }
    hipLaunchKernelGGL(xgather_kernel, xgather_grid(E), dim3(256), 0, s, E.dsys, (const XVec*)dvec, E.B, E.X, true);
    XTRY(hipGetLastError());
    if (info)
        for (int q = 0; q < ns; ++q) {
            info[q].iterations = its[q];
            info[q].converged = hc[q];
            info[q].rel_residual = hb[q] > 0 ? std::sqrt(hr[q] / hb[q]) : 0.0;
            info[q].implicit_rel_residual = 0.0;
        }
    if (!all) {
        g_batch_err = "batched GMRES reached max_cycles";
        return BEM_ERR_NOT_CONVERGED;
    }
    return BEM_OK;
}

// ---- batched mixed-precision solve ---------------------------------------------
int64_t bem_batch_fgmres_work_size(const bem_batch_system* sys, int ns, int outer_restart, int inner_restart)
{
    if (!sys || ns <= 0 || outer_restart <= 0 || inner_restart <= 0) return 0;
    int64_t ntot = 0;
    for (int q = 0; q < ns; ++q) ntot += padded(sys[q].n);
    return (int64_t)flayout(sys, ns, ntot, outer_restart, inner_restart).total;
}

int bem_batch_fgmres(const bem_batch_system* sys, int ns, int dtype, float tol, float inner_tol, int outer_restart,
                     int inner_restart, int max_inner_cycles, int max_outer_cycles, int use_graph, void* work,
                     bem_solve_info* info, int* inner_iterations, void* stream)
{
    if (!sys || ns <= 0 || outer_restart <= 0 || inner_restart <= 0 || max_inner_cycles <= 0 ||
        max_outer_cycles < 0 || !work ||
        !(tol > 0.f) || !(inner_tol > 0.f) || (dtype != kDtBF16 && dtype != kDtFP16 && dtype != kDtFP8))
        return BEM_ERR_BAD_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int mo = outer_restart, mi = inner_restart;
    std::vector<BSys> hs;
    int64_t ntot = 0, nmax = 0;
    if (int rc = check_systems(sys, ns, hs, ntot, nmax)) return rc;
    const FLayout F = flayout(sys, ns, ntot, mo, mi);
    char* W = (char*)work;
    const BSys* dsys = (const BSys*)(W + F.o_at + F.o.sys);
    const XEng Eo = xeng_at(W + F.o_at, F.o, dsys, ns, ntot, nmax, mo);  // FP32 flexible GMRES on A
    XEng Ei = xeng_at(W + F.in_at, F.in, dsys, ns, ntot, nmax, mi);      // 16-bit GMRES per direction
    float2* Z = (float2*)(W + F.z_at);
    BLow* dlow = (BLow*)(W + F.low_at);
    float* scale = (float*)(W + F.scale_at);
    float* inv = (float*)(W + F.inv_at);
    unsigned int* amax = (unsigned int*)(W + F.amax_at);
    Ei.dlow = dlow;
    Ei.dt = dtype;
    Ei.inv_scale = inv;
    Ei.skip = Eo.done;
    // the inner solve is a preconditioner (the outer loop forms true FP32
    // residuals): one Gram-Schmidt pass, the second-pass coefficients kept
    // zero (3-4 % shorter on the batching solves, the same inner step counts)
    Ei.one_pass = true;
    hipLaunchKernelGGL(xfill_kernel, dim3((unsigned)((4 * (size_t)ns * (mi + 1) + 255) / 256)), dim3(256), 0, s,
                       (unsigned int*)Ei.h2, (int64_t)4 * ns * (mi + 1), 0u);

    std::vector<double> hb, hr(ns);
    std::vector<XVec> hv;
    XVec* dvec = (XVec*)(W + F.o_at + F.o.vec);
    if (int rc = gather_systems(sys, hs, Eo, tol, hb, dvec, hv, s)) return rc;
    // 16-bit copies of the matrices
    std::vector<BLow> hl(ns);
    float2* diag = (float2*)(W + F.diag_at);
    float* rinv = (float*)(W + F.rinv_at);
    for (int q = 0; q < ns; ++q) {
        // 16-bit: one word per entry; fp8: two entries per word (the plane
        // sized for 16 bits is half used)
        const int64_t ldl = dtype == kDtFP8 ? padded(sys[q].n) / 2 : padded(sys[q].n);
        hl[q] = BLow{(uint32_t*)(W + F.plane_at[q]), ldl, diag + hs[q].off, rinv + hs[q].off};
    }
    XTRY(hipMemcpyAsync(dlow, hl.data(), sizeof(BLow) * ns, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(xfill_kernel, dim3((ns + 255) / 256), dim3(256), 0, s, amax, (int64_t)ns, 0u);
    const dim3 grow((unsigned)std::min<int64_t>(nmax, 2048), (unsigned)ns);
    if (dtype == kDtFP16) hipLaunchKernelGGL(xamax_kernel, grow, dim3(256), 0, s, dsys, amax);
    hipLaunchKernelGGL(xscale_kernel, dim3((ns + 63) / 64), dim3(64), 0, s, amax, dtype, scale, inv, ns);
    if (dtype == kDtFP8)
        hipLaunchKernelGGL(xto_fp8_kernel, grow, dim3(256), 0, s, dsys, dlow);
    else if (dtype == kDtBF16)
        hipLaunchKernelGGL(xto_lowp_kernel<kDtBF16>, grow, dim3(256), 0, s, dsys, dlow, scale);
    else
        hipLaunchKernelGGL(xto_lowp_kernel<kDtFP16>, grow, dim3(256), 0, s, dsys, dlow, scale);
    XTRY(hipGetLastError());

    XGraphs G;  // the inner solves' graphs, shared by every outer step: each part captured on its third use
    const dim3 gx((unsigned)((nmax + 255) / 256), (unsigned)ns);
    const dim3 gc((unsigned)((nmax + 63) / 64), (unsigned)ns);
    std::vector<int> hc(ns, 0), hd(ns, 0), hdi(ns, 0), hci(ns, 0), hko(ns, 0), hki(ns, 0), its_o(ns, 0),
        its_i(ns, 0);
    bool all = false;
    for (int oc = 0; oc < max_outer_cycles && !all; ++oc) {
        if (oc == 0) XTRY(xresidual(Eo, s));  // later cycles start from the previous cycle's true residual
        XTRY(launch_binit(Eo.g, mo, Eo.rn2, Eo.tgt, Eo.done, Eo.nk, ns, s));
        hipLaunchKernelGGL(xnormalize_kernel, gx, dim3(256), 0, s, dsys, Eo.r, Eo.rn2, Eo.V, (const int*)nullptr);
        for (int j = 0; j < mo; ++j) {
            // z_j ~ A^-1 v_j: 16-bit GMRES from zero to inner_tol (at most max_inner_cycles cycles)
            float2* zj = Z + (int64_t)j * ntot;
            hipLaunchKernelGGL(xcopy_kernel, dim3((unsigned)((ntot + 255) / 256)), dim3(256), 0, s,
                               Eo.V + (int64_t)j * ntot, Ei.B, ntot);
            hipLaunchKernelGGL(xfill_kernel, dim3((unsigned)((2 * ntot + 255) / 256)), dim3(256), 0, s,
                               (unsigned int*)Ei.X, 2 * ntot, 0u);
            hipLaunchKernelGGL(xinner_tgt_kernel, dim3((ns + 63) / 64), dim3(64), 0, s, (const int*)Eo.done,
                               (const double*)Eo.tgt, (const double*)(j == 0 ? Eo.rn2 : Eo.res), j == 0 ? 1 : 0,
                               (double)inner_tol, Ei.tgt, ns);
            XTRY(hipGetLastError());
            for (int ic = 0; ic < max_inner_cycles; ++ic) {
                XTRY(xrun_cycle(Ei, use_graph ? &G : nullptr, s, hdi));
                XTRY(hipMemcpyAsync(hki.data(), Ei.nk, sizeof(int) * ns, hipMemcpyDeviceToHost, s));
                XTRY(hipMemcpyAsync(hci.data(), Ei.conv, sizeof(int) * ns, hipMemcpyDeviceToHost, s));
                XTRY(hipStreamSynchronize(s));
                bool iall = true;
                for (int q = 0; q < ns; ++q) {
                    its_i[q] += hki[q];
                    iall = iall && hci[q];
                }
                if (iall) break;
            }
            hipLaunchKernelGGL(xcopy_kernel, dim3((unsigned)((ntot + 255) / 256)), dim3(256), 0, s, Ei.X, zj, ntot);
            // w = A z_j in FP32, orthogonalised against the FP32 basis
            XTRY(xmatvec(Eo, zj, Eo.w, kOne, kZero, 1, Eo.done, s));
            hipLaunchKernelGGL(xmultidot_kernel, dim3(j + 1, ns), dim3(256), 0, s, dsys, Eo.V, ntot, Eo.w, Eo.h,
                               mo + 1, (const int*)Eo.done);
            hipLaunchKernelGGL(xcombine_kernel, gc, dim3(64 * kXWaves), 0, s, dsys, Eo.V, ntot, j + 1,
                               (const int*)nullptr, Eo.h, mo + 1, -1.0, Eo.w, (const int*)Eo.done);
            hipLaunchKernelGGL(xmultidot_kernel, dim3(j + 1, ns), dim3(256), 0, s, dsys, Eo.V, ntot, Eo.w, Eo.h2,
                               mo + 1, (const int*)Eo.done);
            hipLaunchKernelGGL(xcombine_kernel, gc, dim3(64 * kXWaves), 0, s, dsys, Eo.V, ntot, j + 1,
                               (const int*)nullptr, Eo.h2, mo + 1, -1.0, Eo.w, (const int*)Eo.done);
            hipLaunchKernelGGL(xnorm2_kernel, dim3(ns), dim3(256), 0, s, dsys, Eo.w, Eo.wn2);
            hipLaunchKernelGGL(xnormalize_kernel, gx, dim3(256), 0, s, dsys, Eo.w, Eo.wn2,
                               Eo.V + (int64_t)(j + 1) * ntot, (const int*)Eo.done);
            XTRY(hipGetLastError());
            XTRY(launch_bgivens(Eo.H, mo, j, Eo.h, Eo.h2, mo + 1, Eo.wn2, Eo.cs, Eo.sn, Eo.g, Eo.res, Eo.tgt,
                                Eo.done, Eo.nk, ns, s));
            XTRY(hipMemcpyAsync(hd.data(), Eo.done, sizeof(int) * ns, hipMemcpyDeviceToHost, s));
            XTRY(hipStreamSynchronize(s));
            bool alldone = true;
            for (int q = 0; q < ns; ++q) alldone = alldone && hd[q];
            if (alldone) break;
        }
        // x += Z y, then the true FP32 residual
        XTRY(launch_bbacksolve(Eo.H, mo, Eo.nk, Eo.g, Eo.y, ns, s));
        hipLaunchKernelGGL(xcombine_kernel, gc, dim3(64 * kXWaves), 0, s, dsys, Z, ntot, 0, (const int*)Eo.nk, Eo.y,
                           mo, 1.0, Eo.X, (const int*)nullptr);
        XTRY(hipGetLastError());
        XTRY(xresidual(Eo, s));
        hipLaunchKernelGGL(xconv_kernel, dim3((ns + 63) / 64), dim3(64), 0, s, Eo.rn2, Eo.tgt, Eo.conv, ns);
        XTRY(hipGetLastError());
        XTRY(hipMemcpyAsync(hc.data(), Eo.conv, sizeof(int) * ns, hipMemcpyDeviceToHost, s));
        XTRY(hipMemcpyAsync(hr.data(), Eo.rn2, sizeof(double) * ns, hipMemcpyDeviceToHost, s));
        XTRY(hipMemcpyAsync(hko.data(), Eo.nk, sizeof(int) * ns, hipMemcpyDeviceToHost, s));
        XTRY(hipStreamSynchronize(s));
        all = true;
        for (int q = 0; q < ns; ++q) {
            all = all && hc[q];
            its_o[q] += hko[q];
        }
    }
    hipLaunchKernelGGL(xgather_kernel, xgather_grid(Eo), dim3(256), 0, s, Eo.dsys, (const XVec*)dvec, Eo.B, Eo.X,
                       true);
    XTRY(hipGetLastError());
    if (info)
        for (int q = 0; q < ns; ++q) {
            info[q].iterations = its_o[q];
            info[q].converged = hc[q];
            info[q].rel_residual = hb[q] > 0 ? std::sqrt(hr[q] / hb[q]) : 0.0;
            info[q].implicit_rel_residual = 0.0;
        }
    if (inner_iterations)
        for (int q = 0; q < ns; ++q) inner_iterations[q] = its_i[q];
    if (!all) {
        g_batch_err = "batched mixed-precision solve reached max_outer_cycles";
        return BEM_ERR_NOT_CONVERGED;
    }
    return BEM_OK;
}

}  // extern "C"
