// Dense complex linear algebra for the BEM solve on gfx950.
//
//  * gemv_n: y = alpha A x + beta y for a row-major complex64 matrix; one
//    wavefront streams kGemvRows rows with 16-byte loads, sharing each x
//    fetch between its rows; HBM-bound.
//  * gemv_t: y = alpha op(A) x + beta y with op = transpose or conjugate
//    transpose; each lane owns two columns and a workgroup a row slab, with a
//    second pass summing the slabs (fixed order, no atomics).
//  * vector kernels for GMRES (classical Gram-Schmidt with one
//    re-orthogonalisation, Givens rotations on the device) and for the
//    reference-compatible BiCGSTAB (src/bem/BiCGSTAB.py:5-107).
#include <hip/hip_runtime.h>

#include "bem_device.hpp"
#include "bem_linalg.hpp"

#include <algorithm>

namespace bem {

typedef float v4f __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// GEMV, no transpose
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wave_sum_f(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double block_sum_d(double v, double* sh)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    const int nw = blockDim.x >> 6;
    for (int q = 0; q < nw; ++q) t += sh[q];
    __syncthreads();
    return t;
}

template <int RW>
__global__ __launch_bounds__(256) void gemv_n_vec_kernel(const float4* __restrict__ A, int64_t rows, int64_t cols2,
                                                         int64_t lda2, const float4* __restrict__ x,
                                                         float2* __restrict__ y, float2 alpha, float2 beta,
                                                         int beta_zero)
{
    // cols2 = cols / 2, lda2 = lda / 2 (both even)
    const int lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * RW;
    if (r0 >= rows) return;
    float ar[RW], ai[RW];
#pragma unroll
    for (int q = 0; q < RW; ++q) ar[q] = ai[q] = 0.f;
    const v4f* rowp[RW];
#pragma unroll
    for (int q = 0; q < RW; ++q) rowp[q] = reinterpret_cast<const v4f*>(A + min(r0 + q, rows - 1) * lda2);
    for (int64_t c = lane; c < cols2; c += 64) {
        const float4 xv = x[c];
#pragma unroll
        for (int q = 0; q < RW; ++q) {
            const v4f a = __builtin_nontemporal_load(&rowp[q][c]);
            ar[q] = fmaf(a.x, xv.x, ar[q]);
            ar[q] = fmaf(-a.y, xv.y, ar[q]);
            ai[q] = fmaf(a.x, xv.y, ai[q]);
            ai[q] = fmaf(a.y, xv.x, ai[q]);
            ar[q] = fmaf(a.z, xv.z, ar[q]);
            ar[q] = fmaf(-a.w, xv.w, ar[q]);
            ai[q] = fmaf(a.z, xv.w, ai[q]);
            ai[q] = fmaf(a.w, xv.z, ai[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < RW; ++q) {
        const float sr = wave_sum_f(ar[q]), si = wave_sum_f(ai[q]);
        const int64_t r = r0 + q;
        if (lane == q && r < rows) {
            float2 out = make_float2(alpha.x * sr - alpha.y * si, alpha.x * si + alpha.y * sr);
            if (!beta_zero) {
                const float2 yo = y[r];
                out.x += beta.x * yo.x - beta.y * yo.y;
                out.y += beta.x * yo.y + beta.y * yo.x;
            }
            y[r] = out;
        }
    }
}

// generic (any ld / alignment / odd cols) fallback
__global__ __launch_bounds__(256) void gemv_n_scalar_kernel(const float2* __restrict__ A, int64_t rows, int64_t cols,
                                                            int64_t lda, const float2* __restrict__ x,
                                                            float2* __restrict__ y, float2 alpha, float2 beta,
                                                            int beta_zero)
{
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= rows) return;
    float ar = 0.f, ai = 0.f;
    const float2* row = A + r * lda;
    for (int64_t c = lane; c < cols; c += 64) {
        const float2 a = row[c], xv = x[c];
        ar = fmaf(a.x, xv.x, fmaf(-a.y, xv.y, ar));
        ai = fmaf(a.x, xv.y, fmaf(a.y, xv.x, ai));
    }
    const float sr = wave_sum_f(ar), si = wave_sum_f(ai);
    if (lane == 0) {
        float2 out = make_float2(alpha.x * sr - alpha.y * si, alpha.x * si + alpha.y * sr);
        if (!beta_zero) {
            const float2 yo = y[r];
            out.x += beta.x * yo.x - beta.y * yo.y;
            out.y += beta.x * yo.y + beta.y * yo.x;
        }
        y[r] = out;
    }
}

// ---------------------------------------------------------------------------
// GEMV, transpose / conjugate transpose: partial sums per row slab
// ---------------------------------------------------------------------------
constexpr int kTSlabRows = 256;

__global__ __launch_bounds__(256) void gemv_t_partial_kernel(const float2* __restrict__ A, int64_t rows, int64_t cols,
                                                             int64_t lda, const float2* __restrict__ x, int conj,
                                                             float2* __restrict__ part)
{
    // grid.x: column blocks of 256, grid.y: row slabs
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.y * kTSlabRows;
    const int64_t r1 = min(r0 + (int64_t)kTSlabRows, rows);
    if (c >= cols) return;
    float ar = 0.f, ai = 0.f;
    const float sgn = conj ? -1.f : 1.f;
    for (int64_t r = r0; r < r1; ++r) {
        const float2 a = A[r * lda + c];
        const float2 xv = x[r];  // uniform
        const float aim = sgn * a.y;
        ar = fmaf(a.x, xv.x, fmaf(-aim, xv.y, ar));
        ai = fmaf(a.x, xv.y, fmaf(aim, xv.x, ai));
    }
    part[(int64_t)blockIdx.y * cols + c] = make_float2(ar, ai);
}

// y = op(A)^T-style product, vector form: a workgroup owns 128 columns (two
// per lane, 16-byte loads: each wavefront reads a contiguous 1 KB of a row) and
// a slab of kTVecSlab rows, split over its 4 wavefronts (rows w, w+4, ...);
// the 4 wave partials meet in LDS in a fixed order.  Needs A and lda such that
// every row starts 16-byte aligned (checked by the launcher).
constexpr int kTVecSlab = 128;
__global__ __launch_bounds__(256) void gemv_t_vec_partial_kernel(const float2* __restrict__ A, int64_t rows,
                                                                 int64_t cols, int64_t lda,
                                                                 const float2* __restrict__ x, int conj,
                                                                 float2* __restrict__ part)
{
    __shared__ float4 red[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t c = (int64_t)blockIdx.x * 128 + 2 * lane;
    const int64_t r0 = (int64_t)blockIdx.y * kTVecSlab;
    const int64_t r1 = min(r0 + (int64_t)kTVecSlab, rows);
    const float sgn = conj ? -1.f : 1.f;
    float a0r = 0.f, a0i = 0.f, a1r = 0.f, a1i = 0.f;
    if (c + 1 < cols) {
#pragma unroll 4
        for (int64_t r = r0 + wave; r < r1; r += 4) {
            const float4 a = *reinterpret_cast<const float4*>(A + r * lda + c);
            const float2 xv = x[r];
            const float b0 = sgn * a.y, b1 = sgn * a.w;
            a0r = fmaf(a.x, xv.x, fmaf(-b0, xv.y, a0r));
            a0i = fmaf(a.x, xv.y, fmaf(b0, xv.x, a0i));
            a1r = fmaf(a.z, xv.x, fmaf(-b1, xv.y, a1r));
            a1i = fmaf(a.z, xv.y, fmaf(b1, xv.x, a1i));
        }
    } else if (c < cols) {
        for (int64_t r = r0 + wave; r < r1; r += 4) {
            const float2 a = A[r * lda + c];
            const float2 xv = x[r];
            const float b0 = sgn * a.y;
            a0r = fmaf(a.x, xv.x, fmaf(-b0, xv.y, a0r));
            a0i = fmaf(a.x, xv.y, fmaf(b0, xv.x, a0i));
        }
    }
    red[wave][lane] = make_float4(a0r, a0i, a1r, a1i);
    __syncthreads();
    if (wave == 0 && c < cols) {
        float4 t = red[0][lane];
#pragma unroll
        for (int w = 1; w < 4; ++w) {
            const float4 u = red[w][lane];
            t.x += u.x;
            t.y += u.y;
            t.z += u.z;
            t.w += u.w;
        }
        float2* out = part + (int64_t)blockIdx.y * cols + c;
        out[0] = make_float2(t.x, t.y);
        if (c + 1 < cols) out[1] = make_float2(t.z, t.w);
    }
}

__global__ __launch_bounds__(256) void gemv_t_reduce_kernel(const float2* __restrict__ part, int64_t nslab,
                                                            int64_t cols, float2* __restrict__ y, float2 alpha,
                                                            float2 beta, int beta_zero)
{
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= cols) return;
    double sr = 0.0, si = 0.0;
    for (int64_t s = 0; s < nslab; ++s) {
        const float2 p = part[s * cols + c];
        sr += p.x;
        si += p.y;
    }
    float2 out = make_float2((float)(alpha.x * sr - alpha.y * si), (float)(alpha.x * si + alpha.y * sr));
    if (!beta_zero) {
        const float2 yo = y[c];
        out.x += beta.x * yo.x - beta.y * yo.y;
        out.y += beta.x * yo.y + beta.y * yo.x;
    }
    y[c] = out;
}

// ---------------------------------------------------------------------------
// dense combine: C = a A + b B (either input may be null = 0)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void combine_kernel(const float2* __restrict__ A, int64_t lda,
                                                      const float2* __restrict__ B, int64_t ldb, float2* __restrict__ C,
                                                      int64_t ldc, int64_t rows, int64_t cols, float2 a, float2 b)
{
    const int64_t total = rows * cols;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t r = e / cols, c = e - r * cols;
        float2 o = make_float2(0.f, 0.f);
        if (A) {
            const float2 v = A[r * lda + c];
            o.x += a.x * v.x - a.y * v.y;
            o.y += a.x * v.y + a.y * v.x;
        }
        if (B) {
            const float2 v = B[r * ldb + c];
            o.x += b.x * v.x - b.y * v.y;
            o.y += b.x * v.y + b.y * v.x;
        }
        C[r * ldc + c] = o;
    }
}

// ---------------------------------------------------------------------------
// vector kernels (GMRES / BiCGSTAB)
// ---------------------------------------------------------------------------
// h[q] = sum_e conj(V[q][e]) w[e] for q < nq, one workgroup per q (double
// accumulation, fixed order).  Result written as double2 to hout[q].
__global__ __launch_bounds__(256) void multidot_kernel(const float2* __restrict__ V, int64_t ldv, int64_t n,
                                                       const float2* __restrict__ w, double2* __restrict__ hout)
{
    __shared__ double sh[8];
    const int q = blockIdx.x;
    const float2* v = V + (int64_t)q * ldv;
    double sr = 0.0, si = 0.0;
    for (int64_t e = threadIdx.x; e < n; e += 256) {
        const float2 a = v[e], b = w[e];
        sr += (double)a.x * b.x + (double)a.y * b.y;
        si += (double)a.x * b.y - (double)a.y * b.x;
    }
    sr = block_sum_d(sr, sh);
    si = block_sum_d(si, sh);
    if (threadIdx.x == 0) hout[q] = make_double2(sr, si);
}

// <V_q, w> over element range s of S (grid (nq, S)), 16-byte loads; rows of V
// and w must be 16-byte aligned and n even (checked by the launcher).  Partial
// sums part[s * nq + q], combined in a fixed order by multidot_reduce_kernel:
// a vector split over S workgroups keeps enough loads in flight when the basis
// is short, and the whole chip busy when it is long.
constexpr int kDotMaxSplit = 16;
__global__ __launch_bounds__(256) void multidot_split_kernel(const float2* __restrict__ V, int64_t ldv, int64_t n,
                                                             const float2* __restrict__ w, int nq,
                                                             double2* __restrict__ part)
{
    __shared__ double sh[8];
    const int q = blockIdx.x, sp = blockIdx.y, ns = gridDim.y;
    const int64_t n2 = n / 2;
    const int64_t chunk = (n2 + ns - 1) / ns;
    const int64_t e0 = (int64_t)sp * chunk, e1 = min(n2, e0 + chunk);
    const float4* v = reinterpret_cast<const float4*>(V + (int64_t)q * ldv);
    const float4* w4 = reinterpret_cast<const float4*>(w);
    double sr = 0.0, si = 0.0;
#pragma unroll 4
    for (int64_t e = e0 + threadIdx.x; e < e1; e += 256) {
        const float4 a = v[e], b = w4[e];
        sr += (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z + (double)a.w * b.w;
        si += (double)a.x * b.y - (double)a.y * b.x + (double)a.z * b.w - (double)a.w * b.z;
    }
    sr = block_sum_d(sr, sh);
    si = block_sum_d(si, sh);
    if (threadIdx.x == 0) part[(int64_t)sp * nq + q] = make_double2(sr, si);
}
__global__ __launch_bounds__(256) void multidot_reduce_kernel(const double2* __restrict__ part, int ns, int nq,
                                                              double2* __restrict__ hout)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= nq) return;
    double sr = 0.0, si = 0.0;
    for (int sp = 0; sp < ns; ++sp) {
        const double2 p = part[(int64_t)sp * nq + q];
        sr += p.x;
        si += p.y;
    }
    hout[q] = make_double2(sr, si);
}

// out_s[e] += sign * sum_{q < nq_s} c_s[q] V_s[q][e] for each system s (grid y):
// V_s[q] at V + q qstride + s sstride, c_s at c + s cstride, out_s at
// out + s ostride; nq_s = nq_arr[s] (or nq_fixed); systems with done[s] set are
// skipped.  This is the Gram-Schmidt update (sign -1) and the solution update
// (sign +1) of the single- and multi-system GMRES.  A workgroup owns 64
// consecutive elements; its 8 wavefronts take every 8th basis vector, so a long
// Krylov basis keeps 8x more loads in flight than one thread per element, and
// the 8 FP64 partial sums meet in LDS in a fixed order (deterministic).
constexpr int kBasisWaves = 8;
__global__ __launch_bounds__(64 * kBasisWaves) void basis_combine_kernel(
    const float2* __restrict__ V, int64_t qstride, int64_t sstride, int64_t n, int nq_fixed,
    const int* __restrict__ nq_arr, const double2* __restrict__ c, int64_t cstride, double sign,
    float2* __restrict__ out, int64_t ostride, const int* __restrict__ done)
{
    const int sy = blockIdx.y;
    if (done && done[sy]) return;  // uniform over the workgroup
    __shared__ double2 part[kBasisWaves][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t e = (int64_t)blockIdx.x * 64 + lane;
    const int nq = nq_arr ? nq_arr[sy] : nq_fixed;
    const float2* Vs = V + (int64_t)sy * sstride;
    const double2* cs = c + (int64_t)sy * cstride;
    double ar = 0.0, ai = 0.0;
    if (e < n) {
        for (int q = wv; q < nq; q += kBasisWaves) {
            const double2 cq = cs[q];
            const float2 v = Vs[(int64_t)q * qstride + e];
            ar += cq.x * v.x - cq.y * v.y;
            ai += cq.x * v.y + cq.y * v.x;
        }
    }
    part[wv][lane] = make_double2(ar, ai);
    __syncthreads();
    if (wv == 0 && e < n) {
        double sr = 0.0, si = 0.0;
        for (int q = 0; q < kBasisWaves; ++q) {
            sr += part[q][lane].x;
            si += part[q][lane].y;
        }
        float2* o = out + (int64_t)sy * ostride;
        const float2 ov = o[e];
        o[e] = make_float2((float)((double)ov.x + sign * sr), (float)((double)ov.y + sign * si));
    }
}

// ||w||^2 into out[0] (double), one workgroup of 1024 threads
__global__ __launch_bounds__(1024) void norm2_kernel(const float2* __restrict__ w, int64_t n, double* __restrict__ out)
{
    __shared__ double sh[16];
    double s = 0.0;
    for (int64_t e = threadIdx.x; e < n; e += 1024) {
        const float2 a = w[e];
        s += (double)a.x * a.x + (double)a.y * a.y;
    }
    s = block_sum_d(s, sh);
    if (threadIdx.x == 0) out[0] = s;
}

// dot(a, b) = sum a b (bilinear, conj = 0) or sum conj(a) b (conj = 1) into out (double2)
__global__ __launch_bounds__(1024) void dot_kernel(const float2* __restrict__ a, const float2* __restrict__ b,
                                                   int64_t n, int conj, double2* __restrict__ out)
{
    __shared__ double sh[16];
    double sr = 0.0, si = 0.0;
    const double sg = conj ? -1.0 : 1.0;
    for (int64_t e = threadIdx.x; e < n; e += 1024) {
        const float2 x = a[e], y = b[e];
        sr += (double)x.x * y.x - sg * (double)x.y * y.y;
        si += (double)x.x * y.y + sg * (double)x.y * y.x;
    }
    sr = block_sum_d(sr, sh);
    si = block_sum_d(si, sh);
    if (threadIdx.x == 0) *out = make_double2(sr, si);
}

// out[e] = src[e] * scale, scale = 1 / sqrt(*nrm2) (or 0 if nrm2 == 0)
__global__ __launch_bounds__(256) void normalize_kernel(const float2* __restrict__ src, int64_t n,
                                                        const double* __restrict__ nrm2, float2* __restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const double nn = *nrm2;
    const float s = nn > 0.0 ? (float)(1.0 / sqrt(nn)) : 0.f;
    const float2 v = src[e];
    out[e] = make_float2(v.x * s, v.y * s);
}

// GMRES Arnoldi bookkeeping on the device (one thread):
//   column j of H gets h (length j+1, from CGS2) and ||w||; previous Givens
//   rotations are applied, a new one zeroes H[j+1][j], g is rotated and
//   |g[j+1]| is stored in res[j].  H: (m+1) x m column-major (double2).
__global__ __launch_bounds__(kGivensChunk) void gmres_givens_kernel(
    double2* __restrict__ H, int m, int j, const double2* __restrict__ h, const double2* __restrict__ h2,
    const double* __restrict__ wn2, double2* __restrict__ cs, double2* __restrict__ sn, double2* __restrict__ g,
    double* __restrict__ res)
{
    double2* col = H + (int64_t)j * (m + 1);
    const int t = threadIdx.x;
    const double2 carry = rotate_column_wg(col, j, h, h2, cs, sn);
    if (t != 0) return;
    const double2 a = carry, b = make_double2(sqrt(*wn2), 0.0);
    const double den = sqrt(a.x * a.x + a.y * a.y + b.x * b.x + b.y * b.y);
    double2 c, sv;
    if (den > 0.0) {
        c = make_double2(a.x / den, -a.y / den);   // conj(a)/den
        sv = make_double2(b.x / den, -b.y / den);  // conj(b)/den
    } else {
        c = make_double2(1.0, 0.0);
        sv = make_double2(0.0, 0.0);
    }
    cs[j] = c;
    sn[j] = sv;
    col[j] = make_double2(den, 0.0);
    col[j + 1] = make_double2(0.0, 0.0);
    const double2 gj = g[j];
    const double2 gn = make_double2(-(sv.x * gj.x + sv.y * gj.y), -(sv.x * gj.y - sv.y * gj.x));  // -conj(s) g_j
    g[j + 1] = gn;
    g[j] = make_double2(c.x * gj.x - c.y * gj.y, c.x * gj.y + c.y * gj.x);
    res[j] = sqrt(gn.x * gn.x + gn.y * gn.y);
}


// One Arnoldi step after the product w = A v_j for a small system, in one
// workgroup: classical Gram-Schmidt with one re-orthogonalisation (h, h2),
// |w|^2 into *wn2, v_{j+1} = w / |w|, and the Givens update of column j
// (as gmres_givens_kernel).  One launch instead of nine: for small systems
// the steps are launch-bound (each kernel does microseconds of work).
constexpr int kArnoldiThreads = 1024;  // 16 wavefronts share the dot products and the update
__global__ __launch_bounds__(kArnoldiThreads) void arnoldi_small_kernel(
    float2* __restrict__ V, int64_t n, int j, float2* __restrict__ w, double2* __restrict__ h,
    double2* __restrict__ h2, double2* __restrict__ H, int m, double2* __restrict__ cs, double2* __restrict__ sn,
    double2* __restrict__ g, double* __restrict__ res, double* __restrict__ wn2)
{
    __shared__ double red[kArnoldiThreads / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    constexpr int nw = kArnoldiThreads / 64;
    for (int pass = 0; pass < 2; ++pass) {
        double2* hh = pass ? h2 : h;
        for (int q = wave; q <= j; q += nw) {  // hh[q] = <v_q, w>
            const float2* vq = V + (int64_t)q * n;
            double sr = 0.0, si = 0.0;
            // eight iterations' loads issued before their sums (clamped
            // addresses, masked values): issued one at a time, each waited
            // for before the next, the step took ~95 us at n = 2562
            for (int64_t e0 = lane; e0 < n; e0 += 64 * 8) {
                float2 a[8], b[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int64_t e = min(e0 + 64 * u, n - 1);
                    a[u] = vq[e];
                    b[u] = w[e];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (e0 + 64 * u < n) {
                        sr += (double)a[u].x * b[u].x + (double)a[u].y * b[u].y;
                        si += (double)a[u].x * b[u].y - (double)a[u].y * b[u].x;
                    }
                }
            }
            sr = wave_sum64(sr);
            si = wave_sum64(si);
            if (lane == 0) hh[q] = make_double2(sr, si);
        }
        __syncthreads();
        for (int64_t e = t; e < n; e += kArnoldiThreads) {  // w -= sum_q hh[q] v_q
            double ar = 0.0, ai = 0.0;
            for (int q0 = 0; q0 <= j; q0 += 8) {  // eight basis vectors' loads at once
                float2 v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = V[(int64_t)min(q0 + u, j) * n + e];
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (q0 + u <= j) {
                        const double2 c = hh[q0 + u];
                        ar += c.x * v[u].x - c.y * v[u].y;
                        ai += c.x * v[u].y + c.y * v[u].x;
                    }
                }
            }
            const float2 wo = w[e];
            w[e] = make_float2((float)((double)wo.x - ar), (float)((double)wo.y - ai));
        }
        __syncthreads();
    }
    double a2 = 0.0;
    for (int64_t e = t; e < n; e += kArnoldiThreads) {
        const float2 v = w[e];
        a2 += (double)v.x * v.x + (double)v.y * v.y;
    }
    a2 = wave_sum64(a2);
    if (lane == 0) red[wave] = a2;
    __syncthreads();
    double nrm2 = 0.0;
    for (int q = 0; q < nw; ++q) nrm2 += red[q];
    const float inv = nrm2 > 0.0 ? (float)(1.0 / sqrt(nrm2)) : 0.f;
    float2* vn = V + (int64_t)(j + 1) * n;
    for (int64_t e = t; e < n; e += kArnoldiThreads) {
        const float2 v = w[e];
        vn[e] = make_float2(v.x * inv, v.y * inv);
    }
    double2* col = H + (int64_t)j * (m + 1);
    const double2 carry = rotate_column_wg(col, j, h, h2, cs, sn);
    if (t != 0) return;
    *wn2 = nrm2;
    const double2 aa = carry, bb = make_double2(sqrt(nrm2), 0.0);
    const double den = sqrt(aa.x * aa.x + aa.y * aa.y + bb.x * bb.x + bb.y * bb.y);
    double2 c = make_double2(1.0, 0.0), sv = make_double2(0.0, 0.0);
    if (den > 0.0) {
        c = make_double2(aa.x / den, -aa.y / den);
        sv = make_double2(bb.x / den, -bb.y / den);
    }
    cs[j] = c;
    sn[j] = sv;
    col[j] = make_double2(den, 0.0);
    col[j + 1] = make_double2(0.0, 0.0);
    const double2 gj = g[j];
    const double2 gn = make_double2(-(sv.x * gj.x + sv.y * gj.y), -(sv.x * gj.y - sv.y * gj.x));
    g[j + 1] = gn;
    g[j] = make_double2(c.x * gj.x - c.y * gj.y, c.x * gj.y + c.y * gj.x);
    res[j] = sqrt(gn.x * gn.x + gn.y * gn.y);
}

// g = (beta, 0, ...), beta = sqrt(*nrm2)
__global__ void gmres_init_g_kernel(double2* __restrict__ g, int m, const double* __restrict__ nrm2)
{
    const int q = threadIdx.x;
    for (int e = q; e <= m; e += blockDim.x) g[e] = make_double2(e == 0 ? sqrt(*nrm2) : 0.0, 0.0);
}

// ---- BiCGSTAB helpers (scalars live on the device) ----
// sc layout (double2): 0 rho_prev, 1 alpha, 2 omega, 3 rho, 4 beta, 5 dot(r_hat, v),
//                      6 dot(t, s), 7 dot(t, t)
__global__ void bicg_beta_kernel(double2* sc)
{
    // beta = (rho / rho_prev) * (alpha / omega); rho_prev = rho
    auto cdiv = [](double2 a, double2 b) {
        const double d = b.x * b.x + b.y * b.y;
        return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
    };
    auto cmul = [](double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); };
    sc[4] = cmul(cdiv(sc[3], sc[0]), cdiv(sc[1], sc[2]));
    sc[0] = sc[3];
}
__global__ void bicg_alpha_kernel(double2* sc)
{
    const double2 a = sc[0], b = sc[5];
    const double d = b.x * b.x + b.y * b.y;
    sc[1] = make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__global__ void bicg_omega_kernel(double2* sc)
{
    const double2 a = sc[6], b = sc[7];
    const double d = b.x * b.x + b.y * b.y;
    sc[2] = make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
// p = r + beta (p - omega v)
__global__ __launch_bounds__(256) void bicg_p_kernel(const float2* __restrict__ r, float2* __restrict__ p,
                                                     const float2* __restrict__ v, const double2* __restrict__ sc,
                                                     int64_t n)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const double2 be = sc[4], om = sc[2];
    const float2 pv = p[e], vv = v[e], rv = r[e];
    const double tr = pv.x - (om.x * vv.x - om.y * vv.y), ti = pv.y - (om.x * vv.y + om.y * vv.x);
    p[e] = make_float2((float)(rv.x + be.x * tr - be.y * ti), (float)(rv.y + be.x * ti + be.y * tr));
}
// out = a + coef * b, coef = sign * sc[idx]
__global__ __launch_bounds__(256) void bicg_axpy_kernel(const float2* __restrict__ a, const float2* __restrict__ b,
                                                        const double2* __restrict__ sc, int idx, double sign,
                                                        float2* __restrict__ out, int64_t n)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const double2 c = make_double2(sign * sc[idx].x, sign * sc[idx].y);
    const float2 av = a[e], bv = b[e];
    out[e] = make_float2((float)(av.x + c.x * bv.x - c.y * bv.y), (float)(av.y + c.x * bv.y + c.y * bv.x));
}
// x += alpha p + omega s
__global__ __launch_bounds__(256) void bicg_x_kernel(float2* __restrict__ x, const float2* __restrict__ p,
                                                     const float2* __restrict__ s, const double2* __restrict__ sc,
                                                     int64_t n)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const double2 al = sc[1], om = sc[2];
    const float2 pv = p[e], sv = s[e], xv = x[e];
    x[e] = make_float2((float)(xv.x + al.x * pv.x - al.y * pv.y + om.x * sv.x - om.y * sv.y),
                       (float)(xv.y + al.x * pv.y + al.y * pv.x + om.x * sv.y + om.y * sv.x));
}

// ---------------------------------------------------------------------------
// batched Krylov kernels: ns systems in lockstep; vectors stored system-major
// (system s of basis vector q at V + (q * ns + s) * ldn)
// ---------------------------------------------------------------------------
// h[s * hs + q] = <V_q,s, w_s> for q < nq; grid (nq, ns)
__global__ __launch_bounds__(256) void bmultidot_kernel(const float2* __restrict__ V, int64_t ldn, int ns, int64_t n,
                                                        const float2* __restrict__ w, double2* __restrict__ h, int hs,
                                                        const int* __restrict__ done)
{
    __shared__ double sh[8];
    const int q = blockIdx.x, sy = blockIdx.y;
    if (done[sy]) return;  // uniform per workgroup
    const float2* v = V + ((int64_t)q * ns + sy) * ldn;
    const float2* ww = w + (int64_t)sy * ldn;
    double sr = 0.0, si = 0.0;
    if ((n & 1) == 0 && (ldn & 1) == 0 && (((uintptr_t)V | (uintptr_t)w) & 15) == 0) {
        // 16-byte loads (two entries) when every system vector is 16-byte aligned
        const float4* v4 = reinterpret_cast<const float4*>(v);
        const float4* w4 = reinterpret_cast<const float4*>(ww);
#pragma unroll 4
        for (int64_t e = threadIdx.x; e < n / 2; e += 256) {
            const float4 a = v4[e], b = w4[e];
            sr += (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z + (double)a.w * b.w;
            si += (double)a.x * b.y - (double)a.y * b.x + (double)a.z * b.w - (double)a.w * b.z;
        }
    } else {
        for (int64_t e = threadIdx.x; e < n; e += 256) {
            const float2 a = v[e], b = ww[e];
            sr += (double)a.x * b.x + (double)a.y * b.y;
            si += (double)a.x * b.y - (double)a.y * b.x;
        }
    }
    sr = block_sum_d(sr, sh);
    si = block_sum_d(si, sh);
    if (threadIdx.x == 0) h[(int64_t)sy * hs + q] = make_double2(sr, si);
}


// out[s] = ||w_s||^2 ; grid ns
__global__ __launch_bounds__(1024) void bnorm2_kernel(const float2* __restrict__ w, int64_t ldn, int64_t n,
                                                      double* __restrict__ out)
{
    __shared__ double sh[16];
    const int sy = blockIdx.x;
    const float2* ww = w + (int64_t)sy * ldn;
    double s = 0.0;
    for (int64_t e = threadIdx.x; e < n; e += 1024) {
        const float2 a = ww[e];
        s += (double)a.x * a.x + (double)a.y * a.y;
    }
    s = block_sum_d(s, sh);
    if (threadIdx.x == 0) out[sy] = s;
}

// dst_s = src_s / sqrt(nrm2[s]) ; grid (ceil(n/256), ns)
__global__ __launch_bounds__(256) void bnormalize_kernel(const float2* __restrict__ src, int64_t lds, int64_t n,
                                                         const double* __restrict__ nrm2, float2* __restrict__ dst,
                                                         int64_t ldd, const int* __restrict__ done)
{
    const int sy = blockIdx.y;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n || (done && done[sy])) return;
    const double nn = nrm2[sy];
    const float sc = nn > 0.0 ? (float)(1.0 / sqrt(nn)) : 0.f;
    const float2 v = src[(int64_t)sy * lds + e];
    dst[(int64_t)sy * ldd + e] = make_float2(v.x * sc, v.y * sc);
}

// Per-system restart set-up: g = (beta, 0...), done = (beta <= target), nk = 0.
__global__ void binit_kernel(double2* __restrict__ g, int m, const double* __restrict__ nrm2,
                             const double* __restrict__ target, int* __restrict__ done, int* __restrict__ nk, int ns)
{
    const int sy = blockIdx.x * blockDim.x + threadIdx.x;
    if (sy >= ns) return;
    const double beta = sqrt(nrm2[sy]);
    for (int e = 0; e <= m; ++e) g[(int64_t)sy * (m + 1) + e] = make_double2(e == 0 ? beta : 0.0, 0.0);
    done[sy] = beta <= target[sy] ? 1 : 0;
    nk[sy] = 0;
}

// Per-system Arnoldi bookkeeping (one workgroup per system): as
// gmres_givens_kernel, plus the convergence test |g_{j+1}| <= target -> done,
// nk = j + 1.
__global__ __launch_bounds__(kGivensChunk) void bgivens_kernel(
    double2* __restrict__ H, int m, int j, const double2* __restrict__ h, const double2* __restrict__ h2, int hs,
    const double* __restrict__ wn2, double2* __restrict__ cs, double2* __restrict__ sn, double2* __restrict__ g,
    double* __restrict__ res, const double* __restrict__ target, int* __restrict__ done, int* __restrict__ nk, int ns)
{
    const int sy = blockIdx.x;  // one workgroup per system
    if (done[sy]) return;       // uniform over the workgroup
    double2* col = H + (int64_t)sy * (m + 1) * m + (int64_t)j * (m + 1);
    double2* c_ = cs + (int64_t)sy * m;
    double2* s_ = sn + (int64_t)sy * m;
    const double2 a = rotate_column_wg(col, j, h + (int64_t)sy * hs, h2 + (int64_t)sy * hs, c_, s_);
    if (threadIdx.x != 0) return;
    double2* gg = g + (int64_t)sy * (m + 1);
    const double2 b = make_double2(sqrt(wn2[sy]), 0.0);
    const double den = sqrt(a.x * a.x + a.y * a.y + b.x * b.x + b.y * b.y);
    double2 c = make_double2(1.0, 0.0), sv = make_double2(0.0, 0.0);
    if (den > 0.0) {
        c = make_double2(a.x / den, -a.y / den);
        sv = make_double2(b.x / den, -b.y / den);
    }
    c_[j] = c;
    s_[j] = sv;
    col[j] = make_double2(den, 0.0);
    col[j + 1] = make_double2(0.0, 0.0);
    const double2 gj = gg[j];
    const double2 gn = make_double2(-(sv.x * gj.x + sv.y * gj.y), -(sv.x * gj.y - sv.y * gj.x));
    gg[j + 1] = gn;
    gg[j] = make_double2(c.x * gj.x - c.y * gj.y, c.x * gj.y + c.y * gj.x);
    const double rj = sqrt(gn.x * gn.x + gn.y * gn.y);
    res[sy] = rj;
    nk[sy] = j + 1;
    if (rj <= target[sy]) done[sy] = 1;
}

// The lockstep form of arnoldi_small_kernel: one workgroup per system (grid
// = systems), the Givens bookkeeping of bgivens_kernel.  One launch instead
// of seven per step: with small systems a step's kernels each do microseconds
// of work, so the steps are launch-bound.
__global__ __launch_bounds__(kArnoldiThreads) void barnoldi_small_kernel(
    const BSys* __restrict__ sys, float2* __restrict__ V, int64_t ntot, int j, float2* __restrict__ w,
    double2* __restrict__ h, double2* __restrict__ h2, int hs, double2* __restrict__ H, int m,
    double2* __restrict__ cs, double2* __restrict__ sn, double2* __restrict__ g, double* __restrict__ res,
    const double* __restrict__ target, int* __restrict__ done, int* __restrict__ nk, double* __restrict__ wn2,
    int passes)
{
    __shared__ double red[kArnoldiThreads / 64];
    const int sy = blockIdx.x;
    if (done[sy]) return;  // uniform over the workgroup
    const BSys S = sys[sy];
    const int64_t n = S.n;
    float2* Vs = V + S.off;  // basis vector q at Vs + q ntot
    float2* ws = w + S.off;
    double2* hsy = h + (int64_t)sy * hs;
    double2* h2sy = h2 + (int64_t)sy * hs;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    constexpr int nw = kArnoldiThreads / 64;
    for (int pass = 0; pass < passes; ++pass) {
        double2* hh = pass ? h2sy : hsy;
        for (int q = wave; q <= j; q += nw) {  // hh[q] = <v_q, w>
            const float2* vq = Vs + (int64_t)q * ntot;
            double sr = 0.0, si = 0.0;
            for (int64_t e0 = lane; e0 < n; e0 += 64 * 8) {  // eight loads in flight (clamped, masked)
                float2 a[8], b[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int64_t e = min(e0 + 64 * u, n - 1);
                    a[u] = vq[e];
                    b[u] = ws[e];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (e0 + 64 * u < n) {
                        sr += (double)a[u].x * b[u].x + (double)a[u].y * b[u].y;
                        si += (double)a[u].x * b[u].y - (double)a[u].y * b[u].x;
                    }
                }
            }
            sr = wave_sum64(sr);
            si = wave_sum64(si);
            if (lane == 0) hh[q] = make_double2(sr, si);
        }
        __syncthreads();
        for (int64_t e = t; e < n; e += kArnoldiThreads) {  // w -= sum_q hh[q] v_q
            double ar = 0.0, ai = 0.0;
            for (int q0 = 0; q0 <= j; q0 += 8) {  // eight basis vectors' loads at once
                float2 v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = Vs[(int64_t)min(q0 + u, j) * ntot + e];
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (q0 + u <= j) {
                        const double2 c = hh[q0 + u];
                        ar += c.x * v[u].x - c.y * v[u].y;
                        ai += c.x * v[u].y + c.y * v[u].x;
                    }
                }
            }
            const float2 wo = ws[e];
            ws[e] = make_float2((float)((double)wo.x - ar), (float)((double)wo.y - ai));
        }
        __syncthreads();
    }
    double a2 = 0.0;
    for (int64_t e = t; e < n; e += kArnoldiThreads) {
        const float2 v = ws[e];
        a2 += (double)v.x * v.x + (double)v.y * v.y;
    }
    a2 = wave_sum64(a2);
    if (lane == 0) red[wave] = a2;
    __syncthreads();
    double nrm2 = 0.0;
    for (int q = 0; q < nw; ++q) nrm2 += red[q];
    const float inv = nrm2 > 0.0 ? (float)(1.0 / sqrt(nrm2)) : 0.f;
    float2* vn = Vs + (int64_t)(j + 1) * ntot;
    for (int64_t e = t; e < n; e += kArnoldiThreads) {
        const float2 v = ws[e];
        vn[e] = make_float2(v.x * inv, v.y * inv);
    }
    double2* col = H + (int64_t)sy * (m + 1) * m + (int64_t)j * (m + 1);
    double2* c_ = cs + (int64_t)sy * m;
    double2* s_ = sn + (int64_t)sy * m;
    const double2 carry = rotate_column_wg(col, j, hsy, h2sy, c_, s_);
    if (t != 0) return;
    wn2[sy] = nrm2;
    double2* gg = g + (int64_t)sy * (m + 1);
    const double2 aa = carry, bb = make_double2(sqrt(nrm2), 0.0);
    const double den = sqrt(aa.x * aa.x + aa.y * aa.y + bb.x * bb.x + bb.y * bb.y);
    double2 c = make_double2(1.0, 0.0), sv = make_double2(0.0, 0.0);
    if (den > 0.0) {
        c = make_double2(aa.x / den, -aa.y / den);
        sv = make_double2(bb.x / den, -bb.y / den);
    }
    c_[j] = c;
    s_[j] = sv;
    col[j] = make_double2(den, 0.0);
    col[j + 1] = make_double2(0.0, 0.0);
    const double2 gj = gg[j];
    const double2 gn = make_double2(-(sv.x * gj.x + sv.y * gj.y), -(sv.x * gj.y - sv.y * gj.x));
    gg[j + 1] = gn;
    gg[j] = make_double2(c.x * gj.x - c.y * gj.y, c.x * gj.y + c.y * gj.x);
    const double rj = sqrt(gn.x * gn.x + gn.y * gn.y);
    res[sy] = rj;
    nk[sy] = j + 1;
    if (rj <= target[sy]) done[sy] = 1;
}

// y = R^{-1} g for one system per workgroup (R: leading nk x nk of the
// rotated Hessenberg, column c at H + c (m + 1)).  Column-oriented: once y_q
// is known every thread removes R[t][q] y_q from the right-hand side entries
// t < q (column q is contiguous), one barrier per step; the right-hand side
// lives in LDS.  Beyond kBackLds unknowns thread 0 does the row-oriented sweep.
constexpr int kBackLds = 2048;
__global__ __launch_bounds__(256) void backsolve_wg_kernel(const double2* __restrict__ H, int m, int64_t hstride,
                                                           const int* __restrict__ nk_arr, int nk_fixed,
                                                           const double2* __restrict__ g, int64_t gstride,
                                                           double2* __restrict__ y, int64_t ystride)
{
    __shared__ double2 sv[kBackLds];
    const int sy = blockIdx.x;
    const double2* Hs = H + (int64_t)sy * hstride;
    const double2* gs = g + (int64_t)sy * gstride;
    double2* ys = y + (int64_t)sy * ystride;
    const int n_k = nk_arr ? nk_arr[sy] : nk_fixed;
    const int t = threadIdx.x;
    if (n_k > kBackLds) {
        if (t != 0) return;
        for (int q = n_k - 1; q >= 0; --q) {
            double2 acc = gs[q];
            for (int c = q + 1; c < n_k; ++c) {
                const double2 hqc = Hs[(int64_t)c * (m + 1) + q], yc = ys[c];
                acc.x -= hqc.x * yc.x - hqc.y * yc.y;
                acc.y -= hqc.x * yc.y + hqc.y * yc.x;
            }
            const double2 d = Hs[(int64_t)q * (m + 1) + q];
            const double dd = d.x * d.x + d.y * d.y;
            ys[q] = dd > 0.0 ? make_double2((acc.x * d.x + acc.y * d.y) / dd, (acc.y * d.x - acc.x * d.y) / dd)
                             : make_double2(0.0, 0.0);
        }
        return;
    }
    for (int e = t; e < n_k; e += 256) sv[e] = gs[e];
    __syncthreads();
    for (int q = n_k - 1; q >= 0; --q) {
        const double2* colq = Hs + (int64_t)q * (m + 1);
        const double2 d = colq[q], a = sv[q];
        const double dd = d.x * d.x + d.y * d.y;
        const double2 yq = dd > 0.0 ? make_double2((a.x * d.x + a.y * d.y) / dd, (a.y * d.x - a.x * d.y) / dd)
                                    : make_double2(0.0, 0.0);
        if (t == 0) ys[q] = yq;
        for (int e = t; e < q; e += 256) {
            const double2 r = colq[e];
            sv[e].x -= r.x * yq.x - r.y * yq.y;
            sv[e].y -= r.x * yq.y + r.y * yq.x;
        }
        __syncthreads();
    }
}


// dst_s[e] = src_s[e] (strided copies of ns vectors)
__global__ __launch_bounds__(256) void bcopy_kernel(const float2* __restrict__ src, int64_t lds, float2* __restrict__ dst,
                                                    int64_t ldd, int64_t n)
{
    const int sy = blockIdx.y;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    dst[(int64_t)sy * ldd + e] = src[(int64_t)sy * lds + e];
}

// x_s += d_s
__global__ __launch_bounds__(256) void badd_kernel(float2* __restrict__ x, int64_t ldx, const float2* __restrict__ d,
                                                   int64_t ldd, int64_t n)
{
    const int sy = blockIdx.y;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    float2 a = x[(int64_t)sy * ldx + e];
    const float2 b = d[(int64_t)sy * ldd + e];
    a.x += b.x;
    a.y += b.y;
    x[(int64_t)sy * ldx + e] = a;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline bool gemv_t_vec_ok(const float2* A, int64_t lda)
{
    return ((uintptr_t)A % 16 == 0) && (lda % 2 == 0);
}
static inline unsigned g1(int64_t n, int per = 256) { return (unsigned)((n + per - 1) / per); }

// y_z = alpha A_z x_z + beta y_z for nz matrices (z = blockIdx.y, each at
// its own strides): one wavefront per row, four rows per workgroup, the
// row's products summed in a fixed order (lane partials, then a butterfly).
// y is not read when beta = 0.
__global__ __launch_bounds__(256) void gemv_strided_kernel(const float2* __restrict__ A, int64_t rows, int64_t cols,
                                                           int64_t lda, int64_t sa, const float2* __restrict__ x,
                                                           int64_t sx, float2* __restrict__ y, int64_t sy,
                                                           float2 alpha, float2 beta)
{
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= rows) return;
    const float2* a = A + (int64_t)blockIdx.y * sa + r * lda;
    const float2* xz = x + (int64_t)blockIdx.y * sx;
    float sr = 0.f, si = 0.f;
    for (int64_t c = lane; c < cols; c += 64) {
        const float2 av = a[c], xv = xz[c];
        sr = fmaf(av.x, xv.x, fmaf(-av.y, xv.y, sr));
        si = fmaf(av.x, xv.y, fmaf(av.y, xv.x, si));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off, 64);
        si += __shfl_xor(si, off, 64);
    }
    if (lane == 0) {
        float2* yz = y + (int64_t)blockIdx.y * sy + r;
        float2 v = make_float2(alpha.x * sr - alpha.y * si, alpha.x * si + alpha.y * sr);
        if (beta.x != 0.f || beta.y != 0.f) {
            const float2 o = *yz;
            v.x += beta.x * o.x - beta.y * o.y;
            v.y += beta.x * o.y + beta.y * o.x;
        }
        *yz = v;
    }
}

hipError_t launch_gemv_strided(const float2* A, int64_t rows, int64_t cols, int64_t lda, int64_t sa, const float2* x,
                               int64_t sx, float2* y, int64_t sy, int nz, float2 alpha, float2 beta, hipStream_t s)
{
    if (rows <= 0 || nz <= 0) return hipSuccess;
    for (int z0 = 0; z0 < nz; z0 += 65535) {  // grid y limit
        const int cz = nz - z0 < 65535 ? nz - z0 : 65535;
        hipLaunchKernelGGL(gemv_strided_kernel, dim3((unsigned)((rows + 3) / 4), (unsigned)cz), dim3(256), 0, s,
                           A + (int64_t)z0 * sa, rows, cols, lda, sa, x + (int64_t)z0 * sx, sx, y + (int64_t)z0 * sy,
                           sy, alpha, beta);
    }
    return hipGetLastError();
}

hipError_t launch_gemv(const float2* A, int64_t rows, int64_t cols, int64_t lda, int trans, const float2* x, float2* y,
                       float2 alpha, float2 beta, float2* work, hipStream_t s)
{
    if (rows <= 0 || cols <= 0) return hipSuccess;
    const int bz = (beta.x == 0.f && beta.y == 0.f) ? 1 : 0;
    if (trans == 0) {
        const bool vec = (cols % 2 == 0) && (lda % 2 == 0) && (((uintptr_t)A & 15) == 0) && (((uintptr_t)x & 15) == 0);
        if (vec) {
            constexpr int RW = kGemvRows;
            const int64_t waves = (rows + RW - 1) / RW;
            hipLaunchKernelGGL(gemv_n_vec_kernel<RW>, dim3(g1(waves, 4)), dim3(256), 0, s, (const float4*)A, rows,
                               cols / 2, lda / 2, (const float4*)x, y, alpha, beta, bz);
        } else {
            hipLaunchKernelGGL(gemv_n_scalar_kernel, dim3(g1(rows, 4)), dim3(256), 0, s, A, rows, cols, lda, x, y,
                               alpha, beta, bz);
        }
        return hipGetLastError();
    }
    if (gemv_t_vec_ok(A, lda)) {
        const int64_t nslab = (rows + kTVecSlab - 1) / kTVecSlab;
        hipLaunchKernelGGL(gemv_t_vec_partial_kernel, dim3(g1(cols, 128), (unsigned)nslab), dim3(256), 0, s, A, rows,
                           cols, lda, x, trans == 2 ? 1 : 0, work);
        hipLaunchKernelGGL(gemv_t_reduce_kernel, dim3(g1(cols)), dim3(256), 0, s, work, nslab, cols, y, alpha, beta,
                           bz);
        return hipGetLastError();
    }
    const int64_t nslab = (rows + kTSlabRows - 1) / kTSlabRows;
    hipLaunchKernelGGL(gemv_t_partial_kernel, dim3(g1(cols), (unsigned)nslab), dim3(256), 0, s, A, rows, cols, lda, x,
                       trans == 2 ? 1 : 0, work);
    hipLaunchKernelGGL(gemv_t_reduce_kernel, dim3(g1(cols)), dim3(256), 0, s, work, nslab, cols, y, alpha, beta, bz);
    return hipGetLastError();
}

int64_t gemv_work_elems(int64_t rows, int64_t cols, int trans)
{
    if (trans == 0) return 0;
    // enough for either partial layout (the vector form has the smaller slabs)
    return ((rows + kTVecSlab - 1) / kTVecSlab) * cols;
}

hipError_t launch_combine(const float2* A, int64_t lda, const float2* B, int64_t ldb, float2* C, int64_t ldc,
                          int64_t rows, int64_t cols, float2 a, float2 b, hipStream_t s)
{
    int64_t g = (rows * cols + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(combine_kernel, dim3((unsigned)g), dim3(256), 0, s, A, lda, B, ldb, C, ldc, rows, cols, a, b);
    return hipGetLastError();
}

hipError_t launch_multidot(const float2* V, int64_t ldv, int64_t n, int nq, const float2* w, double2* h,
                           double2* part, hipStream_t s)
{
    const bool vec = part && (n % 2 == 0) && (ldv % 2 == 0) && (((uintptr_t)V | (uintptr_t)w) & 15) == 0;
    if (!vec) {
        hipLaunchKernelGGL(multidot_kernel, dim3(nq), dim3(256), 0, s, V, ldv, n, w, h);
        return hipGetLastError();
    }
    // enough workgroups to fill the chip about four times, each at least 4096 entries
    int ns = (int)std::min<int64_t>(kDotMaxSplit, std::max<int64_t>(1, (1024 + nq - 1) / nq));
    ns = (int)std::min<int64_t>(ns, std::max<int64_t>(1, n / 4096));
    hipLaunchKernelGGL(multidot_split_kernel, dim3(nq, ns), dim3(256), 0, s, V, ldv, n, w, nq, part);
    hipLaunchKernelGGL(multidot_reduce_kernel, dim3((nq + 255) / 256), dim3(256), 0, s, part, ns, nq, h);
    return hipGetLastError();
}

int64_t multidot_part_elems(int nq) { return (int64_t)kDotMaxSplit * nq; }
hipError_t launch_gs_update(const float2* V, int64_t ldv, int64_t n, int nq, const double2* h, float2* w,
                            hipStream_t s)
{
    hipLaunchKernelGGL(basis_combine_kernel, dim3(g1(n, 64)), dim3(64 * kBasisWaves), 0, s, V, ldv, (int64_t)0, n, nq,
                       (const int*)nullptr, h, (int64_t)0, -1.0, w, (int64_t)0, (const int*)nullptr);
    return hipGetLastError();
}
hipError_t launch_norm2(const float2* w, int64_t n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(1024), 0, s, w, n, out);
    return hipGetLastError();
}
hipError_t launch_dot(const float2* a, const float2* b, int64_t n, int conj, double2* out, hipStream_t s)
{
    hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(1024), 0, s, a, b, n, conj, out);
    return hipGetLastError();
}
hipError_t launch_normalize(const float2* src, int64_t n, const double* nrm2, float2* out, hipStream_t s)
{
    hipLaunchKernelGGL(normalize_kernel, dim3(g1(n)), dim3(256), 0, s, src, n, nrm2, out);
    return hipGetLastError();
}
hipError_t launch_givens(double2* H, int m, int j, const double2* h, const double2* h2, const double* wn2,
                         double2* cs, double2* sn, double2* g, double* res, hipStream_t s)
{
    hipLaunchKernelGGL(gmres_givens_kernel, dim3(1), dim3(kGivensChunk), 0, s, H, m, j, h, h2, wn2, cs, sn, g, res);
    return hipGetLastError();
}

hipError_t launch_arnoldi_small(float2* V, int64_t n, int j, float2* w, double2* h, double2* h2, double2* H, int m,
                                double2* cs, double2* sn, double2* g, double* res, double* wn2, hipStream_t s)
{
    hipLaunchKernelGGL(arnoldi_small_kernel, dim3(1), dim3(kArnoldiThreads), 0, s, V, n, j, w, h, h2, H, m, cs, sn, g,
                       res, wn2);
    return hipGetLastError();
}
hipError_t launch_barnoldi_small(const BSys* sys, int ns, float2* V, int64_t ntot, int j, float2* w, double2* h,
                                 double2* h2, int hs, double2* H, int m, double2* cs, double2* sn, double2* g,
                                 double* res, const double* target, int* done, int* nk, double* wn2, int passes,
                                 hipStream_t s)
{
    hipLaunchKernelGGL(barnoldi_small_kernel, dim3((unsigned)ns), dim3(kArnoldiThreads), 0, s, sys, V, ntot, j, w, h,
                       h2, hs, H, m, cs, sn, g, res, target, done, nk, wn2, passes);
    return hipGetLastError();
}
hipError_t launch_backsolve(const double2* H, int m, int nk, const double2* g, double2* y, hipStream_t s)
{
    hipLaunchKernelGGL(backsolve_wg_kernel, dim3(1), dim3(256), 0, s, H, m, (int64_t)0, (const int*)nullptr, nk, g,
                       (int64_t)0, y, (int64_t)0);
    return hipGetLastError();
}
hipError_t launch_update_x(const float2* V, int64_t ldv, int64_t n, int nk, const double2* y, float2* x,
                           hipStream_t s)
{
    hipLaunchKernelGGL(basis_combine_kernel, dim3(g1(n, 64)), dim3(64 * kBasisWaves), 0, s, V, ldv, (int64_t)0, n, nk,
                       (const int*)nullptr, y, (int64_t)0, 1.0, x, (int64_t)0, (const int*)nullptr);
    return hipGetLastError();
}
hipError_t launch_init_g(double2* g, int m, const double* nrm2, hipStream_t s)
{
    hipLaunchKernelGGL(gmres_init_g_kernel, dim3(1), dim3(256), 0, s, g, m, nrm2);
    return hipGetLastError();
}
hipError_t launch_bicg_scalar(int which, double2* sc, hipStream_t s)
{
    if (which == 0) hipLaunchKernelGGL(bicg_beta_kernel, dim3(1), dim3(1), 0, s, sc);
    else if (which == 1) hipLaunchKernelGGL(bicg_alpha_kernel, dim3(1), dim3(1), 0, s, sc);
    else hipLaunchKernelGGL(bicg_omega_kernel, dim3(1), dim3(1), 0, s, sc);
    return hipGetLastError();
}
hipError_t launch_bicg_p(const float2* r, float2* p, const float2* v, const double2* sc, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(bicg_p_kernel, dim3(g1(n)), dim3(256), 0, s, r, p, v, sc, n);
    return hipGetLastError();
}
hipError_t launch_bicg_axpy(const float2* a, const float2* b, const double2* sc, int idx, double sign, float2* out,
                            int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(bicg_axpy_kernel, dim3(g1(n)), dim3(256), 0, s, a, b, sc, idx, sign, out, n);
    return hipGetLastError();
}
hipError_t launch_bicg_x(float2* x, const float2* p, const float2* sv, const double2* sc, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(bicg_x_kernel, dim3(g1(n)), dim3(256), 0, s, x, p, sv, sc, n);
    return hipGetLastError();
}

// ---- batched launchers ----
static inline dim3 gns(int64_t n, int ns) { return dim3((unsigned)((n + 255) / 256), (unsigned)ns); }

hipError_t launch_bmultidot(const float2* V, int64_t ldn, int ns, int64_t n, int nq, const float2* w, double2* h,
                            int hs, const int* done, hipStream_t s)
{
    hipLaunchKernelGGL(bmultidot_kernel, dim3(nq, ns), dim3(256), 0, s, V, ldn, ns, n, w, h, hs, done);
    return hipGetLastError();
}
hipError_t launch_bgs_update(const float2* V, int64_t ldn, int ns, int64_t n, int nq, const double2* h, int hs,
                             float2* w, const int* done, hipStream_t s)
{
    hipLaunchKernelGGL(basis_combine_kernel, dim3(g1(n, 64), (unsigned)ns), dim3(64 * kBasisWaves), 0, s, V,
                       ldn * ns, ldn, n, nq, (const int*)nullptr, h, (int64_t)hs, -1.0, w, ldn, done);
    return hipGetLastError();
}
hipError_t launch_bnorm2(const float2* w, int64_t ldn, int64_t n, int ns, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(bnorm2_kernel, dim3(ns), dim3(1024), 0, s, w, ldn, n, out);
    return hipGetLastError();
}
hipError_t launch_bnormalize(const float2* src, int64_t lds, int64_t n, int ns, const double* nrm2, float2* dst,
                             int64_t ldd, const int* done, hipStream_t s)
{
    hipLaunchKernelGGL(bnormalize_kernel, gns(n, ns), dim3(256), 0, s, src, lds, n, nrm2, dst, ldd, done);
    return hipGetLastError();
}
hipError_t launch_binit(double2* g, int m, const double* nrm2, const double* target, int* done, int* nk, int ns,
                        hipStream_t s)
{
    hipLaunchKernelGGL(binit_kernel, dim3((ns + 63) / 64), dim3(64), 0, s, g, m, nrm2, target, done, nk, ns);
    return hipGetLastError();
}
hipError_t launch_bgivens(double2* H, int m, int j, const double2* h, const double2* h2, int hs, const double* wn2,
                          double2* cs, double2* sn, double2* g, double* res, const double* target, int* done, int* nk,
                          int ns, hipStream_t s)
{
    hipLaunchKernelGGL(bgivens_kernel, dim3(ns), dim3(kGivensChunk), 0, s, H, m, j, h, h2, hs, wn2, cs, sn, g, res,
                       target, done, nk, ns);
    return hipGetLastError();
}
hipError_t launch_bbacksolve(const double2* H, int m, const int* nk, const double2* g, double2* y, int ns,
                             hipStream_t s)
{
    hipLaunchKernelGGL(backsolve_wg_kernel, dim3(ns), dim3(256), 0, s, H, m, (int64_t)(m + 1) * m, nk, 0, g,
                       (int64_t)(m + 1), y, (int64_t)m);
    return hipGetLastError();
}
hipError_t launch_bupdate_x(const float2* V, int64_t ldn, int ns, int64_t n, const int* nk, const double2* y, int m,
                            float2* x, int64_t ldx, hipStream_t s)
{
    hipLaunchKernelGGL(basis_combine_kernel, dim3(g1(n, 64), (unsigned)ns), dim3(64 * kBasisWaves), 0, s, V,
                       ldn * ns, ldn, n, 0, nk, y, (int64_t)m, 1.0, x, ldx, (const int*)nullptr);
    return hipGetLastError();
}
hipError_t launch_bcopy(const float2* src, int64_t lds, float2* dst, int64_t ldd, int64_t n, int ns, hipStream_t s)
{
    hipLaunchKernelGGL(bcopy_kernel, gns(n, ns), dim3(256), 0, s, src, lds, dst, ldd, n);
    return hipGetLastError();
}
hipError_t launch_badd(float2* x, int64_t ldx, const float2* d, int64_t ldd, int64_t n, int ns, hipStream_t s)
{
    hipLaunchKernelGGL(badd_kernel, gns(n, ns), dim3(256), 0, s, x, ldx, d, ldd, n);
    return hipGetLastError();
}

}  // namespace bem
