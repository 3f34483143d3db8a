// Low-precision storage and matrix-core products for the mixed-precision
// solve (gfx950).
//
// A complex matrix is stored as two planes Ar, Ai of 16-bit values (bf16 or
// fp16) holding round(scale * A); scale is a power of two, so fp16 keeps the
// small far-field entries in its normal range.  Columns [cols, ld) of each
// plane are zero so that kernels may read whole 32-column steps.
//
//  * gemv_lowp_kernel: one right-hand side, VALU; 16-byte loads of 8 entries
//    per plane, 4 rows per wavefront share each x fetch; HBM-bound at 4 B per
//    complex entry.
//  * gemm_lowp_mfma_kernel: 16 right-hand sides on v_mfma_f32_16x16x32_{bf16,f16}.
//    A wavefront owns 32 rows (two 16-row tiles).  Lane l loads 8 consecutive
//    entries of row (l % 16) at column offset 8 * (l / 16) of the current
//    32-column step -- exactly one MFMA A operand -- and the matching 8 entries
//    of right-hand side (l % 16) as the B operand.  Four products per tile:
//    Ar Xr, Ai Xi, Ar Xi, Ai Xr, accumulated in FP32; Y = (rr - ii) + i (ri + ir).
//  * gemm_c64_mfma_kernel: the same 16-RHS product on the complex64 matrix with
//    v_mfma_f32_16x16x4f32 (the FP32 residual of iterative refinement).
// Right-hand-side blocks are stored per system: X[s * ldx + k].
#include <hip/hip_runtime.h>

#include "bem_lowp.hpp"

namespace bem {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f32_to_f16(float f)
{
    _Float16 h = (_Float16)f;
    return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float f16_to_f32(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

template <int DT>
__device__ __forceinline__ uint16_t to_lp(float f)
{
    return DT == kDtBF16 ? f32_to_bf16_rne(f) : f32_to_f16(f);
}
template <int DT>
__device__ __forceinline__ float from_lp(uint16_t h)
{
    return DT == kDtBF16 ? bf16_to_f32(h) : f16_to_f32(h);
}

// ---------------------------------------------------------------------------
// conversions
// ---------------------------------------------------------------------------
template <int DT>
__global__ __launch_bounds__(256) void to_lowp_kernel(const float2* __restrict__ A, int64_t rows, int64_t cols,
                                                      int64_t lda, float scale, uint16_t* __restrict__ Ar,
                                                      uint16_t* __restrict__ Ai, int64_t ldo)
{
    const int64_t total = rows * ldo;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t r = e / ldo, c = e - r * ldo;
        float2 v = make_float2(0.f, 0.f);
        if (c < cols) v = A[r * lda + c];
        Ar[e] = to_lp<DT>(v.x * scale);
        Ai[e] = to_lp<DT>(v.y * scale);
    }
}

// |A|_max over the real and imaginary parts (block maxima, then one block)
__global__ __launch_bounds__(256) void absmax_kernel(const float2* __restrict__ A, int64_t rows, int64_t cols,
                                                     int64_t lda, float* __restrict__ out)
{
    __shared__ float sh[4];
    float m = 0.f;
    const int64_t total = rows * cols;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t r = e / cols, c = e - r * cols;
        const float2 v = A[r * lda + c];
        m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
}

// X (complex64, systems of length n at stride ldx_in) -> 16-bit planes
// (ldx, zero-padded) with a second, low-order plane per part (x = hi + lo,
// both in the 16-bit format), so the right-hand sides carry about twice the
// format's mantissa (fp16 low parts may be subnormal: their absolute
// resolution 2^-24 is far below a normalised Krylov vector's entries):
// planes Xr, Xi, Xr_lo = Xr + 16 ldx, Xi_lo = Xi + 16 ldx.
template <int DT>
__global__ __launch_bounds__(256) void vec_to_lowp_kernel(const float2* __restrict__ X, int64_t n, int64_t ldx_in,
                                                          int nrhs, uint16_t* __restrict__ Xr,
                                                          uint16_t* __restrict__ Xi, int64_t ldx)
{
    const int64_t total = (int64_t)nrhs * ldx;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t s = e / ldx, c = e - s * ldx;
        float2 v = make_float2(0.f, 0.f);
        if (c < n) v = X[s * ldx_in + c];
        const uint16_t hr = to_lp<DT>(v.x), hi = to_lp<DT>(v.y);
        Xr[e] = hr;
        Xi[e] = hi;
        Xr[e + 16 * ldx] = to_lp<DT>(v.x - from_lp<DT>(hr));
        Xi[e + 16 * ldx] = to_lp<DT>(v.y - from_lp<DT>(hi));
    }
}

// ---------------------------------------------------------------------------
// one right-hand side, VALU
// ---------------------------------------------------------------------------
template <int DT, int RW>
__global__ __launch_bounds__(256) void gemv_lowp_kernel(const uint16_t* __restrict__ Ar, const uint16_t* __restrict__ Ai,
                                                        int64_t rows, int64_t cols, int64_t ld, float inv_scale,
                                                        const float2* __restrict__ x, float2* __restrict__ y,
                                                        float2 alpha, float2 beta, int beta_zero)
{
    const int lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * RW;
    if (r0 >= rows) return;
    float accr[RW], acci[RW];
#pragma unroll
    for (int q = 0; q < RW; ++q) accr[q] = acci[q] = 0.f;
    const int64_t nchunk = (cols + 7) / 8;  // planes are zero beyond cols
    for (int64_t c8 = lane; c8 < nchunk; c8 += 64) {
        const int64_t c = c8 * 8;
        float xr[8], xi[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float2 v = (c + e < cols) ? x[c + e] : make_float2(0.f, 0.f);
            xr[e] = v.x;
            xi[e] = v.y;
        }
#pragma unroll
        for (int q = 0; q < RW; ++q) {
            const int64_t r = min(r0 + q, rows - 1);
            const u16x8 ar = __builtin_nontemporal_load(reinterpret_cast<const u16x8*>(Ar + r * ld + c));
            const u16x8 ai = __builtin_nontemporal_load(reinterpret_cast<const u16x8*>(Ai + r * ld + c));
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float a = from_lp<DT>(ar[e]), b = from_lp<DT>(ai[e]);
                accr[q] = fmaf(a, xr[e], fmaf(-b, xi[e], accr[q]));
                acci[q] = fmaf(a, xi[e], fmaf(b, xr[e], acci[q]));
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RW; ++q) {
        float sr = accr[q] * inv_scale, si = acci[q] * inv_scale;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            sr += __shfl_xor(sr, off, 64);
            si += __shfl_xor(si, off, 64);
        }
        const int64_t r = r0 + q;
        if (lane == q && r < rows) {
            float2 o = make_float2(alpha.x * sr - alpha.y * si, alpha.x * si + alpha.y * sr);
            if (!beta_zero) {
                const float2 yo = y[r];
                o.x += beta.x * yo.x - beta.y * yo.y;
                o.y += beta.x * yo.y + beta.y * yo.x;
            }
            y[r] = o;
        }
    }
}

// ---------------------------------------------------------------------------
// 16 right-hand sides on the matrix cores
// ---------------------------------------------------------------------------
template <int DT>
__device__ __forceinline__ f32x4 mfma16x16x32(u16x8 a, u16x8 b, f32x4 c)
{
    if constexpr (DT == kDtBF16)
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                       c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c,
                                                      0, 0, 0);
}

template <int DT>
__global__ __launch_bounds__(256) void gemm_lowp_mfma_kernel(const uint16_t* __restrict__ Ar,
                                                             const uint16_t* __restrict__ Ai, int64_t rows,
                                                             int64_t ld, float inv_scale,
                                                             const uint16_t* __restrict__ Xr,
                                                             const uint16_t* __restrict__ Xi, int64_t ldx, int nrhs,
                                                             float2* __restrict__ Y, int64_t ldy, float2 alpha,
                                                             float2 beta, int beta_zero)
{
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, q = lane >> 4;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 32;
    if (r0 >= rows) return;
    const int64_t ra = min(r0 + li, rows - 1), rb = min(r0 + 16 + li, rows - 1);
    const uint16_t* par = Ar + ra * ld + 8 * q;
    const uint16_t* pai = Ai + ra * ld + 8 * q;
    const uint16_t* pbr = Ar + rb * ld + 8 * q;
    const uint16_t* pbi = Ai + rb * ld + 8 * q;
    const int s = li < nrhs ? li : 0;  // right-hand side of this lane's B operand
    const uint16_t* pxr = Xr + (int64_t)s * ldx + 8 * q;
    const uint16_t* pxi = Xi + (int64_t)s * ldx + 8 * q;
    const u16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
    const bool live = li < nrhs;
    constexpr bool kSplit = true;  // hi + lo right-hand-side planes
    f32x4 a_rr = {0, 0, 0, 0}, a_ii = a_rr, a_ri = a_rr, a_ir = a_rr;
    f32x4 b_rr = a_rr, b_ii = a_rr, b_ri = a_rr, b_ir = a_rr;
    const int64_t kend = ld;  // ld is a multiple of 32; padded columns are zero
    for (int64_t kb = 0; kb < kend; kb += 32) {
        const u16x8 xr = live ? *reinterpret_cast<const u16x8*>(pxr + kb) : zero8;
        const u16x8 xi = live ? *reinterpret_cast<const u16x8*>(pxi + kb) : zero8;
        const u16x8 ar = __builtin_nontemporal_load(reinterpret_cast<const u16x8*>(par + kb));
        const u16x8 ai = __builtin_nontemporal_load(reinterpret_cast<const u16x8*>(pai + kb));
        const u16x8 br = __builtin_nontemporal_load(reinterpret_cast<const u16x8*>(pbr + kb));
        const u16x8 bi = __builtin_nontemporal_load(reinterpret_cast<const u16x8*>(pbi + kb));
        a_rr = mfma16x16x32<DT>(ar, xr, a_rr);
        a_ii = mfma16x16x32<DT>(ai, xi, a_ii);
        a_ri = mfma16x16x32<DT>(ar, xi, a_ri);
        a_ir = mfma16x16x32<DT>(ai, xr, a_ir);
        b_rr = mfma16x16x32<DT>(br, xr, b_rr);
        b_ii = mfma16x16x32<DT>(bi, xi, b_ii);
        b_ri = mfma16x16x32<DT>(br, xi, b_ri);
        b_ir = mfma16x16x32<DT>(bi, xr, b_ir);
        if constexpr (kSplit) {
            const u16x8 xrl = live ? *reinterpret_cast<const u16x8*>(pxr + 16 * ldx + kb) : zero8;
            const u16x8 xil = live ? *reinterpret_cast<const u16x8*>(pxi + 16 * ldx + kb) : zero8;
            a_rr = mfma16x16x32<DT>(ar, xrl, a_rr);
            a_ii = mfma16x16x32<DT>(ai, xil, a_ii);
            a_ri = mfma16x16x32<DT>(ar, xil, a_ri);
            a_ir = mfma16x16x32<DT>(ai, xrl, a_ir);
            b_rr = mfma16x16x32<DT>(br, xrl, b_rr);
            b_ii = mfma16x16x32<DT>(bi, xil, b_ii);
            b_ri = mfma16x16x32<DT>(br, xil, b_ri);
            b_ir = mfma16x16x32<DT>(bi, xrl, b_ir);
        }
    }
    // D layout: lane l holds column (right-hand side) l % 16, rows 4 (l / 16) + e
    if (!live) return;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int64_t r = r0 + 16 * t + 4 * q + e;
            if (r >= rows) continue;
            const float yr = (t ? (b_rr[e] - b_ii[e]) : (a_rr[e] - a_ii[e])) * inv_scale;
            const float yi = (t ? (b_ri[e] + b_ir[e]) : (a_ri[e] + a_ir[e])) * inv_scale;
            float2 o = make_float2(alpha.x * yr - alpha.y * yi, alpha.x * yi + alpha.y * yr);
            float2* dst = &Y[(int64_t)li * ldy + r];
            if (!beta_zero) {
                const float2 yo = *dst;
                o.x += beta.x * yo.x - beta.y * yo.y;
                o.y += beta.x * yo.y + beta.y * yo.x;
            }
            *dst = o;
        }
    }
}

// complex64 matrix, 16 right-hand sides, FP32 matrix cores.  Lane l owns
// row r0 + (l % 16) and the columns kb + 8 (l / 16) + [0, 8) of each 32-wide
// step; each of the 8 sub-steps is one K = 4 slice over the four lane groups.
__global__ __launch_bounds__(256) void gemm_c64_mfma_kernel(const float2* __restrict__ A, int64_t rows, int64_t cols,
                                                            int64_t lda, const float2* __restrict__ X, int64_t ldx,
                                                            int nrhs, float2* __restrict__ Y, int64_t ldy,
                                                            float2 alpha, float2 beta, int beta_zero)
{
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, q = lane >> 4;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (r0 >= rows) return;
    const int64_t ra = min(r0 + li, rows - 1);
    const float2* pa = A + ra * lda;
    const bool live = li < nrhs;
    const float2* px = X + (int64_t)(live ? li : 0) * ldx;
    f32x4 rr = {0, 0, 0, 0}, ii = rr, ri = rr, ir = rr;
    for (int64_t kb = 0; kb < cols; kb += 32) {
        float2 a[8], b[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int64_t c = kb + 8 * q + e;
            const bool in = c < cols;
            a[e] = in ? pa[c] : make_float2(0.f, 0.f);
            b[e] = (in && live) ? px[c] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            rr = __builtin_amdgcn_mfma_f32_16x16x4f32(a[e].x, b[e].x, rr, 0, 0, 0);
            ii = __builtin_amdgcn_mfma_f32_16x16x4f32(a[e].y, b[e].y, ii, 0, 0, 0);
            ri = __builtin_amdgcn_mfma_f32_16x16x4f32(a[e].x, b[e].y, ri, 0, 0, 0);
            ir = __builtin_amdgcn_mfma_f32_16x16x4f32(a[e].y, b[e].x, ir, 0, 0, 0);
        }
    }
    if (!live) return;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int64_t r = r0 + 4 * q + e;
        if (r >= rows) continue;
        const float yr = rr[e] - ii[e], yi = ri[e] + ir[e];
        float2 o = make_float2(alpha.x * yr - alpha.y * yi, alpha.x * yi + alpha.y * yr);
        float2* dst = &Y[(int64_t)li * ldy + r];
        if (!beta_zero) {
            const float2 yo = *dst;
            o.x += beta.x * yo.x - beta.y * yo.y;
            o.y += beta.x * yo.y + beta.y * yo.x;
        }
        *dst = o;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static unsigned gcap(int64_t n, int64_t cap = 65536)
{
    int64_t g = (n + 255) / 256;
    return (unsigned)(g < 1 ? 1 : (g > cap ? cap : g));
}

hipError_t launch_absmax(const float2* A, int64_t rows, int64_t cols, int64_t lda, float* partial, int nblocks,
                         hipStream_t s)
{
    hipLaunchKernelGGL(absmax_kernel, dim3(nblocks), dim3(256), 0, s, A, rows, cols, lda, partial);
    return hipGetLastError();
}

hipError_t launch_to_lowp(int dt, const float2* A, int64_t rows, int64_t cols, int64_t lda, float scale, uint16_t* Ar,
                          uint16_t* Ai, int64_t ldo, hipStream_t s)
{
    if (dt == kDtBF16)
        hipLaunchKernelGGL(to_lowp_kernel<kDtBF16>, dim3(gcap(rows * ldo)), dim3(256), 0, s, A, rows, cols, lda, scale,
                           Ar, Ai, ldo);
    else
        hipLaunchKernelGGL(to_lowp_kernel<kDtFP16>, dim3(gcap(rows * ldo)), dim3(256), 0, s, A, rows, cols, lda, scale,
                           Ar, Ai, ldo);
    return hipGetLastError();
}

hipError_t launch_vec_to_lowp(int dt, const float2* X, int64_t n, int64_t ldx_in, int nrhs, uint16_t* Xr, uint16_t* Xi,
                              int64_t ldx, hipStream_t s)
{
    if (dt == kDtBF16)
        hipLaunchKernelGGL(vec_to_lowp_kernel<kDtBF16>, dim3(gcap(nrhs * ldx)), dim3(256), 0, s, X, n, ldx_in, nrhs, Xr,
                           Xi, ldx);
    else
        hipLaunchKernelGGL(vec_to_lowp_kernel<kDtFP16>, dim3(gcap(nrhs * ldx)), dim3(256), 0, s, X, n, ldx_in, nrhs, Xr,
                           Xi, ldx);
    return hipGetLastError();
}

hipError_t launch_gemv_lowp(int dt, const uint16_t* Ar, const uint16_t* Ai, int64_t rows, int64_t cols, int64_t ld,
                            float inv_scale, const float2* x, float2* y, float2 alpha, float2 beta, hipStream_t s)
{
    constexpr int RW = 4;
    const int bz = (beta.x == 0.f && beta.y == 0.f) ? 1 : 0;
    const unsigned g = (unsigned)(((rows + RW - 1) / RW + 3) / 4);
    if (dt == kDtBF16)
        hipLaunchKernelGGL((gemv_lowp_kernel<kDtBF16, RW>), dim3(g), dim3(256), 0, s, Ar, Ai, rows, cols, ld, inv_scale,
                           x, y, alpha, beta, bz);
    else
        hipLaunchKernelGGL((gemv_lowp_kernel<kDtFP16, RW>), dim3(g), dim3(256), 0, s, Ar, Ai, rows, cols, ld, inv_scale,
                           x, y, alpha, beta, bz);
    return hipGetLastError();
}

hipError_t launch_gemm_lowp(int dt, const uint16_t* Ar, const uint16_t* Ai, int64_t rows, int64_t ld, float inv_scale,
                            const uint16_t* Xr, const uint16_t* Xi, int64_t ldx, int nrhs, float2* Y, int64_t ldy,
                            float2 alpha, float2 beta, hipStream_t s)
{
    const int bz = (beta.x == 0.f && beta.y == 0.f) ? 1 : 0;
    const unsigned g = (unsigned)(((rows + 31) / 32 + 3) / 4);
    if (dt == kDtBF16)
        hipLaunchKernelGGL(gemm_lowp_mfma_kernel<kDtBF16>, dim3(g), dim3(256), 0, s, Ar, Ai, rows, ld, inv_scale, Xr,
                           Xi, ldx, nrhs, Y, ldy, alpha, beta, bz);
    else
        hipLaunchKernelGGL(gemm_lowp_mfma_kernel<kDtFP16>, dim3(g), dim3(256), 0, s, Ar, Ai, rows, ld, inv_scale, Xr,
                           Xi, ldx, nrhs, Y, ldy, alpha, beta, bz);
    return hipGetLastError();
}

hipError_t launch_gemm_c64(const float2* A, int64_t rows, int64_t cols, int64_t lda, const float2* X, int64_t ldx,
                           int nrhs, float2* Y, int64_t ldy, float2 alpha, float2 beta, hipStream_t s)
{
    const int bz = (beta.x == 0.f && beta.y == 0.f) ? 1 : 0;
    const unsigned g = (unsigned)(((rows + 15) / 16 + 3) / 4);
    hipLaunchKernelGGL(gemm_c64_mfma_kernel, dim3(g), dim3(256), 0, s, A, rows, cols, lda, X, ldx, nrhs, Y, ldy, alpha,
                       beta, bz);
    return hipGetLastError();
}

}  // namespace bem
