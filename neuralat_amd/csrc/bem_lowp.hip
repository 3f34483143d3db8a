// Low-precision storage and matrix-core products for the mixed-precision
// solve (gfx950).
//
// A complex matrix is stored as two planes Ar, Ai of 16-bit values (bf16 or
// fp16) holding round(scale * A); scale is a power of two, so fp16 keeps the
// small far-field entries in its normal range.  The planes are TILED for the
// matrix cores: a tile covers 16 rows x 32 columns (1 KB per plane), tiles are
// ordered row-tile major ((T, K) at (T * ld / 32 + K) * 512 elements), and
// inside a tile lane l = 16 q + i of a wavefront finds its MFMA A operand --
// the 8 entries of row 16 T + i at columns 32 K + 8 q + [0, 8) -- at
// elements [8 l, 8 l + 8).  One wavefront step therefore reads one
// contiguous 1 KB block per plane, and a wavefront streams contiguous memory
// (row-major planes made each step touch 16 rows 64 bytes apiece, which ran
// at a third of HBM bandwidth).  Rows are padded to a multiple of 16 and
// columns to ld (a multiple of 32) with zeros.
//
//  * gemm_lowp_mfma_kernel: up to 16 right-hand sides on
//    v_mfma_f32_16x16x32_{bf16,f16}; a wavefront owns two row tiles (32 rows)
//    and a range of column tiles (split-K); four products per tile:
//    Ar Xr, Ai Xi, Ar Xi, Ai Xr, accumulated in FP32; Y = (rr - ii) + i (ri + ir).
//    One right-hand side uses the same kernel (the product is HBM-bound).
//  * gemm_c64_mfma3_kernel: the same 16-RHS product on the complex64 matrix
//    with v_mfma_f32_16x16x4f32, three real products per complex entry (the
//    FP32 residual of the refinement), right-hand sides staged in LDS per
//    workgroup, loads one step ahead of the products.
// Right-hand-side blocks are stored per system: X[s * ldx + k].
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bem_lowp.hpp"
#include "bem_lowp_cvt.hpp"

namespace bem {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// conversions
// ---------------------------------------------------------------------------
template <int DT>
__global__ __launch_bounds__(256) void to_lowp_kernel(const float2* __restrict__ A, int64_t rows, int64_t cols,
                                                      int64_t lda, float scale, uint16_t* __restrict__ Ar,
                                                      uint16_t* __restrict__ Ai, int64_t ldo)
{
    // one thread per lane slot (8 entries) of the tiled planes
    const int64_t nk = ldo / 32;
    const int64_t total = (rows + 15) / 16 * nk * 64;
    for (int64_t sl = (int64_t)blockIdx.x * 256 + threadIdx.x; sl < total; sl += (int64_t)gridDim.x * 256) {
        const int64_t tile = sl >> 6;
        const int l = (int)(sl & 63);
        const int64_t T = tile / nk, K = tile - T * nk;
        const int64_t r = 16 * T + (l & 15), c0 = 32 * K + 8 * (l >> 4);
        u16x8 hr, hi;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float2 v = make_float2(0.f, 0.f);
            if (r < rows && c0 + e < cols) v = A[r * lda + c0 + e];
            hr[e] = to_lp<DT>(v.x * scale);
            hi[e] = to_lp<DT>(v.y * scale);
        }
        *reinterpret_cast<u16x8*>(Ar + sl * 8) = hr;
        *reinterpret_cast<u16x8*>(Ai + sl * 8) = hi;
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
                                                             float2 beta, int beta_zero, int64_t kchunk,
                                                             float2* __restrict__ part)
{
    // The right-hand-side planes (hi and lo parts of Xr, Xi) of each 32-column
    // step are staged in LDS once per workgroup (double-buffered, one barrier
    // per step): read per wavefront they cost as many bytes as the matrix.
    __shared__ u16x8 xs[2][4][16][4];
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, q = lane >> 4;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 32;
    const int64_t nk = ld / 32, ntile = (rows + 15) / 16;
    const int64_t ta = min(r0 / 16, ntile - 1), tb = min(ta + 1, ntile - 1);  // clamped: tiles may not exist
    // tile (T, K) of a plane at (T nk + K) 512; this lane's 8 entries at 8 lane
    const uint16_t* par = Ar + ta * nk * 512 + 8 * lane;
    const uint16_t* pai = Ai + ta * nk * 512 + 8 * lane;
    const uint16_t* pbr = Ar + tb * nk * 512 + 8 * lane;
    const uint16_t* pbi = Ai + tb * nk * 512 + 8 * lane;
    // this thread's share of the staging: plane (t / 64), right-hand side ((t / 4) % 16), 8 columns (t % 4)
    const int tp = threadIdx.x >> 6, ts = (threadIdx.x >> 2) & 15, tg = threadIdx.x & 3;
    const uint16_t* px = ((tp & 1) ? Xi : Xr) + (tp >> 1) * 16 * ldx + (int64_t)ts * ldx + 8 * tg;
    const bool tlive = ts < nrhs;
    const u16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
    const bool live = li < nrhs;
    f32x4 a_rr = {0, 0, 0, 0}, a_ii = a_rr, a_ri = a_rr, a_ir = a_rr;
    f32x4 b_rr = a_rr, b_ii = a_rr, b_ri = a_rr, b_ir = a_rr;
    // columns [k0, k1) of this split (ld and kchunk are multiples of 32; padded columns are zero)
    const int64_t k0 = (int64_t)blockIdx.y * kchunk;
    const int64_t kend = min(ld, k0 + kchunk);
    int buf = 0;
    for (int64_t kb = k0; kb < kend; kb += 32) {
        const int64_t to = kb * 16;  // (kb / 32) tiles of 512 entries; matrix loads issued before the barrier
        const u16x8 ar = __builtin_nontemporal_load(reinterpret_cast<const u16x8*>(par + to));
        const u16x8 ai = __builtin_nontemporal_load(reinterpret_cast<const u16x8*>(pai + to));
        const u16x8 br = __builtin_nontemporal_load(reinterpret_cast<const u16x8*>(pbr + to));
        const u16x8 bi = __builtin_nontemporal_load(reinterpret_cast<const u16x8*>(pbi + to));
        xs[buf][tp][ts][tg] = tlive ? *reinterpret_cast<const u16x8*>(px + kb) : zero8;
        __syncthreads();
        const u16x8 xr = xs[buf][0][li][q], xi = xs[buf][1][li][q];
        const u16x8 xrl = xs[buf][2][li][q], xil = xs[buf][3][li][q];
        a_rr = mfma16x16x32<DT>(ar, xr, a_rr);
        a_ii = mfma16x16x32<DT>(ai, xi, a_ii);
        a_ri = mfma16x16x32<DT>(ar, xi, a_ri);
        a_ir = mfma16x16x32<DT>(ai, xr, a_ir);
        b_rr = mfma16x16x32<DT>(br, xr, b_rr);
        b_ii = mfma16x16x32<DT>(bi, xi, b_ii);
        b_ri = mfma16x16x32<DT>(br, xi, b_ri);
        b_ir = mfma16x16x32<DT>(bi, xr, b_ir);
        a_rr = mfma16x16x32<DT>(ar, xrl, a_rr);
        a_ii = mfma16x16x32<DT>(ai, xil, a_ii);
        a_ri = mfma16x16x32<DT>(ar, xil, a_ri);
        a_ir = mfma16x16x32<DT>(ai, xrl, a_ir);
        b_rr = mfma16x16x32<DT>(br, xrl, b_rr);
        b_ii = mfma16x16x32<DT>(bi, xil, b_ii);
        b_ri = mfma16x16x32<DT>(br, xil, b_ri);
        b_ir = mfma16x16x32<DT>(bi, xrl, b_ir);
        buf ^= 1;  // the next step writes the other buffer: every wave has left it (previous barrier)
    }
    if (r0 >= rows) return;  // after the loop: all waves take part in every barrier
    // D layout: lane l holds column (right-hand side) l % 16, rows 4 (l / 16) + e
    if (!live) return;
    if (part) {  // split-K partial, combined by splitk_reduce_kernel
        float2* P = part + ((int64_t)blockIdx.y * nrhs + li) * rows;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int64_t r = r0 + 16 * t + 4 * q + e;
                if (r < rows)
                    P[r] = t ? make_float2(b_rr[e] - b_ii[e], b_ri[e] + b_ir[e])
                             : make_float2(a_rr[e] - a_ii[e], a_ri[e] + a_ir[e]);
            }
        }
        return;
    }
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

// complex64 matrix, up to 16 right-hand sides, FP32 matrix cores.  A
// workgroup of kC64Waves wavefronts owns kC64Waves row tiles of 16 rows and a
// column range (split-K); lane l owns row r0 + (l % 16).  Three real
// matrix-core products per complex entry
// (P1 = Ar Xr, P2 = Ai Xi, P3 = (Ar + Ai)(Xr + Xi); Yr = P1 - P2,
// Yi = P3 - P1 - P2): at 16 right-hand sides four FP32 products per entry
// need 16 FLOP per matrix byte, three need 12 -- 644 GFLOP for an 81920 x
// 81920 matrix, 4.1 ms of the FP32 matrix-core peak against 6.7 ms of HBM
// time, so the loads must overlap the products.
//  * The right-hand sides of kC64Steps 32-column steps are staged in LDS per
//    barrier (double-buffered) as two planes (Xr, Xi; Xr + Xi is formed at
//    the read) laid out [plane][lane group q][row i][8 kC64Steps + 1]: the
//    lanes of each 32-lane half of a wavefront read 32 different banks.  The
//    next stage's entries are read from global memory while a stage computes.
//  * The matrix entries of step s + 1 are in flight while step s computes.
template <bool VEC, int kC64Waves, int kC64Steps>
__global__ __launch_bounds__(64 * kC64Waves) __attribute__((amdgpu_waves_per_eu(4))) void gemm_c64_mfma3_kernel(
    const float2* __restrict__ A, int64_t rows, int64_t cols, int64_t lda, const float2* __restrict__ X, int64_t ldx,
    int nrhs, float2* __restrict__ Y, int64_t ldy, float2 alpha, float2 beta, int beta_zero, int64_t kchunk,
    float2* __restrict__ part)
{
    constexpr int kC64Pad = 8 * kC64Steps + 1;
    __shared__ float xs[2][2][4][16][kC64Pad];
    constexpr int W = 32 * kC64Steps;                 // columns per stage
    constexpr int kXPer = 16 * W / (64 * kC64Waves);  // right-hand-side entries staged per thread and stage
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int li = lane & 15, q = lane >> 4;
    const int64_t r0 = ((int64_t)blockIdx.x * kC64Waves + wv) * 16;
    const float2* pa = A + min(r0 + li, rows - 1) * lda;
    f32x4 p1 = {0, 0, 0, 0}, p2 = p1, p3 = p1;
    const int64_t k0 = (int64_t)blockIdx.y * kchunk;  // columns [k0, k0 + kchunk) of this split
    const int64_t kend = min(cols, k0 + kchunk);

    // Right-hand sides: the next stage's entries are read (clamped addresses)
    // while this stage computes; entries outside the split are staged as zero.
    float2 xv[kXPer];
    auto load_x = [&](int64_t kb) {
#pragma unroll
        for (int i = 0; i < kXPer; ++i) {
            const int t = threadIdx.x + i * 64 * kC64Waves;
            const int sx = t / W, c = t - sx * W;  // right-hand side, column in the stage
            xv[i] = X[(int64_t)min(sx, nrhs - 1) * ldx + min(kb + c, kend - 1)];
        }
    };
    auto store_x = [&](int b, int64_t kb) {
#pragma unroll
        for (int i = 0; i < kXPer; ++i) {
            const int t = threadIdx.x + i * 64 * kC64Waves;
            const int sx = t / W, c = t - sx * W;
            const bool ok = sx < nrhs && kb + c < kend;
            const float xr = ok ? xv[i].x : 0.f, xi = ok ? xv[i].y : 0.f;
            // column 32 st + r is sub-step e of lane group cq (see the matrix loads)
            const int st = c >> 5, r = c & 31, cq = (r >> 1) & 3, e = 2 * (r >> 3) + (r & 1);
            xs[b][0][cq][sx][8 * st + e] = xr;
            xs[b][1][cq][sx][8 * st + e] = xi;
        }
    };
    // sub-step e (0..7) of a 32-column step at kc: 16 rows x 4 columns, lane
    // group q holding column kc + 8 (e / 2) + 2 q + e % 2
    auto mfma_step = [&](int b, int st, float2 (&a)[8]) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float br = xs[b][0][q][li][8 * st + e];
            const float bi = xs[b][1][q][li][8 * st + e];
            const float bs = br + bi;
            p1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[e].x, br, p1, 0, 0, 0);
            p2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[e].y, bi, p2, 0, 0, 0);
            p3 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[e].x + a[e].y, bs, p3, 0, 0, 0);
        }
    };

    // guarded loads, no look-ahead: every stage without VEC, the last
    // (partial) stage with it
    auto plain_stage = [&](int b, int64_t kb) {
#pragma unroll
        for (int st = 0; st < kC64Steps; ++st) {
            const int64_t kc = kb + 32 * st;
            float2 a[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int64_t c = kc + 8 * (e >> 1) + 2 * q + (e & 1);
                a[e] = c < kend ? pa[c] : make_float2(0.f, 0.f);
            }
            mfma_step(b, st, a);
        }
    };

    load_x(k0);
    int buf = 0;
    int64_t kb = k0;
    if constexpr (VEC) {
        // Whole stages (rows 16-byte aligned, kchunk a multiple of 32, cols
        // even): each lane reads four 16-byte pieces per step (piece e / 2 of
        // the four lanes of a row covers 64 contiguous bytes; the two halves
        // of a 128-byte line follow each other).  The pieces of step s + 1 are
        // in flight while step s computes: two register sets alternating with
        // the step's parity, no branch or select between load and use (the
        // look-ahead past the last whole stage reads a clamped address and is
        // not used).
        auto load_a = [&](int64_t kc, float2 (&d)[8]) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const f32x4 v = *reinterpret_cast<const f32x4*>(pa + min(kc + 8 * e + 2 * q, kend - 2));
                d[2 * e] = make_float2(v.x, v.y);
                d[2 * e + 1] = make_float2(v.z, v.w);
            }
        };
        static_assert(kC64Steps % 2 == 0, "two register sets alternate over the steps of a stage");
        const int64_t kfull = k0 + (kend - k0) / W * W;
        float2 a0[8], a1[8];
        load_a(k0, a0);
        for (; kb < kfull; kb += W) {
            store_x(buf, kb);
            __syncthreads();
            load_x(kb + W);
#pragma unroll
            for (int st = 0; st < kC64Steps; st += 2) {
                // scheduling barriers keep each look-ahead load ahead of the
                // step it overlaps (left alone, the compiler sinks the loads
                // next to their use to save registers)
                load_a(kb + 32 * (st + 1), a1);
                __builtin_amdgcn_sched_barrier(0);
                mfma_step(buf, st, a0);
                __builtin_amdgcn_sched_barrier(0);
                load_a(kb + 32 * (st + 2), a0);
                __builtin_amdgcn_sched_barrier(0);
                mfma_step(buf, st + 1, a1);
                __builtin_amdgcn_sched_barrier(0);
            }
            buf ^= 1;  // the next stage writes the other buffer: every wave has left it (previous barrier)
        }
    }
    for (; kb < kend; kb += W) {
        store_x(buf, kb);
        __syncthreads();
        load_x(kb + W);
        plain_stage(buf, kb);
        buf ^= 1;
    }
    if (li >= nrhs) return;  // after the loop: all waves take part in every barrier
    if (part) {  // split-K partial, combined by splitk_reduce_kernel
        float2* P = part + ((int64_t)blockIdx.y * nrhs + li) * rows;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t r = r0 + 4 * q + e;
            if (r < rows) P[r] = make_float2(p1[e] - p2[e], p3[e] - p1[e] - p2[e]);
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int64_t r = r0 + 4 * q + e;
        if (r >= rows) continue;
        const float yr = p1[e] - p2[e], yi = p3[e] - p1[e] - p2[e];
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

// Y[s] = alpha * scale * sum_k part[k][s] + beta * Y[s]; fixed order over the
// ks column splits (deterministic).  One thread per (row, right-hand side).
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const float2* __restrict__ part, int ks, int nrhs,
                                                            int64_t rows, float scale, float2* __restrict__ Y,
                                                            int64_t ldy, float2 alpha, float2 beta, int beta_zero)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int sy = blockIdx.y;
    if (r >= rows) return;
    float yr = 0.f, yi = 0.f;
    for (int k = 0; k < ks; ++k) {
        const float2 p = part[((int64_t)k * nrhs + sy) * rows + r];
        yr += p.x;
        yi += p.y;
    }
    yr *= scale;
    yi *= scale;
    float2 o = make_float2(alpha.x * yr - alpha.y * yi, alpha.x * yi + alpha.y * yr);
    float2* dst = &Y[(int64_t)sy * ldy + r];
    if (!beta_zero) {
        const float2 yo = *dst;
        o.x += beta.x * yo.x - beta.y * yo.y;
        o.y += beta.x * yo.y + beta.y * yo.x;
    }
    *dst = o;
}

// Column splits for a product over `tiles` row tiles and `cols` columns: enough
// wavefronts to cover the chip about 16 times (a single split leaves 2-5 per
// SIMD at 81920 rows, too few to hide HBM latency), at most kMaxSplit, each
// split at least 256 columns, split width a multiple of 32.
static void choose_split(int64_t tiles, int64_t cols, int& ks, int64_t& kchunk)
{
    constexpr int64_t kTargetWaves = 16384;
    int64_t k = (kTargetWaves + tiles - 1) / tiles;
    k = std::min<int64_t>(k, kMaxSplit);
    k = std::min<int64_t>(k, std::max<int64_t>(1, cols / 256));
    k = std::max<int64_t>(k, 1);
    kchunk = ((cols + k - 1) / k + 31) / 32 * 32;
    ks = (int)((cols + kchunk - 1) / kchunk);
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
        hipLaunchKernelGGL(to_lowp_kernel<kDtBF16>, dim3(gcap(lowp_plane_elems(rows, ldo) / 8)), dim3(256), 0, s, A,
                           rows, cols, lda, scale, Ar, Ai, ldo);
    else
        hipLaunchKernelGGL(to_lowp_kernel<kDtFP16>, dim3(gcap(lowp_plane_elems(rows, ldo) / 8)), dim3(256), 0, s, A,
                           rows, cols, lda, scale, Ar, Ai, ldo);
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

hipError_t launch_gemm_lowp(int dt, const uint16_t* Ar, const uint16_t* Ai, int64_t rows, int64_t ld, float inv_scale,
                            const uint16_t* Xr, const uint16_t* Xi, int64_t ldx, int nrhs, float2* Y, int64_t ldy,
                            float2 alpha, float2 beta, float2* part, hipStream_t s)
{
    const int bz = (beta.x == 0.f && beta.y == 0.f) ? 1 : 0;
    const int64_t tiles = (rows + 31) / 32;
    const unsigned g = (unsigned)((tiles + 3) / 4);
    int ks = 1;
    int64_t kchunk = ld;
    if (part) choose_split(tiles, ld, ks, kchunk);
    float2* pp = ks > 1 ? part : nullptr;
    if (dt == kDtBF16)
        hipLaunchKernelGGL(gemm_lowp_mfma_kernel<kDtBF16>, dim3(g, (unsigned)ks), dim3(256), 0, s, Ar, Ai, rows, ld,
                           inv_scale, Xr, Xi, ldx, nrhs, Y, ldy, alpha, beta, bz, kchunk, pp);
    else
        hipLaunchKernelGGL(gemm_lowp_mfma_kernel<kDtFP16>, dim3(g, (unsigned)ks), dim3(256), 0, s, Ar, Ai, rows, ld,
                           inv_scale, Xr, Xi, ldx, nrhs, Y, ldy, alpha, beta, bz, kchunk, pp);
    if (pp)
        hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)((rows + 255) / 256), (unsigned)nrhs), dim3(256), 0, s,
                           pp, ks, nrhs, rows, inv_scale, Y, ldy, alpha, beta, bz);
    return hipGetLastError();
}

hipError_t launch_gemm_c64(const float2* A, int64_t rows, int64_t cols, int64_t lda, const float2* X, int64_t ldx,
                           int nrhs, float2* Y, int64_t ldy, float2 alpha, float2 beta, float2* part, hipStream_t s)
{
    constexpr int kWaves = 8;
    const int bz = (beta.x == 0.f && beta.y == 0.f) ? 1 : 0;
    const int64_t tiles = (rows + 15) / 16;
    const unsigned g = (unsigned)((tiles + kWaves - 1) / kWaves);
    int ks = 1;
    int64_t kchunk = (cols + 31) / 32 * 32;
    if (part) choose_split(tiles, cols, ks, kchunk);
    float2* pp = ks > 1 ? part : nullptr;
    // 16-byte pieces: rows 16-byte aligned and an even number of columns
    // (splits start at multiples of 32 columns)
    const bool vec = (lda % 2 == 0) && (cols % 2 == 0) && ((uintptr_t)A % 16 == 0);
    if (vec)
        hipLaunchKernelGGL((gemm_c64_mfma3_kernel<true, kWaves, 8>), dim3(g, (unsigned)ks), dim3(64 * kWaves), 0, s, A,
                           rows, cols, lda, X, ldx, nrhs, Y, ldy, alpha, beta, bz, kchunk, pp);
    else
        hipLaunchKernelGGL((gemm_c64_mfma3_kernel<false, kWaves, 4>), dim3(g, (unsigned)ks), dim3(64 * kWaves), 0, s, A,
                           rows, cols, lda, X, ldx, nrhs, Y, ldy, alpha, beta, bz, kchunk, pp);
    if (pp)
        hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)((rows + 255) / 256), (unsigned)nrhs), dim3(256), 0, s,
                           pp, ks, nrhs, rows, 1.f, Y, ldy, alpha, beta, bz);
    return hipGetLastError();
}

int64_t splitk_part_elems(int64_t rows) { return (int64_t)kMaxSplit * 16 * rows; }

int64_t lowp_plane_elems(int64_t rows, int64_t ld) { return (rows + 15) / 16 * 16 * ld; }

}  // namespace bem
