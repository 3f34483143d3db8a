// Boundary-element assembly kernels for gfx950.
//
// Regular pairs (triangles sharing no vertex; reference: bind.cu:67-79,
// integrand.h:187-235,280-329):
//   * one lane owns one COLUMN triangle j (its Q quadrature points and
//     weights live in VGPRs); the ROW triangle i is wave-uniform, so its data
//     is fetched with scalar loads and used as SGPR operands;
//   * DP0: each (i, j) entry is produced by exactly one lane and stored
//     directly -- a wave writes 64 consecutive complex entries of a row;
//   * P1: a workgroup owns a (row tile x column tile) of a locality-ordered
//     triangle list; the 3x3 basis blocks of its pairs are summed in LDS on
//     the tile's compacted vertex slots, then flushed once per tile entry.
// Neighbour pairs (sharing >= 1 vertex; bind.cu:81-152, integrand.h:63-185,
// 331-394) go to a separate Sauter-Schwab kernel: one wavefront per pair,
// the n^4 tensor Gauss points spread over the lanes, FP64 accumulation.
#include "bem_device.hpp"
#include "bem_internal.hpp"

namespace bem {

// ---------------------------------------------------------------------------
// utility kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void fill_c64_kernel(float2* __restrict__ out, int64_t rows, int64_t cols,
                                                       int64_t ld, float2 val, int64_t zstride)
{
    out += (int64_t)blockIdx.z * zstride;
    int64_t total = rows * cols;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = e / cols, c = e - r * cols;
        out[r * ld + c] = val;
    }
}

// ---------------------------------------------------------------------------
// regular pairs, DP0 basis
// ---------------------------------------------------------------------------
// Grid: x = column blocks of 256 triangles, y = row tiles of kDp0RowTile rows
// (rows [row0, row1) of the operator).  out row r holds triangle row0 + r.
template <int OP, int Q, bool CLAMP>
__global__ __launch_bounds__(256) void dp0_regular_kernel(const float* __restrict__ rec, int64_t nt, int64_t row0,
                                                          int64_t row1, const KBatch kb, float2* __restrict__ out,
                                                          int64_t ld)
{
    const float k = kb.k[blockIdx.z];
    out += (int64_t)blockIdx.z * kb.zstride;
    rec += (int64_t)blockIdx.z * kb.gstride;
    constexpr int S = Rec<Q>::kFloats;
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool active = j < nt;
    const int64_t jj = active ? j : nt - 1;
    const float* cr = rec + jj * S;
    float py[3 * Q], wy[Q];
#pragma unroll
    for (int b = 0; b < 3 * Q; ++b) py[b] = cr[b];
#pragma unroll
    for (int b = 0; b < Q; ++b) wy[b] = cr[3 * Q + b] * kInv4Pi;
    const float ny0 = cr[4 * Q], ny1 = cr[4 * Q + 1], ny2 = cr[4 * Q + 2];
    const float krev = k * kInv2Pi;

    const int64_t i_begin = row0 + (int64_t)blockIdx.y * kDp0RowTile;
    const int64_t i_end = min(i_begin + (int64_t)kDp0RowTile, row1);
    for (int64_t i = i_begin; i < i_end; ++i) {
        const float* rr = rec + i * S;  // wave-uniform
        const float nx0 = rr[4 * Q], nx1 = rr[4 * Q + 1], nx2 = rr[4 * Q + 2];
        float accr = 0.f, acci = 0.f;
        if constexpr (OP == SLP || OP == DLP || OP == ADLP) {
            // Specialised evaluation.  The normal factor is constant along one
            // triangle: (y - x).n_y = (y_0 - x_a).n_y depends on the row point only
            // (DLP), (y - x).n_x = (y_b - x_0).n_x on the column point only (ADLP);
            // both are formed from a point difference first to stay accurate for
            // nearly coplanar pairs.
            float qb[Q];
            if constexpr (OP == ADLP) {
#pragma unroll
                for (int b = 0; b < Q; ++b)
                    qb[b] = wy[b] * ((py[3 * b] - rr[0]) * nx0 + (py[3 * b + 1] - rr[1]) * nx1 +
                                     (py[3 * b + 2] - rr[2]) * nx2);
            }
#pragma unroll
            for (int a = 0; a < Q; ++a) {
                const float xa0 = rr[3 * a], xa1 = rr[3 * a + 1], xa2 = rr[3 * a + 2];
                float sr = 0.f, si = 0.f;
#pragma unroll
                for (int b = 0; b < Q; ++b) {
                    const float dx = py[3 * b] - xa0, dy = py[3 * b + 1] - xa1, dz = py[3 * b + 2] - xa2;
                    float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    if (CLAMP) r2 = fmaxf(r2, kEps2);  // max(r, eps) = sqrt(max(r^2, eps^2))
                    const float rinv = __builtin_amdgcn_rsqf(r2);
                    const float r = r2 * rinv;
                    float sn, cs;
                    sincos_rev(r * krev, sn, cs);
                    if constexpr (OP == SLP) {
                        const float g = rinv * wy[b];
                        sr = fmaf(g, cs, sr);
                        si = fmaf(g, sn, si);
                    } else {
                        const float th = r * k;
                        const float re = fmaf(th, sn, cs), im = fmaf(-th, cs, sn);
                        const float p = rinv * rinv * (rinv * (OP == DLP ? wy[b] : qb[b]));
                        sr = fmaf(p, re, sr);
                        si = fmaf(p, im, si);
                    }
                }
                float ua = rr[3 * Q + a];
                if constexpr (OP == DLP)  // -(y - x).n_y = (x_a - y_0).n_y
                    ua *= (xa0 - py[0]) * ny0 + (xa1 - py[1]) * ny1 + (xa2 - py[2]) * ny2;
                else if constexpr (OP == ADLP)  // -G (1 - ikr)(x - y).n_x / r^2 = +G (...) (y - x).n_x / r^2
                    ;
                accr = fmaf(ua, sr, accr);
                acci = fmaf(ua, si, acci);
            }
        } else {
#pragma unroll
            for (int a = 0; a < Q; ++a) {
                const float xa0 = rr[3 * a], xa1 = rr[3 * a + 1], xa2 = rr[3 * a + 2];
                float sr = 0.f, si = 0.f;
#pragma unroll
                for (int b = 0; b < Q; ++b) {
                    float vr, vi, tr = 0.f, ti = 0.f;
                    kernel4pi<OP, CLAMP>(py[3 * b] - xa0, py[3 * b + 1] - xa1, py[3 * b + 2] - xa2, nx0, nx1, nx2,
                                         ny0, ny1, ny2, k, krev, vr, vi, tr, ti);
                    sr = fmaf(wy[b], vr, sr);
                    si = fmaf(wy[b], vi, si);
                }
                const float ua = rr[3 * Q + a];
                accr = fmaf(ua, sr, accr);
                acci = fmaf(ua, si, acci);
            }
        }
        if (active) out[(i - row0) * ld + j] = make_float2(accr, acci);
    }
}

// ---------------------------------------------------------------------------
// fused DP0 CBIE system: lhs = -K (regular pairs) and partial sums of V psi
// ---------------------------------------------------------------------------
// The single- and double-layer kernels share the point difference, r and
// e^{ikr}; the single-layer entries are contracted with psi on the fly and
// never stored.  Both operators are evaluated once per UNORDERED pair {i, j}:
// the 6 x 6 Gauss-point pairs of (i, j) and (j, i) are the same points with the
// same weight products (w_a |J_i| / 2) (w_b |J_j| / 2), so
//   V_ji = V_ij,
//   K_ij = sum_a (x_a - y_0).n_j  sum_b F_ab,
//   K_ji = sum_b (y_b - x_0).n_i  sum_a F_ab,
// with F_ab = w_a w_b e^{ikr} (1 - ikr) / r^3 (x_0, y_0: any point of the
// plane of T_i, T_j).  One evaluation of F_ab feeds both entries, which halves
// the transcendental work of the pair sweep.
//
// A task is a wavefront: columns J*64 + lane (lane = column triangle, its
// points in VGPRs), rows [I*64 + sub*RT, +RT) with I <= J (row data
// wave-uniform, in SGPRs).  In diagonal blocks (I == J) only j > i is taken.
// K_ij is stored directly (64 consecutive entries of row i); K_ji is staged
// in LDS for 16 rows and written as 16-entry row segments (4 full 128-byte
// lines per store).  Pairs sharing a vertex are left out of the V psi sums
// (the singular kernel provides them); their lhs entries are overwritten by
// it.  Partial sums (deterministic, independent of RT):
//   prow[J * nt + i]               = sum over columns j in block J of V_ij psi_j,
//   pcol[((I * 64 / RT) + sub) * nt + j] = sum over the task's rows of V_ij psi_i,
// the latter formed per 16-row chunk and combined as ((c0 + c1) + (c2 + c3)).
constexpr int kSymBlock = 64;
constexpr int kStageCols = 16;

__device__ __forceinline__ float wave_sum_f32(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block pair index p -> (I, J), I <= J, p = J (J + 1) / 2 + I
__device__ __forceinline__ void tri_block_pair(int64_t p, int64_t& I, int64_t& J)
{
    int64_t jb = (int64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while ((jb + 1) * (jb + 2) / 2 <= p) ++jb;
    while (jb * (jb + 1) / 2 > p) --jb;
    J = jb;
    I = p - jb * (jb + 1) / 2;
}

// a wave-uniform 64-bit value held in SGPRs (v must be uniform, or come from lane 0)
__device__ __forceinline__ int64_t uniform_i64(int64_t v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

template <int Q, bool CLAMP, int RT>
__device__ __forceinline__ void cbie_sym_task(const float* __restrict__ rec, const int4* __restrict__ tris, int64_t nt,
                                              float k, const float2* __restrict__ psi, float2* __restrict__ lhs,
                                              int64_t ld, float2* __restrict__ prow, float2* __restrict__ pcol,
                                              int64_t I, int64_t sub, int64_t J,
                                              float2 (*__restrict__ stage)[kStageCols + 1])
{
    static_assert(RT % kStageCols == 0 && kSymBlock % RT == 0, "row tile");
    constexpr int S = Rec<Q>::kFloats;
    const int lane = threadIdx.x & 63;
    const int64_t j = J * kSymBlock + lane;
    const bool jin = j < nt;
    const int64_t jj = jin ? j : nt - 1;
    const float* cr = rec + jj * S;
    float py[3 * Q], wy[Q];
#pragma unroll
    for (int b = 0; b < 3 * Q; ++b) py[b] = cr[b];
#pragma unroll
    for (int b = 0; b < Q; ++b) wy[b] = cr[3 * Q + b] * kInv4Pi;
    const float ny0 = cr[4 * Q], ny1 = cr[4 * Q + 1], ny2 = cr[4 * Q + 2];
    const int4 tj = tris[jj];
    const float2 pj = jin ? psi[jj] : make_float2(0.f, 0.f);
    const float krev = k * kInv2Pi;
    const bool diag = I == J;

    const int64_t i_begin = I * kSymBlock + sub * RT;
    const int64_t i_end = min(i_begin + (int64_t)RT, nt);
    // column sums of 16-row chunks: pa = c0 (+ c1), pb = c2 (+ c3)
    float par = 0.f, pai = 0.f, pbr = 0.f, pbi = 0.f;

#pragma unroll 1
    for (int64_t i0 = i_begin; i0 < i_end; i0 += kStageCols) {
        const int chunk = (int)((i0 - i_begin) / kStageCols);
        const int64_t i1 = min(i0 + (int64_t)kStageCols, i_end);
        float cr_acc = 0.f, ci_acc = 0.f;
#pragma unroll 1
        for (int64_t i = i0; i < i1; ++i) {
            const float* rr = rec + i * S;  // wave-uniform
            const float nx0 = rr[4 * Q], nx1 = rr[4 * Q + 1], nx2 = rr[4 * Q + 2];
            float er[Q], ei[Q];
#pragma unroll
            for (int b = 0; b < Q; ++b) er[b] = ei[b] = 0.f;
            float kr = 0.f, ki = 0.f, vr = 0.f, vi = 0.f;
#pragma unroll 1  // 6 independent evaluations per iteration; unrolling a as well costs registers
            for (int a = 0; a < Q; ++a) {
                const float xa0 = rr[3 * a], xa1 = rr[3 * a + 1], xa2 = rr[3 * a + 2];
                const float ua = rr[3 * Q + a];
                float dr = 0.f, di = 0.f, sr = 0.f, si = 0.f;
#pragma unroll
                for (int b = 0; b < Q; ++b) {
                    const float dx = py[3 * b] - xa0, dy = py[3 * b + 1] - xa1, dz = py[3 * b + 2] - xa2;
                    float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    if (CLAMP) r2 = fmaxf(r2, kEps2);
                    const float rinv = __builtin_amdgcn_rsqf(r2);
                    const float r = r2 * rinv;
                    float sn, cs;
                    sincos_rev(r * krev, sn, cs);
                    const float g = rinv * wy[b];
                    sr = fmaf(g, cs, sr);
                    si = fmaf(g, sn, si);
                    const float th = r * k;
                    const float p = g * rinv * rinv;
                    const float t1 = fmaf(th, sn, cs), t2 = fmaf(-th, cs, sn);
                    dr = fmaf(p, t1, dr);
                    di = fmaf(p, t2, di);
                    const float pu = p * ua;
                    er[b] = fmaf(pu, t1, er[b]);
                    ei[b] = fmaf(pu, t2, ei[b]);
                }
                const float un = ua * ((xa0 - py[0]) * ny0 + (xa1 - py[1]) * ny1 + (xa2 - py[2]) * ny2);
                kr = fmaf(un, dr, kr);
                ki = fmaf(un, di, ki);
                vr = fmaf(ua, sr, vr);
                vi = fmaf(ua, si, vi);
            }
            // K_ji = sum_b (y_b - x_0).n_i E_b
            float tr = 0.f, ti = 0.f;
#pragma unroll
            for (int b = 0; b < Q; ++b) {
                const float mb = (py[3 * b] - rr[0]) * nx0 + (py[3 * b + 1] - rr[1]) * nx1 +
                                 (py[3 * b + 2] - rr[2]) * nx2;
                tr = fmaf(mb, er[b], tr);
                ti = fmaf(mb, ei[b], ti);
            }
            const bool valid = jin && (!diag || j > i);
            if (valid) lhs[i * ld + j] = make_float2(-kr, -ki);
            stage[lane][i - i0] = make_float2(-tr, -ti);
            const int4 ti4 = tris[i];
            const bool shared = ti4.x == tj.x || ti4.x == tj.y || ti4.x == tj.z || ti4.y == tj.x ||
                                ti4.y == tj.y || ti4.y == tj.z || ti4.z == tj.x || ti4.z == tj.y || ti4.z == tj.z;
            const bool use = valid && !shared;
            const float2 pi_ = psi[i];
            float rsr = use ? vr * pj.x - vi * pj.y : 0.f;
            float rsi = use ? vr * pj.y + vi * pj.x : 0.f;
            rsr = wave_sum_f32(rsr);
            rsi = wave_sum_f32(rsi);
            if (lane == 0) prow[J * nt + i] = make_float2(rsr, rsi);
            if (use) {
                cr_acc = fmaf(vr, pi_.x, fmaf(-vi, pi_.y, cr_acc));
                ci_acc = fmaf(vr, pi_.y, fmaf(vi, pi_.x, ci_acc));
            }
        }
        if (chunk == 0) {
            par = cr_acc;
            pai = ci_acc;
        } else if (chunk == 1) {
            par = par + cr_acc;
            pai = pai + ci_acc;
        } else if (chunk == 2) {
            pbr = cr_acc;
            pbi = ci_acc;
        } else {
            pbr = pbr + cr_acc;
            pbi = pbi + ci_acc;
        }
        // transposed entries: stage[jl][c] is (row J*64 + jl, column i0 + c)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int c = lane & (kStageCols - 1);
        const int64_t ic = i0 + c;
#pragma unroll 4
        for (int s4 = 0; s4 < kSymBlock / 4; ++s4) {
            const int jl = s4 * 4 + (lane >> 4);
            const int64_t jr = J * kSymBlock + jl;
            if (ic < i1 && jr < nt && (!diag || jr > ic)) lhs[jr * ld + ic] = stage[jl][c];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (jin) {
        // chunk sums combined pairwise, so that tasks of 16, 32 or 64 rows give
        // the same bits after the finalize combines their slots the same way
        if constexpr (RT == 64)
            pcol[I * nt + j] = make_float2(par + pbr, pai + pbi);
        else
            pcol[(I * (kSymBlock / RT) + sub) * nt + j] = make_float2(par, pai);
    }
}

// Tasks of one mesh: task t -> block pair t / (64 / RT), row sub-tile t % (64 / RT).
__host__ __device__ inline int64_t cbie_sym_tasks(int64_t nt, int rt)
{
    const int64_t nb = (nt + kSymBlock - 1) / kSymBlock;
    return nb * (nb + 1) / 2 * (kSymBlock / rt);
}

template <int Q, bool CLAMP, int RT>
__global__ __launch_bounds__(256) void dp0_cbie_kernel(const float* __restrict__ rec, const int4* __restrict__ tris,
                                                       int64_t nt, float k, const float2* __restrict__ psi,
                                                       float2* __restrict__ lhs, int64_t ld,
                                                       float2* __restrict__ prow, float2* __restrict__ pcol,
                                                       int64_t ntasks)
{
    __shared__ float2 stage[4][kSymBlock][kStageCols + 1];
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t t = (int64_t)blockIdx.x * 4 + wave;
    if (t >= ntasks) return;  // wave-uniform; the kernel has no workgroup barrier
    int64_t I, J;
    tri_block_pair(t / (kSymBlock / RT), I, J);
    cbie_sym_task<Q, CLAMP, RT>(rec, tris, nt, k, psi, lhs, ld, prow, pcol, I, t % (kSymBlock / RT), J,
                                stage[wave]);
}

// ---------------------------------------------------------------------------
// batched (mesh, k) items: one persistent launch over a queue of tasks
// ---------------------------------------------------------------------------
// Each wavefront pulls task indices from a device counter until the queue is
// empty; task t belongs to the item whose task0 is the largest <= t.  Items
// are ordered largest first, so the long-running items start early.  Waves
// are independent (no workgroup barrier), so a wave that finds the queue
// empty simply exits.
template <int Q, bool CLAMP, int RT>
__global__ __launch_bounds__(256) void batch_cbie_persistent(const BatchItemDev* __restrict__ items, int nitems,
                                                             int64_t total_tasks,
                                                             unsigned long long* __restrict__ counter)
{
    __shared__ float2 stage[4][kSymBlock][kStageCols + 1];
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    while (true) {
        unsigned long long got = 0ull;
        if (lane == 0) got = atomicAdd(counter, 1ull);
        const int64_t t = uniform_i64((int64_t)got);  // lane 0 is the first active lane
        if (t >= total_tasks) break;
        int lo = 0, hi = nitems - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (items[mid].task0 <= t) lo = mid;
            else hi = mid - 1;
        }
        const BatchItemDev it = items[lo];
        const int64_t local = t - it.task0;
        int64_t I, J;
        tri_block_pair(local / (kSymBlock / RT), I, J);
        cbie_sym_task<Q, CLAMP, RT>(it.rec, it.tris, it.nt, it.k, it.psi, it.lhs, it.ld, it.prow, it.pcol, I,
                                    local % (kSymBlock / RT), J, stage[wave]);
    }
}

__global__ void zero_u64_kernel(unsigned long long* p) { *p = 0ull; }

template <int RT>
static hipError_t batch_persistent_t(const BatchItemDev* items, int nitems, int64_t total, unsigned long long* counter,
                                     int clamp, hipStream_t s)
{
    int dev = 0, ncu = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (clamp)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, batch_cbie_persistent<6, true, RT>, 256, 0);
    else
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, batch_cbie_persistent<6, false, RT>, 256, 0);
    if (ncu <= 0) ncu = 256;
    if (per <= 0) per = 2;
    int64_t grid = (int64_t)ncu * per;
    if (grid * 4 > total) grid = (total + 3) / 4;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(1), 0, s, counter);
    if (clamp)
        hipLaunchKernelGGL((batch_cbie_persistent<6, true, RT>), dim3((unsigned)grid), dim3(256), 0, s, items, nitems,
                           total, counter);
    else
        hipLaunchKernelGGL((batch_cbie_persistent<6, false, RT>), dim3((unsigned)grid), dim3(256), 0, s, items, nitems,
                           total, counter);
    return hipGetLastError();
}

hipError_t launch_batch_cbie_persistent(const BatchItemDev* items, int nitems, int64_t total, int rt,
                                        unsigned long long* counter, int clamp, hipStream_t s)
{
    switch (rt) {
    case 16: return batch_persistent_t<16>(items, nitems, total, counter, clamp, s);
    case 32: return batch_persistent_t<32>(items, nitems, total, counter, clamp, s);
    case 64: return batch_persistent_t<64>(items, nitems, total, counter, clamp, s);
    default: return hipErrorInvalidValue;
    }
}

int64_t batch_cbie_tasks(int64_t nt, int rt) { return cbie_sym_tasks(nt, rt); }

// rhs_i = -(row partials + column partials + singular V psi terms of row i);
// lhs_ii += |T_i| / 2 (the M/2 of the CBIE).  One thread per row, fixed
// summation order; column slots of one 64-row block are combined pairwise
// exactly as a 64-row task combines its 16-row chunks.
__global__ __launch_bounds__(256) void cbie_finalize_kernel(const float2* __restrict__ prow,
                                                            const float2* __restrict__ pcol, int nsub,
                                                            const float2* __restrict__ vpsi,
                                                            const int64_t* __restrict__ pair_ptr,
                                                            const float* __restrict__ geo, int64_t nt,
                                                            float2* __restrict__ lhs, int64_t ld,
                                                            float2* __restrict__ rhs)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nt) return;
    const int64_t nb = (nt + kSymBlock - 1) / kSymBlock;
    const int64_t bi = i / kSymBlock;
    double sr = 0.0, si = 0.0;
    for (int64_t J = bi; J < nb; ++J) {
        const float2 p = prow[J * nt + i];
        sr += p.x;
        si += p.y;
    }
    for (int64_t I = 0; I <= bi; ++I) {
        const float2* c = pcol + I * nsub * nt + i;
        float2 v;
        if (nsub == 1) {
            v = c[0];
        } else if (nsub == 2) {
            v = make_float2(c[0].x + c[nt].x, c[0].y + c[nt].y);
        } else {
            v = make_float2((c[0].x + c[nt].x) + (c[2 * nt].x + c[3 * nt].x),
                            (c[0].y + c[nt].y) + (c[2 * nt].y + c[3 * nt].y));
        }
        sr += v.x;
        si += v.y;
    }
    for (int64_t e = pair_ptr[i]; e < pair_ptr[i + 1]; ++e) {
        sr += vpsi[e].x;
        si += vpsi[e].y;
    }
    rhs[i] = make_float2((float)-sr, (float)-si);
    lhs[i * ld + i].x += 0.25f * geo[i * 16 + 12];  // M_ii / 2 = |T_i| / 2 = |J_i| / 4
}

// ---------------------------------------------------------------------------
// regular pairs, P1 basis (LDS tile accumulation)
// ---------------------------------------------------------------------------
// Workgroup = (row tile, column tile) of a locality-ordered triangle list.
// For each row triangle of the tile, phase 1: lane p evaluates the pair with
// column triangle p and writes its 3 x 3 block (row-basis a, column-basis b)
// to an LDS stage, no atomics; phase 2: lane w owns column vertex w of the
// tile, gathers from the stage the entries of the triangles incident to w
// (host-built incidence lists) and adds them to its own accumulator column,
// so every LDS accumulator entry has exactly one writer.  (Adding the blocks
// with LDS float atomics left the kernel waiting on LDS issue 64 % of its wave
// cycles.)  Each tile entry is flushed once with a global float atomic.
// SYM (operators whose pair blocks satisfy B(j, i) = B(i, j)^T: single layer,
// hypersingular): each unordered pair is evaluated once.  Work item w of a
// 1-D grid pairs row tile r with a column tile c >= the column tile holding r
// (row tiles nest in column tiles); in that own column tile only columns
// after the row's position are taken.  The blocks are flushed at (row
// vertices, column vertices) only; p1_symmetrize_kernel then adds the
// transpose (flushing the transposed block with atomics as well scattered
// them over every row of the matrix, and the atomics missed in cache).
// PAIR (double layer and its adjoint): the same half sweep, and each
// Gauss-point pair's exp(ikr)(1 - ikr)/r^3 serves both blocks of the
// unordered pair: B(i, j) with the normal factor (x - y).n_y and B(j, i)^T
// with (y - x).n_x.  B(i, j) goes through the LDS accumulator to out;
// B(j, i)^T through the same stage and owner gather into 2 x 20 registers
// per lane (one per row-tile vertex slot, real and imaginary: the LDS has no
// room for a second accumulator at 3 workgroups per CU; the slot numbers are
// the same on every lane, so the registers are addressed by them through
// indexed register moves); after B(i, j) is flushed, the LDS accumulator takes
// B(j, i)^T and flushes it at the transposed positions (column vertices, row
// vertices) of out.  The adjoint double layer is the transpose of the
// double layer, so it evaluates the double layer and swaps the two blocks.
// kSweepAllRg: every ordered pair of the Burton-Miller LHS, with the
// Burton-Miller RHS contracted with a vertex vector g in the same pass
// (the RHS operator itself is never stored)
enum { kSweepAll = 0, kSweepSym = 1, kSweepPair = 2, kSweepAllRg = 3 };

// The 3 x 3 block B(i, j) (row-basis a, column-basis b) of one regular pair
// of the P1 operator OP with the Q-point rule: rr = the row triangle's
// quadrature record, rci its curls (HYP, BM_LHS); py, wy, ny, cj the column
// triangle's points, weights / 4 pi, normal and curls.  br, bi are added to
// (zero them first); with PAIR also tr, tii = B(j, i)^T of the paired sweep.
// EOP is the evaluated operator (the paired adjoint double layer evaluates
// the double layer).
template <int OP, int Q, bool CLAMP, bool PAIR, int EOP, bool RG = false>
__device__ __forceinline__ void p1_pair_block(const float* __restrict__ rr, const float* __restrict__ rci,
                                              const float* py, const float* wy, float ny0, float ny1, float ny2,
                                              const float* cj, float k, float krev, float invk, float* br, float* bi,
                                              float* tr, float* tii, const float* ghat = nullptr,
                                              float* rgs = nullptr)
{
    static_assert(!RG || OP == BM_LHS, "the right-hand-side contraction rides on the Burton-Miller LHS");
    // RG: the Burton-Miller RHS kernel -G + (i/k) G (1 - ikr) (x - y).n_x / r^2
    // at the same Gauss-point pairs, contracted with ghat (g at the column
    // points, 2Q floats) into rgs (the three row corners, complex)
    float qR[RG ? Q : 1], qRk[RG ? Q : 1];
    if constexpr (RG) {
        const float nx0 = rr[4 * Q], nx1 = rr[4 * Q + 1], nx2 = rr[4 * Q + 2];
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            const float yn = (py[3 * b] - rr[0]) * nx0 + (py[3 * b + 1] - rr[1]) * nx1 +
                             (py[3 * b + 2] - rr[2]) * nx2;
            qR[b] = -yn;  // (x_0 - y_b).n_x = (x_a - y_b).n_x on the row plane
            qRk[b] = -yn * invk;
        }
    }
    float gr9 = 0.f, gi9 = 0.f;  // plain-G sum (HYP/BM_LHS secondary)
    // ADLP: the normal factor (y_b - x_0).n_x is constant along the row
    // triangle's plane, so it is formed once per column point
    float qb[Q], qbk[Q];
    if constexpr (EOP == ADLP || EOP == BM_RHS) {
        const float nx0 = rr[4 * Q], nx1 = rr[4 * Q + 1], nx2 = rr[4 * Q + 2];
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            const float yn = (py[3 * b] - rr[0]) * nx0 + (py[3 * b + 1] - rr[1]) * nx1 +
                             (py[3 * b + 2] - rr[2]) * nx2;
            if constexpr (EOP == ADLP) {
                qb[b] = wy[b] * yn;
            } else {  // BM_RHS: (x_0 - y_b).n_x and its 1/k multiple
                qb[b] = -yn;
                qbk[b] = -yn * invk;
            }
        }
    }
    // PAIR: the transposed block's normal factor (y_b - x_0).n_x per
    // column point
    float qt[PAIR ? Q : 1];
    if constexpr (PAIR) {
        const float nx0 = rr[4 * Q], nx1 = rr[4 * Q + 1], nx2 = rr[4 * Q + 2];
#pragma unroll
        for (int b = 0; b < Q; ++b)
            qt[b] = (py[3 * b] - rr[0]) * nx0 + (py[3 * b + 1] - rr[1]) * nx1 + (py[3 * b + 2] - rr[2]) * nx2;
    }
    // BM_LHS: -k n_x.n_y once per pair; (x_a - y_0).n_y once per row point
    float mknn = 0.f;
    if constexpr (OP == BM_LHS)
        mknn = -k * (rr[4 * Q] * ny0 + rr[4 * Q + 1] * ny1 + rr[4 * Q + 2] * ny2);
#pragma unroll
    for (int a = 0; a < Q; ++a) {
        const float xa0 = rr[3 * a], xa1 = rr[3 * a + 1], xa2 = rr[3 * a + 2];
        float sr[3] = {0.f, 0.f, 0.f}, si[3] = {0.f, 0.f, 0.f};
        float ur[3] = {0.f, 0.f, 0.f}, ui[3] = {0.f, 0.f, 0.f};  // PAIR: transposed sums
        float tr9 = 0.f, ti9 = 0.f;
        float sRr = 0.f, sRi = 0.f;  // RG: sum_b (RHS kernel) w_b ghat_b
        float mca = 0.f, kca = 0.f;
        if constexpr (OP == BM_LHS) {
            const float ca = (xa0 - py[0]) * ny0 + (xa1 - py[1]) * ny1 + (xa2 - py[2]) * ny2;
            mca = -ca;
            kca = k * ca;
        }
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            float vr, vi, gr = 0.f, gi = 0.f;
            float wr = 0.f, wi = 0.f;  // PAIR: the transposed block's value
            const float dx = py[3 * b] - xa0, dy = py[3 * b + 1] - xa1, dz = py[3 * b + 2] - xa2;
            if constexpr (EOP == SLP || EOP == DLP || EOP == ADLP) {
                // specialised evaluations (as the DP0 kernels); the DLP normal
                // factor (x_a - y_0).n_y is applied per row point below
                float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                if (CLAMP) r2 = fmaxf(r2, kEps2);
                const float rinv = __builtin_amdgcn_rsqf(r2);
                const float r = r2 * rinv;
                float sn, cs;
                sincos_rev(r * krev, sn, cs);
                if constexpr (EOP == SLP) {
                    const float g = rinv * wy[b];
                    vr = g * cs;
                    vi = g * sn;
                } else {
                    const float th = r * k;
                    const float pw = rinv * rinv * (rinv * (EOP == DLP ? wy[b] : qb[b]));
                    vr = pw * fmaf(th, sn, cs);
                    vi = pw * fmaf(-th, cs, sn);
                    if constexpr (PAIR) {  // (y_b - x_0).n_x replaces (x_a - y_0).n_y
                        wr = vr * qt[b];
                        wi = vi * qt[b];
                    }
                }
            } else if constexpr (OP == BM_LHS || OP == BM_RHS) {
                // specialised Burton-Miller evaluations: value = G (er + i ei)
                // with the normal factors constant per row point (BM_LHS:
                // (y_b - x_a).n_y = -(x_a - y_0).n_y) or per column point
                // (BM_RHS: (x_a - y_b).n_x = (x_0 - y_b).n_x)
                float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                if (CLAMP) r2 = fmaxf(r2, kEps2);
                const float rinv = __builtin_amdgcn_rsqf(r2);
                const float r = r2 * rinv;
                float sn, cs;
                sincos_rev(r * krev, sn, cs);
                const float rw = rinv * wy[b];
                const float g_r = cs * rw, g_i = sn * rw;  // w_b G
                float er, ei;
                if constexpr (OP == BM_LHS) {
                    // (1 - i k r)(y - x).n_y / r^2 - i k n_x.n_y
                    er = mca * rinv * rinv;
                    ei = fmaf(kca, rinv, mknn);
                    gr = g_r;  // w_b G for the curl term
                    gi = g_i;
                    if constexpr (RG) {
                        // RHS: w_b G [(x - y).n_x / r - 1 + i (x - y).n_x / (k r^2)], times g(y_b)
                        const float e2r = fmaf(qR[b], rinv, -1.0f), e2i = qRk[b] * rinv * rinv;
                        const float v2r = g_r * e2r - g_i * e2i, v2i = g_r * e2i + g_i * e2r;
                        sRr = fmaf(v2r, ghat[2 * b], fmaf(-v2i, ghat[2 * b + 1], sRr));
                        sRi = fmaf(v2r, ghat[2 * b + 1], fmaf(v2i, ghat[2 * b], sRi));
                    }
                } else {
                    // -1 + (i/k)(1 - i k r)(x - y).n_x / r^2
                    er = fmaf(qb[b], rinv, -1.0f);
                    ei = qbk[b] * rinv * rinv;
                }
                vr = g_r * er - g_i * ei;
                vi = g_r * ei + g_i * er;
            } else {
                kernel4pi<OP, CLAMP>(dx, dy, dz, rr[4 * Q], rr[4 * Q + 1], rr[4 * Q + 2], ny0, ny1, ny2, k,
                                     krev, vr, vi, gr, gi);
                vr *= wy[b];
                vi *= wy[b];
                if constexpr (OP == (BM_LHS | kDk)) {  // w_b H for the curl term
                    gr *= wy[b];
                    gi *= wy[b];
                }
            }
            const float p0 = 1.0f - TriRule<Q>::x1[b] - TriRule<Q>::x2[b];
            const float p1 = TriRule<Q>::x1[b], p2 = TriRule<Q>::x2[b];
            sr[0] = fmaf(vr, p0, sr[0]); si[0] = fmaf(vi, p0, si[0]);
            sr[1] = fmaf(vr, p1, sr[1]); si[1] = fmaf(vi, p1, si[1]);
            sr[2] = fmaf(vr, p2, sr[2]); si[2] = fmaf(vi, p2, si[2]);
            if constexpr (PAIR) {
                ur[0] = fmaf(wr, p0, ur[0]); ui[0] = fmaf(wi, p0, ui[0]);
                ur[1] = fmaf(wr, p1, ur[1]); ui[1] = fmaf(wi, p1, ui[1]);
                ur[2] = fmaf(wr, p2, ur[2]); ui[2] = fmaf(wi, p2, ui[2]);
            }
            if constexpr (base_op(OP) == BM_LHS) {
                tr9 += gr;  // gr, gi: w_b G (w_b H for the derivative)
                ti9 += gi;
            }
        }
        const float ua0 = rr[3 * Q + a];
        float ua = ua0;
        if constexpr (EOP == DLP)  // -(y - x).n_y = (x_a - y_0).n_y
            ua *= (xa0 - py[0]) * ny0 + (xa1 - py[1]) * ny1 + (xa2 - py[2]) * ny2;
        const float q0 = ua * (1.0f - TriRule<Q>::x1[a] - TriRule<Q>::x2[a]);
        const float q1 = ua * TriRule<Q>::x1[a], q2 = ua * TriRule<Q>::x2[a];
        if constexpr (PAIR) {
            const float t0 = ua0 * (1.0f - TriRule<Q>::x1[a] - TriRule<Q>::x2[a]);
            const float t1 = ua0 * TriRule<Q>::x1[a], t2 = ua0 * TriRule<Q>::x2[a];
#pragma unroll
            for (int be = 0; be < 3; ++be) {
                tr[be] = fmaf(t0, ur[be], tr[be]); tii[be] = fmaf(t0, ui[be], tii[be]);
                tr[3 + be] = fmaf(t1, ur[be], tr[3 + be]); tii[3 + be] = fmaf(t1, ui[be], tii[3 + be]);
                tr[6 + be] = fmaf(t2, ur[be], tr[6 + be]); tii[6 + be] = fmaf(t2, ui[be], tii[6 + be]);
            }
        }
#pragma unroll
        for (int be = 0; be < 3; ++be) {
            br[be] = fmaf(q0, sr[be], br[be]); bi[be] = fmaf(q0, si[be], bi[be]);
            br[3 + be] = fmaf(q1, sr[be], br[3 + be]); bi[3 + be] = fmaf(q1, si[be], bi[3 + be]);
            br[6 + be] = fmaf(q2, sr[be], br[6 + be]); bi[6 + be] = fmaf(q2, si[be], bi[6 + be]);
        }
        if constexpr (base_op(OP) == BM_LHS) {
            gr9 = fmaf(ua, tr9, gr9);
            gi9 = fmaf(ua, ti9, gi9);
        }
        if constexpr (RG) {
            rgs[0] = fmaf(q0, sRr, rgs[0]); rgs[1] = fmaf(q0, sRi, rgs[1]);
            rgs[2] = fmaf(q1, sRr, rgs[2]); rgs[3] = fmaf(q1, sRi, rgs[3]);
            rgs[4] = fmaf(q2, sRr, rgs[4]); rgs[5] = fmaf(q2, sRi, rgs[5]);
        }
    }
    if constexpr (OP == HYP) {
        // curl.curl * (sum of G) - k^2 (n_x.n_y) * (phi-weighted G) (integrand.h:304-313)
        float sgr = 0.f, sgi = 0.f;
#pragma unroll
        for (int e = 0; e < 9; ++e) { sgr += br[e]; sgi += bi[e]; }
        float ci[9], cp[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) ci[e] = rci[e];
        curl_product(ci, cj, cp);
        const float nn = rr[4 * Q] * ny0 + rr[4 * Q + 1] * ny1 + rr[4 * Q + 2] * ny2;
        const float kk = k * k * nn;
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            br[e] = sgr * cp[e] - br[e] * kk;
            bi[e] = sgi * cp[e] - bi[e] * kk;
        }
    } else if constexpr (base_op(OP) == BM_LHS) {
        // (i/k) G summed, times curl.curl, plus the phi-weighted BM kernel (integrand.h:314-323)
        const float invk = 1.0f / k;
        const float s9r = -gi9 * invk, s9i = gr9 * invk;
        float ci[9], cp[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) ci[e] = rci[e];
        curl_product(ci, cj, cp);
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            br[e] = fmaf(s9r, cp[e], br[e]);
            bi[e] = fmaf(s9i, cp[e], bi[e]);
        }
    }
    if constexpr (PAIR && OP == ADLP) {  // adjoint double layer = double layer transposed
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            const float xr = br[e], xi = bi[e];
            br[e] = tr[e];
            bi[e] = tii[e];
            tr[e] = xr;
            tii[e] = xi;
        }
    }
}

// The same block for the two Burton-Miller operators with the 3-point rule
// (the reference's approximate rules), built from sums.  The rule's points
// give every corner the basis value (1 + 3 delta)/6 (2/3 at its own point,
// 1/6 at the others), and its three weights are equal, so with u_ab the
// kernel value at row point a and column point b (column weight w folded
// in) and ua the row weight,
//   B[c][d] = sum_ab ua (1 + 3 d_ca)(1 + 3 d_db) u_ab / 36
//           = (ua / 36) (9 u_cd + 3 R_c + 3 C_d + S)
// with R, C the row and column sums of u and S its total: 28 additions and 36
// fused multiply-adds per pair instead of 117 basis-weighted accumulations.
// w = the column weight / 4 pi (all three equal), kw = k w.  RG as above.
template <int Q3>
__device__ __forceinline__ void bm3_block(const float (&uR)[Q3][Q3], const float (&uI)[Q3][Q3], float ua, float* br,
                                          float* bi)
{
    const float f = ua * (1.0f / 36.0f), f3 = 3.0f * f, f9 = 9.0f * f;
    float Rr[Q3], Ri[Q3], Cr[Q3], Ci[Q3];
#pragma unroll
    for (int a = 0; a < Q3; ++a) {
        Rr[a] = uR[a][0] + uR[a][1] + uR[a][2];
        Ri[a] = uI[a][0] + uI[a][1] + uI[a][2];
        Cr[a] = uR[0][a] + uR[1][a] + uR[2][a];
        Ci[a] = uI[0][a] + uI[1][a] + uI[2][a];
    }
    const float Sr = Rr[0] + Rr[1] + Rr[2], Si = Ri[0] + Ri[1] + Ri[2];
#pragma unroll
    for (int c = 0; c < Q3; ++c) {
        const float pr = fmaf(f3, Rr[c], f * Sr), pi = fmaf(f3, Ri[c], f * Si);
#pragma unroll
        for (int d = 0; d < Q3; ++d) {
            br[3 * c + d] = fmaf(f9, uR[c][d], fmaf(f3, Cr[d], pr));
            bi[3 * c + d] = fmaf(f9, uI[c][d], fmaf(f3, Ci[d], pi));
        }
    }
}

template <int OP, bool CLAMP, bool RG>
__device__ __forceinline__ void p1_pair_block_bm3(const float* __restrict__ rr, const float* __restrict__ rci,
                                                  const float* py, float w, float kw, float ny0, float ny1, float ny2,
                                                  const float* cj, float k, float krev, float invk, float* br,
                                                  float* bi, const float* ghat, float* rgs)
{
    static_assert(OP == BM_LHS || OP == BM_RHS, "Burton-Miller operators");
    static_assert(!RG || OP == BM_LHS, "the right-hand-side contraction rides on the Burton-Miller LHS");
    constexpr int Q = 3;
    const float nx0 = rr[4 * Q], nx1 = rr[4 * Q + 1], nx2 = rr[4 * Q + 2];
    // (x_0 - y_b).n_x per column point (constant along the row plane)
    float qn[Q];
    if constexpr (OP == BM_RHS || RG) {
#pragma unroll
        for (int b = 0; b < Q; ++b)
            qn[b] = -((py[3 * b] - rr[0]) * nx0 + (py[3 * b + 1] - rr[1]) * nx1 + (py[3 * b + 2] - rr[2]) * nx2);
    }
    float qw[Q], qkw[Q];  // BM_RHS: w qn and w qn / k
    if constexpr (OP == BM_RHS) {
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            qw[b] = w * qn[b];
            qkw[b] = qw[b] * invk;
        }
    }
    // BM_LHS: -i k w n_x.n_y once per pair
    float mknn = 0.f;
    if constexpr (OP == BM_LHS) mknn = -kw * (nx0 * ny0 + nx1 * ny1 + nx2 * ny2);
    float uR[Q][Q], uI[Q][Q];
    float gsr = 0.f, gsi = 0.f;  // BM_LHS: sum of exp(ikr)/r (the curl term)
    float tr_[RG ? Q : 1], ti_[RG ? Q : 1];  // RG: per row point, sum_b exp(ikr)/r e2_ab ghat_b
#pragma unroll
    for (int a = 0; a < Q; ++a) {
        const float xa0 = rr[3 * a], xa1 = rr[3 * a + 1], xa2 = rr[3 * a + 2];
        // -w (x_a - y_0).n_y and k w (x_a - y_0).n_y
        float mca = 0.f, kca = 0.f;
        if constexpr (OP == BM_LHS) {
            const float ca = (xa0 - py[0]) * ny0 + (xa1 - py[1]) * ny1 + (xa2 - py[2]) * ny2;
            mca = -w * ca;
            kca = kw * ca;
        }
        if constexpr (RG) tr_[a] = ti_[a] = 0.f;
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            const float dx = py[3 * b] - xa0, dy = py[3 * b + 1] - xa1, dz = py[3 * b + 2] - xa2;
            float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            if (CLAMP) r2 = fmaxf(r2, kEps2);
            const float rinv = __builtin_amdgcn_rsqf(r2);
            const float r = r2 * rinv;
            float sn, cs;
            sincos_rev(r * krev, sn, cs);
            const float gr = cs * rinv, gi = sn * rinv;  // exp(ikr)/r
            const float rinv2 = rinv * rinv;
            float er, ei;
            if constexpr (OP == BM_LHS) {
                // w [(1 - i k r)(y - x).n_y / r^2 - i k n_x.n_y]
                er = mca * rinv2;
                ei = fmaf(kca, rinv, mknn);
                gsr += gr;
                gsi += gi;
                if constexpr (RG) {
                    // RHS: (x - y).n_x / r - 1 + i (x - y).n_x / (k r^2), times g(y_b)
                    const float e2r = fmaf(qn[b], rinv, -1.0f), e2i = qn[b] * invk * rinv2;
                    const float v2r = gr * e2r - gi * e2i, v2i = gr * e2i + gi * e2r;
                    tr_[a] = fmaf(v2r, ghat[2 * b], fmaf(-v2i, ghat[2 * b + 1], tr_[a]));
                    ti_[a] = fmaf(v2r, ghat[2 * b + 1], fmaf(v2i, ghat[2 * b], ti_[a]));
                }
            } else {
                // w [-1 + (i/k)(1 - i k r)(x - y).n_x / r^2]
                er = fmaf(qw[b], rinv, -w);
                ei = qkw[b] * rinv2;
            }
            uR[a][b] = gr * er - gi * ei;
            uI[a][b] = gr * ei + gi * er;
        }
    }
    const float ua = rr[3 * Q];  // the row weight (all three equal)
    bm3_block<Q>(uR, uI, ua, br, bi);
    if constexpr (OP == BM_LHS) {
        // (i/k) w ua (sum of G) times curl.curl
        const float sc = invk * w * ua;
        const float s9r = -gsi * sc, s9i = gsr * sc;
        // a triangle's three curls sum to zero, so the 2 x 2 leading dot
        // products fix the other five
        float ci[9], cp[9];
#pragma unroll
        for (int e = 0; e < 6; ++e) ci[e] = rci[e];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int d = 0; d < 2; ++d)
                cp[3 * c + d] = ci[3 * c] * cj[3 * d] + ci[3 * c + 1] * cj[3 * d + 1] + ci[3 * c + 2] * cj[3 * d + 2];
        cp[2] = -(cp[0] + cp[1]);
        cp[5] = -(cp[3] + cp[4]);
        cp[6] = -(cp[0] + cp[3]);
        cp[7] = -(cp[1] + cp[4]);
        cp[8] = -(cp[6] + cp[7]);
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            br[e] = fmaf(s9r, cp[e], br[e]);
            bi[e] = fmaf(s9i, cp[e], bi[e]);
        }
    }
    if constexpr (RG) {
        // corner c: sum_a ua w (1 + 3 d_ca) / 6 T_a
        const float h = ua * w * (1.0f / 6.0f), h3 = 3.0f * h;
        const float Tr = tr_[0] + tr_[1] + tr_[2], Ti = ti_[0] + ti_[1] + ti_[2];
#pragma unroll
        for (int c = 0; c < Q; ++c) {
            rgs[2 * c] = fmaf(h3, tr_[c], h * Tr);
            rgs[2 * c + 1] = fmaf(h3, ti_[c], h * Ti);
        }
    }
}


template <int OP, int Q, bool CLAMP, int SWEEP>
__global__ __launch_bounds__(256) void p1_regular_kernel(const float* __restrict__ rec, const int4* __restrict__ tris,
                                                         const float* __restrict__ curls, P1Tiles tl,
                                                         const KBatch kb, float2* __restrict__ out, int64_t ld,
                                                         const int32_t* __restrict__ vmap,
                                                         const float2* __restrict__ rg_g,
                                                         float2* __restrict__ rg_out, int64_t plane)
{
    const float k = kb.k[blockIdx.z];
    out += (int64_t)blockIdx.z * kb.zstride;
    rec += (int64_t)blockIdx.z * kb.gstride;
    curls += (int64_t)blockIdx.z * kb.gstride;
    constexpr int CT = kP1ColTris;  // one lane per column triangle and per column vertex
    constexpr int NCV = kP1ColVerts;
    static_assert(CT == 256 && NCV <= CT, "one lane per column triangle and per column vertex");
    static_assert(SWEEP != kSweepSym || OP == SLP || OP == HYP, "symmetric sweep only for symmetric pair blocks");
    static_assert(SWEEP != kSweepPair || OP == DLP || OP == ADLP, "paired sweep operators");
    static_assert(SWEEP != kSweepAllRg || OP == BM_LHS, "the right-hand-side contraction rides on the LHS");
    constexpr bool HALF = SWEEP == kSweepSym || SWEEP == kSweepPair;  // one evaluation per unordered pair
    constexpr bool PAIR = SWEEP == kSweepPair;
    constexpr bool RG = SWEEP == kSweepAllRg;
    // the evaluated operator: the paired adjoint double layer evaluates the
    // double layer and swaps the blocks
    constexpr int EOP = (PAIR && OP == ADLP) ? DLP : OP;
    // the Burton-Miller operators with the 3-point rule: blocks built from sums
    constexpr bool kBm3 = Q == 3 && !PAIR && (OP == BM_LHS || OP == BM_RHS);
    constexpr int S = Rec<Q>::kFloats;
    __shared__ float2 acc[kP1RowVerts][NCV];
    __shared__ float2 stage[9][CT];
    // the row tile's triangles, records and curls, read by every wavefront at
    // each row step (from LDS rather than a chain of dependent scalar loads)
    __shared__ float rrec[kP1RowTris][S];
    __shared__ int4 rtri[kP1RowTris];
    __shared__ int4 rslot[kP1RowTris];  // the row triangles' vertex slots in the row tile
    __shared__ float rcurl[kP1RowTris][9];
    // RG: the right-hand-side contraction per row-tile vertex slot, summed
    // over the workgroup's columns, then stored as this column tile's partial
    __shared__ float2 rgslot[RG ? kP1RowVerts : 1];
    int ct = blockIdx.x, rt = blockIdx.y;
    bool own = false;  // half sweeps: the column tile holds the row tile
    if constexpr (HALF) {
        const int64_t wi = blockIdx.x;
        int lo = 0, hi = tl.nrow_tiles - 1;  // last r with sym_ptr[r] <= wi
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tl.sym_ptr[mid] <= wi) lo = mid;
            else hi = mid - 1;
        }
        rt = lo;
        ct = tl.row_ctile[rt] + (int)(wi - tl.sym_ptr[rt]);
        own = ct == tl.row_ctile[rt];
    }
    const int c0 = tl.col_ptr[ct], c1 = tl.col_ptr[ct + 1];
    const int r0 = tl.row_ptr[rt], r1 = tl.row_ptr[rt + 1];
    const int rv0 = tl.row_vptr[rt], nrv = tl.row_vptr[rt + 1] - rv0;
    const int cv0 = tl.col_vptr[ct], ncv = tl.col_vptr[ct + 1] - cv0;
    for (int e = threadIdx.x; e < nrv * ncv; e += CT) acc[e / ncv][e % ncv] = make_float2(0.f, 0.f);
    if constexpr (RG)
        if (threadIdx.x < kP1RowVerts) rgslot[threadIdx.x] = make_float2(0.f, 0.f);
    for (int e = threadIdx.x; e < (r1 - r0) * S; e += CT) {  // r1 - r0 <= kP1RowTris
        const int q = e / S;
        rrec[q][e - q * S] = rec[(int64_t)tl.perm[r0 + q] * S + (e - q * S)];
    }
    for (int q = threadIdx.x; q < r1 - r0; q += CT) {
        rtri[q] = tris[tl.perm[r0 + q]];
        rslot[q] = tl.row_slot[r0 + q];
    }
    if constexpr (OP == HYP || base_op(OP) == BM_LHS)
        for (int e = threadIdx.x; e < (r1 - r0) * 9; e += CT) {
            const int q = e / 9;
            rcurl[q][e - q * 9] = curls[(int64_t)tl.perm[r0 + q] * 12 + (e - q * 9)];
        }

    // phase-1 role: column triangle c0 + threadIdx.x
    const int pos = c0 + (int)threadIdx.x;
    const bool active = pos < c1;
    const int j = tl.perm[active ? pos : c0];
    const float* cr = rec + (int64_t)j * S;
    float py[3 * Q], wy[Q];
#pragma unroll
    for (int b = 0; b < 3 * Q; ++b) py[b] = cr[b];
#pragma unroll
    for (int b = 0; b < Q; ++b) wy[b] = cr[3 * Q + b] * kInv4Pi;
    const float ny0 = cr[4 * Q], ny1 = cr[4 * Q + 1], ny2 = cr[4 * Q + 2];
    const int4 tj = tris[j];
    float cj[9];
    if constexpr (OP == HYP || base_op(OP) == BM_LHS) {
#pragma unroll
        for (int e = 0; e < 9; ++e) cj[e] = curls[(int64_t)j * 12 + e];
    }
    const float krev = k * kInv2Pi;
    const float invk = 1.0f / k;
    if constexpr (RG) {  // this launch's vertex vector and partial rows
        rg_g += (int64_t)blockIdx.z * kb.vstride;
        rg_out += (int64_t)blockIdx.z * tl.ncol_tiles * (int64_t)tl.row_vptr[tl.nrow_tiles];
    }
    // RG: g at the column triangle's quadrature points (caller's vertex ids)
    float ghat[RG ? 2 * Q : 1];
    float2 rgacc = make_float2(0.f, 0.f);  // RG: lane 3 r + c keeps row triangle r's corner c
    if constexpr (RG) {
        const float2 g0 = rg_g[tj.x], g1 = rg_g[tj.y], g2 = rg_g[tj.z];
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            const float p1 = TriRule<Q>::x1[b], p2 = TriRule<Q>::x2[b], p0 = 1.0f - p1 - p2;
            ghat[2 * b] = p0 * g0.x + p1 * g1.x + p2 * g2.x;
            ghat[2 * b + 1] = p0 * g0.y + p1 * g1.y + p2 * g2.y;
        }
    }
    // phase-2 role: column vertex slot w = threadIdx.x
    const int w = (int)threadIdx.x;
    const bool owner = w < ncv;
    const int ib = owner ? tl.col_inc_ptr[cv0 + w] : 0;
    const int ie = owner ? tl.col_inc_ptr[cv0 + w + 1] : 0;
    // the owner's first kIncRegs incidence codes in registers for the whole
    // workgroup (read from memory at every row triangle they left the 3-point
    // kernels waiting on the loads)
    constexpr int kIncRegs = 8;
    int cinc[kIncRegs];
#pragma unroll
    for (int u = 0; u < kIncRegs; ++u) cinc[u] = ib + u < ie ? tl.col_inc[ib + u] : 0;
    // s_c = the sum of the stage entries of row corner c over the owner's
    // incident column triangles, in incidence order (an owner has at least
    // one: its vertex belongs to a triangle of the tile)
    auto gather = [&](float2& s0, float2& s1, float2& s2) {
        {
            const int p = cinc[0] >> 2, b = cinc[0] & 3;
            s0 = stage[b][p];
            s1 = stage[3 + b][p];
            s2 = stage[6 + b][p];
        }
#pragma unroll
        for (int u = 1; u < kIncRegs; ++u) {
            if (ib + u < ie) {
                const int code = cinc[u];
                const int p = code >> 2, b = code & 3;
                const float2 e0 = stage[b][p], e1 = stage[3 + b][p], e2 = stage[6 + b][p];
                s0.x += e0.x; s0.y += e0.y;
                s1.x += e1.x; s1.y += e1.y;
                s2.x += e2.x; s2.y += e2.y;
            }
        }
        for (int q = ib + kIncRegs; q < ie; ++q) {
            const int code = tl.col_inc[q];
            const int p = code >> 2, b = code & 3;
            const float2 e0 = stage[b][p], e1 = stage[3 + b][p], e2 = stage[6 + b][p];
            s0.x += e0.x; s0.y += e0.y;
            s1.x += e1.x; s1.y += e1.y;
            s2.x += e2.x; s2.y += e2.y;
        }
    };
    // PAIR: this lane's column of B(j, i)^T by row slot, real and imaginary
    float accTr[PAIR ? kP1RowVerts : 1], accTi[PAIR ? kP1RowVerts : 1];
#pragma unroll
    for (int q = 0; q < (PAIR ? kP1RowVerts : 1); ++q) accTr[q] = accTi[q] = 0.f;
    __syncthreads();

    for (int rp = r0; rp < r1; ++rp) {
        const int4 ti = rtri[rp - r0];
        const bool shared = ti.x == tj.x || ti.x == tj.y || ti.x == tj.z || ti.y == tj.x || ti.y == tj.y ||
                            ti.y == tj.z || ti.z == tj.x || ti.z == tj.y || ti.z == tj.z || (HALF && own && pos <= rp);
        float br[9], bi[9];
        float tr[PAIR ? 9 : 1], tii[PAIR ? 9 : 1];  // PAIR: B(j, i)^T
        // the sum-built blocks are assigned whole: skipped lanes stage zeros
        // from a branch of their own instead of clearing 18 registers at
        // every row triangle (the kernels are bound by vector issue)
        if constexpr (!kBm3) {
#pragma unroll
            for (int e = 0; e < 9; ++e) br[e] = bi[e] = 0.f;
        }
#pragma unroll
        for (int e = 0; e < (PAIR ? 9 : 1); ++e) tr[e] = tii[e] = 0.f;
        float rgs[RG ? 6 : 1];
#pragma unroll
        for (int e = 0; e < (RG ? 6 : 1); ++e) rgs[e] = 0.f;
        const bool eval = active && !shared;
        if (eval) {
            float ci[9];
            if constexpr (OP == HYP || base_op(OP) == BM_LHS) {
#pragma unroll
                for (int e = 0; e < 9; ++e) ci[e] = rcurl[rp - r0][e];
            }
            if constexpr (kBm3)
                p1_pair_block_bm3<OP, CLAMP, RG>(rrec[rp - r0], ci, py, wy[0], k * wy[0], ny0, ny1, ny2, cj, k, krev,
                                                 invk, br, bi, ghat, rgs);
            else
                p1_pair_block<OP, Q, CLAMP, PAIR, EOP, RG>(rrec[rp - r0], ci, py, wy, ny0, ny1, ny2, cj, k, krev,
                                                            invk, br, bi, tr, tii, ghat, rgs);
        }
        if constexpr (RG) {  // the row triangle's three corner sums over this wavefront's columns
            const int lane = threadIdx.x & 63;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float sr = wave_sum_f32(rgs[2 * c]), si = wave_sum_f32(rgs[2 * c + 1]);
                if (lane == 3 * (rp - r0) + c) rgacc = make_float2(sr, si);
            }
        }
        if (!kBm3 || eval) {
#pragma unroll
            for (int e = 0; e < 9; ++e) stage[e][threadIdx.x] = make_float2(br[e], bi[e]);
        } else {
#pragma unroll
            for (int e = 0; e < 9; ++e) stage[e][threadIdx.x] = make_float2(0.f, 0.f);
        }
        __syncthreads();
        if (owner) {
            const int4 rs = rslot[rp - r0];
            float2 s0, s1, s2;
            gather(s0, s1, s2);
            acc[rs.x][w].x += s0.x; acc[rs.x][w].y += s0.y;
            acc[rs.y][w].x += s1.x; acc[rs.y][w].y += s1.y;
            acc[rs.z][w].x += s2.x; acc[rs.z][w].y += s2.y;
        }
        __syncthreads();
        if constexpr (PAIR) {  // the transposed block through the same stage
#pragma unroll
            for (int e = 0; e < 9; ++e) stage[e][threadIdx.x] = make_float2(tr[e], tii[e]);
            __syncthreads();
            const int4 rs = rslot[rp - r0];
            float2 s0 = make_float2(0.f, 0.f), s1 = s0, s2 = s0;
            if (owner) gather(s0, s1, s2);
            // the slots are the same on every lane: indexed register moves
            // (a select per slot costs 8 vector instructions x 20 slots)
            const int sx = __builtin_amdgcn_readfirstlane(rs.x), sy = __builtin_amdgcn_readfirstlane(rs.y),
                      sz = __builtin_amdgcn_readfirstlane(rs.z);
            accTr[sx] += s0.x; accTi[sx] += s0.y;
            accTr[sy] += s1.x; accTi[sy] += s1.y;
            accTr[sz] += s2.x; accTi[sz] += s2.y;
            __syncthreads();
        }
    }
    for (int e = threadIdx.x; e < nrv * ncv; e += CT) {
        const int rvi = e / ncv, cvi = e - rvi * ncv;
        // out is in the tiles' P1 numbering, or (vmap) in the caller's
        int64_t gr = tl.row_verts[rv0 + rvi], gc = tl.col_verts[cv0 + cvi];
        if (vmap) {
            gr = vmap[gr];
            gc = vmap[gc];
        }
        const float2 v = acc[rvi][cvi];
        if (plane) {  // real and imaginary planes: a wavefront's 64 atomics span half the cache lines
            float* dst = reinterpret_cast<float*>(out) + (gr * ld + gc);
            unsafeAtomicAdd(dst, v.x);
            unsafeAtomicAdd(dst + plane, v.y);
        } else {
            float* dst = reinterpret_cast<float*>(&out[gr * ld + gc]);
            unsafeAtomicAdd(dst, v.x);
            unsafeAtomicAdd(dst + 1, v.y);
        }
    }
    if constexpr (RG) {
        // the wavefronts' corner sums into the row tile's vertex slots (LDS),
        // then one plain store per slot into this column tile's partial row
        // (rg_out [ncol_tiles, row slots]; p1_rg_reduce_kernel sums the
        // column tiles in a fixed order): no global atomics, which at level 6
        // (every row vertex hit from 320 column tiles) serialised the kernel
        const int lane = threadIdx.x & 63;
        if (lane < 3 * (r1 - r0)) {
            const int4 rs = rslot[lane / 3];
            const int c = lane % 3;
            const int slot = c == 0 ? rs.x : (c == 1 ? rs.y : rs.z);
            atomicAdd(&rgslot[slot].x, rgacc.x);
            atomicAdd(&rgslot[slot].y, rgacc.y);
        }
        __syncthreads();
        const int64_t nslots = tl.row_vptr[tl.nrow_tiles];
        if ((int)threadIdx.x < nrv) rg_out[(int64_t)ct * nslots + rv0 + threadIdx.x] = rgslot[threadIdx.x];
    }
    if constexpr (PAIR) {
        // B(j, i)^T lands at the transposed positions of out: the owners'
        // registers through the LDS accumulator (free once flushed), then a
        // flush with consecutive lanes along the row-tile slots of one column
        // vertex (a few cache lines per wavefront in the tiles' numbering)
        __syncthreads();
        if (owner) {
#pragma unroll
            for (int q = 0; q < kP1RowVerts; ++q)
                if (q < nrv) acc[q][w] = make_float2(accTr[q], accTi[q]);
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nrv * ncv; e += CT) {
            const int cvi = e / nrv, rvi = e - cvi * nrv;
            int64_t gr = tl.row_verts[rv0 + rvi], gc = tl.col_verts[cv0 + cvi];
            if (vmap) {
                gr = vmap[gr];
                gc = vmap[gc];
            }
            const float2 v = acc[rvi][cvi];
            if (plane) {
                float* dst = reinterpret_cast<float*>(out) + (gc * ld + gr);
                unsafeAtomicAdd(dst, v.x);
                unsafeAtomicAdd(dst + plane, v.y);
            } else {
                float* dst = reinterpret_cast<float*>(&out[gc * ld + gr]);
                unsafeAtomicAdd(dst, v.x);
                unsafeAtomicAdd(dst + 1, v.y);
            }
        }
    }
}

// out <- out + out^T in place on an n x n matrix (the one-sided regular part
// of a symmetric P1 operator, before the neighbour pairs are added).  One
// workgroup per 32 x 32 tile pair (I <= J): both tiles through LDS, written
// back as A_IJ + A_JI^T and its transpose.
constexpr int kSymT = 32;
__global__ __launch_bounds__(256) void p1_symmetrize_kernel(float2* __restrict__ out, int64_t n, int64_t ld,
                                                            int64_t zstride)
{
    out += (int64_t)blockIdx.z * zstride;
    __shared__ float2 a[kSymT][kSymT + 1], b[kSymT][kSymT + 1];
    int64_t I, J;
    tri_block_pair(blockIdx.x, I, J);
    const int64_t r0 = I * kSymT, c0 = J * kSymT;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 8 rows per pass
    for (int rr = ty; rr < kSymT; rr += 8) {
        const int64_t ra = r0 + rr, ca = c0 + tx, rb = c0 + rr, cb = r0 + tx;
        a[rr][tx] = (ra < n && ca < n) ? out[ra * ld + ca] : make_float2(0.f, 0.f);
        b[rr][tx] = (rb < n && cb < n) ? out[rb * ld + cb] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    for (int rr = ty; rr < kSymT; rr += 8) {
        const int64_t ra = r0 + rr, ca = c0 + tx, rb = c0 + rr, cb = r0 + tx;
        const float2 va = a[rr][tx], vb = b[tx][rr];  // A(rr, tx) + B(tx, rr)
        const float2 s = make_float2(va.x + vb.x, va.y + vb.y);
        if (ra < n && ca < n) out[ra * ld + ca] = s;
        if (I != J) {
            const float2 wb = b[rr][tx], wa = a[tx][rr];  // B(rr, tx) + A(tx, rr)
            if (rb < n && cb < n) out[rb * ld + cb] = make_float2(wb.x + wa.x, wb.y + wa.y);
        }
    }
}

// ---------------------------------------------------------------------------
// neighbour (singular) pairs: Sauter-Schwab, one wavefront per pair
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// sum over aligned groups of N lanes (N a power of two <= 64), in every lane
// (16 and 64: DPP row moves, bem_device.hpp; others: cross-lane permutes)
template <int N>
__device__ __forceinline__ double group_sum(double v)
{
    if constexpr (N == 16) {
        return row_sum16(v);
    } else if constexpr (N == 64) {
        return wave_sum64(v);
    } else {
#pragma unroll
        for (int off = N / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        return v;
    }
}

// MODE 0: write (DP0: out = alpha * entry) or add (P1) the entries;
// MODE 1 (DP0): vpsi[p] = entry * psi[j], the pair's term of V psi.
template <int OP, int NL, int BASIS, bool CLAMP, int MODE = 0>
__global__ __launch_bounds__(256) void singular_kernel(const float* __restrict__ geo, const int4* __restrict__ tris,
                                                       const float* __restrict__ curls,
                                                       const int4* __restrict__ pairs, int64_t p0, int64_t p1,
                                                       const KBatch kb, float2* __restrict__ out, int64_t ld,
                                                       int64_t row0, float alpha, const float2* __restrict__ psi,
                                                       float2* __restrict__ vpsi)
{
    const float k = kb.k[blockIdx.z];
    out += (int64_t)blockIdx.z * kb.zstride;
    geo += (int64_t)blockIdx.z * kb.gstride;
    curls += (int64_t)blockIdx.z * kb.gstride;
    if constexpr (MODE == 2) {
        psi += (int64_t)blockIdx.z * kb.vstride;
        vpsi += (int64_t)blockIdx.z * kb.vstride;
    }
    const int lane = threadIdx.x & 63;
    // lanes per pair: a wavefront per pair, except the 2-point line rule of the
    // P1 operators (16 points): 16 lanes per pair, 4 pairs per wavefront, so
    // the FP64 reductions run over 16 lanes and serve four pairs
    constexpr int LPP = (NL == 2 && BASIS == P1 && MODE != 1) ? 16 : 64;
    constexpr int PPW = 64 / LPP;
    const int ll = lane % LPP;
    const int64_t pw = p0 + ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * PPW;  // the wavefront's first pair
    if (pw >= p1) return;  // uniform per wavefront
    const int64_t p = pw + lane / LPP;
    const bool valid = p < p1;  // groups past the end compute on the last pair and write nothing
    const int4 pr = pairs[valid ? p : p1 - 1];
    const int i = PPW == 1 ? __builtin_amdgcn_readfirstlane(pr.x) : pr.x;
    const int j = PPW == 1 ? __builtin_amdgcn_readfirstlane(pr.y) : pr.y;
    const int shared = PPW == 1 ? __builtin_amdgcn_readfirstlane(pr.z) : pr.z;
    const int4 ti = tris[i], tj = tris[j];
    const int ia[3] = {ti.x, ti.y, ti.z}, ja[3] = {tj.x, tj.y, tj.z};
    int pi[3] = {0, 1, 2}, pj[3] = {0, 1, 2};
    if (shared == 2) {
        // common edge first, same order on both triangles, ordered by position
        // in the column triangle (integrand.h:351-364)
        int c = 0;
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a)
                if (ia[a] == ja[b]) { pi[c] = a; pj[c] = b; ++c; }
        pi[2] = 3 - pi[0] - pi[1];
        pj[2] = 3 - pj[0] - pj[1];
    } else if (shared == 1) {
        // common vertex swapped to local 0 (integrand.h:365-382)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                if (ia[a] == ja[b]) {
                    if (a != 0) { pi[0] = a; pi[a] = 0; }
                    if (b != 0) { pj[0] = b; pj[b] = 0; }
                }
    }
    // geo record: v0 v1 v2 (9), n (3), jac (1), pad
    const float* gi = geo + (int64_t)i * 16;
    const float* gj = geo + (int64_t)j * 16;
    float v[3][3], w[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            v[a][d] = gi[3 * pi[a] + d];
            w[a][d] = gj[3 * pj[a] + d];
        }
    // x = v0 + x1 (v1 - v0) + x2 (v2 - v1) on (0,0),(1,0),(1,1); the shared
    // local vertex 0 (v0 = w0 in all three cases) cancels in d = y - x.
    float e1[3], e2[3], f1[3], f2[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        e1[d] = v[1][d] - v[0][d];
        e2[d] = v[2][d] - v[1][d];
        f1[d] = w[1][d] - w[0][d];
        f2[d] = w[2][d] - w[1][d];
    }
    const float nx0 = gi[9], nx1 = gi[10], nx2 = gi[11];
    const float ny0 = gj[9], ny1 = gj[10], ny2 = gj[11];
    const float jij = gi[12] * gj[12];
    const float krev = k * kInv2Pi;

    double ar[10], ai[10];
#pragma unroll
    for (int e = 0; e < 10; ++e) ar[e] = ai[e] = 0.0;

    constexpr int NP = NL * NL * NL * NL;
    // The pair kind (shared vertices) picks a fully unrolled region loop.
    ss_dispatch(shared, [&](auto kind) {
        constexpr int S = decltype(kind)::value;
        constexpr int NR = kSSRegions<S>;
        for (int q = ll; q < NP; q += LPP) {
            const int q0 = q / (NL * NL * NL), q1 = (q / (NL * NL)) % NL, q2 = (q / NL) % NL, q3 = q % NL;
            const float wq =
                LineRule01<NL>::w[q0] * LineRule01<NL>::w[q1] * LineRule01<NL>::w[q2] * LineRule01<NL>::w[q3];
            float reg[NR][7];
            const float jac = ss_regions<S>(LineRule01<NL>::t[q0], LineRule01<NL>::t[q1], LineRule01<NL>::t[q2],
                                            LineRule01<NL>::t[q3], reg);
            float br[10], bi[10];
#pragma unroll
            for (int e = 0; e < 10; ++e) br[e] = bi[e] = 0.f;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const float x1 = reg[r][0], x2 = reg[r][1], y1 = reg[r][2], y2 = reg[r][3];
                float dx, dy, dz;
                ss_disp<S>(reg[r], e1, e2, f1, f2, dx, dy, dz);
                float vr, vi, gr = 0.f, gg = 0.f;
                kernel4pi<OP, CLAMP>(dx, dy, dz, nx0, nx1, nx2, ny0, ny1, ny2, k, krev, vr, vi, gr, gg);
                const float ww = reg[r][4];
                vr *= ww;
                vi *= ww;
                if constexpr (BASIS == P1) {
                    const float px[3] = {1.f - x1, x1 - x2, x2};
                    const float py[3] = {1.f - y1, y1 - y2, y2};
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int b = 0; b < 3; ++b) {
                            const float pp = px[a] * py[b];
                            br[3 * a + b] = fmaf(vr, pp, br[3 * a + b]);
                            bi[3 * a + b] = fmaf(vi, pp, bi[3 * a + b]);
                        }
                }
                if constexpr (base_op(OP) == BM_LHS) {
                    br[9] = fmaf(gr, ww, br[9]);
                    bi[9] = fmaf(gg, ww, bi[9]);
                } else {
                    br[9] += vr;
                    bi[9] += vi;
                }
            }
            const double sc = (double)(jac * wq) * (double)jij;
#pragma unroll
            for (int e = 0; e < 10; ++e) {
                ar[e] += sc * (double)br[e];
                ai[e] += sc * (double)bi[e];
            }
        }
    });
#pragma unroll
    for (int e = 0; e < 10; ++e) {
        ar[e] = group_sum<LPP>(ar[e]) * (double)kInv4Pi;
        ai[e] = group_sum<LPP>(ai[e]) * (double)kInv4Pi;
    }
    if constexpr (MODE == 1) {
        if (lane == 0) {
            const float2 pj = psi[j];
            vpsi[p] = make_float2((float)(ar[9] * pj.x - ai[9] * pj.y), (float)(ar[9] * pj.y + ai[9] * pj.x));
        }
    } else if constexpr (BASIS == DP0) {
        if (lane == 0)
            out[((int64_t)i - row0) * ld + j] = make_float2((float)(alpha * ar[9]), (float)(alpha * ai[9]));
    } else if constexpr (MODE == 2) {
        // P1 contraction (the operator is not stored): entry (a, b) times psi
        // at its column vertex; the three lanes 3a, 3a + 1, 3a + 2 of the
        // pair's group are summed, so each pair adds three values (one per
        // row corner) into vpsi instead of nine (on small meshes every row
        // vertex takes hundreds of these atomics)
        static_assert(OP == BM_RHS, "the contraction is the Burton-Miller right-hand side's");
        float cr = 0.f, cim = 0.f;
        int oa = 0;
        if (valid && ll < 9) {
            const int a = ll / 3, b = ll % 3;
            double vr = 0.0, vi = 0.0;
#pragma unroll
            for (int e = 0; e < 9; ++e)
                if (e == ll) { vr = ar[e]; vi = ai[e]; }
            oa = pi[a];
            const float2 gv = psi[ja[pj[b]]];
            cr = (float)(vr * gv.x - vi * gv.y);
            cim = (float)(vr * gv.y + vi * gv.x);
        }
        // every lane takes part in the shuffles; lanes 3a + 1, 3a + 2 stay in the group (ll <= 8 < 16)
        const float r1 = __shfl_down(cr, 1, 64), i1 = __shfl_down(cim, 1, 64);
        const float r2 = __shfl_down(cr, 2, 64), i2 = __shfl_down(cim, 2, 64);
        if (valid && ll < 9 && ll % 3 == 0) {
            float* dst = reinterpret_cast<float*>(&vpsi[ia[oa]]);
            unsafeAtomicAdd(dst, cr + r1 + r2);
            unsafeAtomicAdd(dst + 1, cim + i1 + i2);
        }
    } else {
        if (valid && ll < 9) {
            // ll = reordered (a, b) pair; its original local indices are (pi[a], pj[b])
            const int a = ll / 3, b = ll % 3;
            double vr = 0.0, vi = 0.0;
#pragma unroll
            for (int e = 0; e < 9; ++e)
                if (e == ll) { vr = ar[e]; vi = ai[e]; }
            const int oa = pi[a], ob = pj[b];
            if constexpr (OP == HYP || base_op(OP) == BM_LHS) {
                float ci[9], cj[9], cp[9];
#pragma unroll
                for (int e = 0; e < 9; ++e) {
                    ci[e] = curls[(int64_t)i * 12 + e];
                    cj[e] = curls[(int64_t)j * 12 + e];
                }
                curl_product(ci, cj, cp);
                const double c = cp[3 * oa + ob];
                if constexpr (OP == HYP) {
                    const double kk = (double)k * (double)k * (double)(nx0 * ny0 + nx1 * ny1 + nx2 * ny2);
                    // sum over all 9 reordered entries = plain-G integral (phi sums to 1)
                    vr = ar[9] * c - vr * kk;
                    vi = ai[9] * c - vi * kk;
                } else {
                    const double invk = 1.0 / (double)k;
                    vr += -ai[9] * invk * c;
                    vi += ar[9] * invk * c;
                }
            }
            float* dst = reinterpret_cast<float*>(&out[(int64_t)ia[oa] * ld + ja[ob]]);
            unsafeAtomicAdd(dst, (float)vr);
            unsafeAtomicAdd(dst + 1, (float)vi);
        }
    }
}

// rhs[caller id of slot s] += sum over column tiles (fixed order) of
// part[ct][s]; one thread per row-tile vertex slot (a vertex owns one slot in
// each row tile that touches it, so the atomics meet at most a few times)
__global__ __launch_bounds__(256) void p1_rg_reduce_kernel(const float2* __restrict__ part, int ncol_tiles,
                                                           int64_t nslots, const int32_t* __restrict__ row_verts,
                                                           const int32_t* __restrict__ vold,
                                                           float2* __restrict__ rhs, int64_t vstride)
{
    const int64_t sl = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (sl >= nslots) return;
    part += (int64_t)blockIdx.y * ncol_tiles * nslots;  // system blockIdx.y
    rhs += (int64_t)blockIdx.y * vstride;
    float sx = 0.f, sy = 0.f;
    for (int c = 0; c < ncol_tiles; ++c) {
        const float2 v = part[(int64_t)c * nslots + sl];
        sx += v.x;
        sy += v.y;
    }
    float* dst = reinterpret_cast<float*>(&rhs[vold[row_verts[sl]]]);
    unsafeAtomicAdd(dst, sx);
    unsafeAtomicAdd(dst + 1, sy);
}

// out[v] += alpha (M g)[v] for the P1 mass matrix M_ab = |c| (1 + delta_ab) / 24
// per triangle (the closed form of identity_p1_kernel's entries); one thread
// per triangle, atomics into the three corners.
struct AlphaBatch {
    float2 a[kMaxBatchK];
};

// system z = blockIdx.y: geometry z (gstride floats on), g and out z * vstride
// elements on, factor ab.a[z]
__global__ __launch_bounds__(256) void p1_mass_apply_kernel(const float* __restrict__ geo, int64_t gstride,
                                                            const int4* __restrict__ tris, int64_t nt,
                                                            const float2* __restrict__ g, const AlphaBatch ab,
                                                            float2* __restrict__ out, int64_t vstride)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nt) return;
    const int z = blockIdx.y;
    geo += (int64_t)z * gstride;
    g += (int64_t)z * vstride;
    out += (int64_t)z * vstride;
    const float2 alpha = ab.a[z];
    const int4 tr = tris[t];
    const float c = geo[t * 16 + 12] * (1.0f / 24.0f);
    const float2 g0 = g[tr.x], g1 = g[tr.y], g2 = g[tr.z];
    const float sx = g0.x + g1.x + g2.x, sy = g0.y + g1.y + g2.y;
    const int v3[3] = {tr.x, tr.y, tr.z};
    const float2 gv[3] = {g0, g1, g2};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float mx = c * (sx + gv[a].x), my = c * (sy + gv[a].y);  // (M g)_a from this triangle
        float* dst = reinterpret_cast<float*>(&out[v3[a]]);
        unsafeAtomicAdd(dst, alpha.x * mx - alpha.y * my);
        unsafeAtomicAdd(dst + 1, alpha.x * my + alpha.y * mx);
    }
}

// ---------------------------------------------------------------------------
// mass matrix, triangle -> vertex average, potentials
// ---------------------------------------------------------------------------
// P1 mass matrix row va = sum over triangles T around va of |T|/12 (1 + delta)
// (closed form of integrand.h:257-278); one thread per row, so each entry has
// one writer and a fixed summation order.
__global__ __launch_bounds__(256) void identity_p1_kernel(const float* __restrict__ geo, const int4* __restrict__ tris,
                                                          const int32_t* __restrict__ vt_ptr,
                                                          const int32_t* __restrict__ vt_idx, int64_t nv,
                                                          float2* __restrict__ out, int64_t ld)
{
    const int64_t va = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (va >= nv) return;
    for (int e = vt_ptr[va]; e < vt_ptr[va + 1]; ++e) {
        const int t = vt_idx[e];
        const float area = 0.5f * geo[(int64_t)t * 16 + 12];
        const int4 tv = tris[t];
        const int vs[3] = {tv.x, tv.y, tv.z};
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            float2* dst = &out[va * ld + vs[b]];
            const float add = area * (vs[b] == va ? (2.0f / 12.0f) : (1.0f / 12.0f));
            dst->x += add;
        }
    }
}

__global__ __launch_bounds__(256) void identity_dp0_kernel(const float* __restrict__ geo, int64_t nt,
                                                           float2* __restrict__ out, int64_t ld)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nt) return;
    out[t * ld + t] = make_float2(0.5f * geo[t * 16 + 12], 0.f);
}

// vertex value = mean of the values of the triangles around it (bind.cu:23-48)
__global__ __launch_bounds__(256) void tri2vert_kernel(const float2* __restrict__ tv, const int32_t* __restrict__ vt_ptr,
                                                       const int32_t* __restrict__ vt_idx, int64_t nv,
                                                       float2* __restrict__ out)
{
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= nv) return;
    float sr = 0.f, si = 0.f;
    const int e0 = vt_ptr[v], e1 = vt_ptr[v + 1];
    for (int e = e0; e < e1; ++e) {
        const float2 x = tv[vt_idx[e]];
        sr += x.x;
        si += x.y;
    }
    const float n = (float)(e1 - e0);
    out[v] = make_float2(sr / n, si / n);
}

// Stateless triangle -> vertex average (no mesh handle): the vertex ->
// triangle lists are built here.  work (int32): count/cursor [nv], ptr [nv+1],
// idx [3 nt], flag [1].  The fill order of a vertex's list depends on the
// atomics, so the gather sums each list in ascending triangle order (a
// repeated triangle, from a degenerate one, counts with its multiplicity):
// the result does not depend on the order.
__global__ __launch_bounds__(256) void fill_i32_kernel(int32_t* __restrict__ out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) out[i] = 0;
}

__global__ __launch_bounds__(256) void t2v_count_kernel(const int32_t* __restrict__ tris, int64_t nt, int64_t nv,
                                                        int32_t* __restrict__ cnt, int32_t* __restrict__ flag)
{
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < 3 * nt; e += (int64_t)gridDim.x * 256) {
        const int32_t v = tris[e];
        if (v < 0 || v >= nv) {
            atomicOr(flag, 1);
            continue;
        }
        atomicAdd(&cnt[v], 1);
    }
}

// exclusive scan of cnt [nv] into ptr [nv+1] by one workgroup of 1024
// threads (each owns a contiguous chunk); cnt becomes the fill cursor
__global__ __launch_bounds__(1024) void t2v_scan_kernel(int32_t* __restrict__ cnt, int64_t nv,
                                                        int32_t* __restrict__ ptr)
{
    __shared__ int32_t part[1024];
    const int tid = threadIdx.x;
    const int64_t chunk = (nv + 1023) / 1024;
    const int64_t b = tid * chunk, e = b + chunk < nv ? b + chunk : nv;
    int32_t sum = 0;
    for (int64_t i = b; i < e; ++i) sum += cnt[i];
    part[tid] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int32_t add = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += add;
        __syncthreads();
    }
    int32_t run = part[tid] - sum;
    for (int64_t i = b; i < e; ++i) {
        const int32_t c = cnt[i];
        ptr[i] = run;
        cnt[i] = run;
        run += c;
    }
    if (tid == 1023) ptr[nv] = part[1023];
}

__global__ __launch_bounds__(256) void t2v_fill_kernel(const int32_t* __restrict__ tris, int64_t nt, int64_t nv,
                                                       int32_t* __restrict__ cursor, int32_t* __restrict__ idx)
{
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < 3 * nt; e += (int64_t)gridDim.x * 256) {
        const int32_t v = tris[e];
        if (v < 0 || v >= nv) continue;
        idx[atomicAdd(&cursor[v], 1)] = (int32_t)(e / 3);
    }
}

__global__ __launch_bounds__(256) void t2v_gather_kernel(const float2* __restrict__ tv, const int32_t* __restrict__ ptr,
                                                         const int32_t* __restrict__ idx, int64_t nv,
                                                         float2* __restrict__ out)
{
    for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < nv; v += (int64_t)gridDim.x * 256) {
        const int e0 = ptr[v], e1 = ptr[v + 1];
        float sr = 0.f, si = 0.f;
        int32_t last = -1;
        for (int done = e0; done < e1;) {
            // next larger triangle index in the list and its multiplicity
            int32_t next = 0x7fffffff;
            for (int e = e0; e < e1; ++e) {
                const int32_t t = idx[e];
                if (t > last && t < next) next = t;
            }
            int mult = 0;
            for (int e = e0; e < e1; ++e) mult += (idx[e] == next);
            const float2 x = tv[next];
            for (int m = 0; m < mult; ++m) {
                sr += x.x;
                si += x.y;
            }
            done += mult;
            last = next;
        }
        const float n = (float)(e1 - e0);
        out[v] = make_float2(sr / n, si / n);
    }
}

// Surface -> point operator (bind.cu:156-176, integrand.h:237-255,396-406):
// row p = evaluation point (n_x = 0), column = triangle (DP0) or its three
// vertices (P1, summed with atomics into a zeroed output).
// P1 with GATHER: the three corner values of each (point, triangle) pair go
// to a scratch array [np, nt, 3] (out, zstride per wavenumber) for
// p1_point_gather_kernel; without it they are added to out [np, ld] with
// float atomics (six per pair).
template <int OP, int BASIS, bool CLAMP, bool GATHER = false>
__global__ __launch_bounds__(256) void potential_kernel(const float* __restrict__ rec6, const int4* __restrict__ tris,
                                                        int64_t nt, const float* __restrict__ pts, int64_t np,
                                                        const KBatch kb, float2* __restrict__ out, int64_t ld)
{
    const float k = kb.k[blockIdx.z];
    out += (int64_t)blockIdx.z * kb.zstride;
    rec6 += (int64_t)blockIdx.z * kb.gstride;
    pts += (int64_t)blockIdx.z * kb.pstride;
    constexpr int Q = 6;
    constexpr int S = Rec<Q>::kFloats;
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool active = j < nt;
    const float* cr = rec6 + (active ? j : nt - 1) * S;
    float py[3 * Q], wy[Q];
#pragma unroll
    for (int b = 0; b < 3 * Q; ++b) py[b] = cr[b];
#pragma unroll
    for (int b = 0; b < Q; ++b) wy[b] = cr[3 * Q + b] * kInv4Pi;
    const float ny0 = cr[4 * Q], ny1 = cr[4 * Q + 1], ny2 = cr[4 * Q + 2];
    const int4 tj = tris[active ? j : nt - 1];
    const float krev = k * kInv2Pi;
    const int64_t p_begin = (int64_t)blockIdx.y * kPotRowTile;
    const int64_t p_end = min(p_begin + (int64_t)kPotRowTile, np);
    for (int64_t p = p_begin; p < p_end; ++p) {
        const float x0 = pts[3 * p], x1 = pts[3 * p + 1], x2 = pts[3 * p + 2];
        float s[3][2] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
        float tr = 0.f, ti = 0.f;
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            float vr, vi, g0 = 0.f, g1 = 0.f;
            kernel4pi<OP, CLAMP>(py[3 * b] - x0, py[3 * b + 1] - x1, py[3 * b + 2] - x2, 0.f, 0.f, 0.f, ny0, ny1,
                                 ny2, k, krev, vr, vi, g0, g1);
            vr *= wy[b];
            vi *= wy[b];
            if constexpr (BASIS == P1) {
                const float ph[3] = {1.0f - TriRule<Q>::x1[b] - TriRule<Q>::x2[b], TriRule<Q>::x1[b],
                                     TriRule<Q>::x2[b]};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    s[c][0] = fmaf(vr, ph[c], s[c][0]);
                    s[c][1] = fmaf(vi, ph[c], s[c][1]);
                }
            } else {
                tr += vr;
                ti += vi;
            }
        }
        if (!active) continue;
        if constexpr (BASIS == DP0) {
            out[p * ld + j] = make_float2(tr, ti);
        } else if constexpr (GATHER) {
            float2* dst = out + (p * nt + j) * 3;
            dst[0] = make_float2(s[0][0], s[0][1]);
            dst[1] = make_float2(s[1][0], s[1][1]);
            dst[2] = make_float2(s[2][0], s[2][1]);
        } else {
            const int vs[3] = {tj.x, tj.y, tj.z};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float* dst = reinterpret_cast<float*>(&out[p * ld + vs[c]]);
                unsafeAtomicAdd(dst, s[c][0]);
                unsafeAtomicAdd(dst + 1, s[c][1]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned grid1d(int64_t n, int64_t per = 256, int64_t cap = 1 << 20)
{
    int64_t g = (n + per - 1) / per;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// nz matrices at out + z * zstride
static hipError_t launch_fill_z(float2* out, int64_t rows, int64_t cols, int64_t ld, int64_t zstride, int nz,
                                float2 val, hipStream_t s)
{
    if (rows <= 0 || cols <= 0 || nz <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_c64_kernel, dim3(grid1d(rows * cols, 256, 65536), 1, (unsigned)nz), dim3(256), 0, s,
                       out, rows, cols, ld, val, zstride);
    return hipGetLastError();
}

hipError_t launch_fill(float2* out, int64_t rows, int64_t cols, int64_t ld, float2 val, hipStream_t s)
{
    return launch_fill_z(out, rows, cols, ld, 0, 1, val, s);
}

template <int OP, int Q, bool CLAMP>
static hipError_t dp0_regular_t(const MeshDev& m, int64_t row0, int64_t row1, const KBatch& kb, int nz, float2* out,
                                int64_t ld, hipStream_t s)
{
    const float* rec = (Q == 6) ? m.rec6 : (Q == 3) ? m.rec3 : m.rec1;
    dim3 grid((unsigned)((m.nt + 255) / 256), (unsigned)((row1 - row0 + kDp0RowTile - 1) / kDp0RowTile),
              (unsigned)nz);
    hipLaunchKernelGGL((dp0_regular_kernel<OP, Q, CLAMP>), grid, dim3(256), 0, s, rec, m.nt, row0, row1, kb, out, ld);
    return hipGetLastError();
}

// out[c, d] = S[vnew[c], vnew[d]] for each wavenumber z = blockIdx.y: the
// P1 regular part accumulated in the tiles' numbering (S: [n, n] per z,
// contiguous) moved to the caller's.  One workgroup per output row: the
// source row is gathered (it stays in L2), the output row written whole.
__global__ __launch_bounds__(256) void p1_permute_kernel(const float2* __restrict__ S, int64_t n,
                                                         const int32_t* __restrict__ vnew, float2* __restrict__ out,
                                                         int64_t ld, int64_t zstride)
{
    const int64_t c = blockIdx.x;
    const float2* src = S + (int64_t)blockIdx.y * n * n + (int64_t)vnew[c] * n;
    float2* dst = out + (int64_t)blockIdx.y * zstride + c * ld;
    for (int64_t d = threadIdx.x; d < n; d += 256) dst[d] = src[vnew[d]];
}

// The same move from a planar scratch: per z, the real parts [n, n], then
// the imaginary parts [n, n].
__global__ __launch_bounds__(256) void p1_permute_planar_kernel(const float* __restrict__ S, int64_t n,
                                                                const int32_t* __restrict__ vnew,
                                                                float2* __restrict__ out, int64_t ld, int64_t zstride)
{
    const int64_t c = blockIdx.x;
    const float* re = S + (int64_t)blockIdx.y * 2 * n * n + (int64_t)vnew[c] * n;
    const float* im = re + n * n;
    float2* dst = out + (int64_t)blockIdx.y * zstride + c * ld;
    for (int64_t d = threadIdx.x; d < n; d += 256) {
        const int64_t q = vnew[d];
        dst[d] = make_float2(re[q], im[q]);
    }
}

// Half sweeps where they pay (single layer and hypersingular: symmetric;
// double layer and its adjoint with the 6-point rule on meshes of at least
// kP1PairMinVerts vertices: paired -- on smaller meshes the halved grid
// under-fills the GPU; with the 3-point rule the per-pair exchange and the
// transposed flush, whose destination rows take 20 values each, cost more
// than the halved evaluations save); one_sided evaluates every ordered pair.
constexpr int64_t kP1PairMinVerts = 4096;
template <int OP, int Q, bool CLAMP>
static hipError_t p1_regular_t(const MeshDev& m, const KBatch& kb, int nz, float2* out, int64_t ld, bool one_sided,
                               const int32_t* vmap, hipStream_t s, const float2* rg_g = nullptr,
                               float2* rg_out = nullptr, bool* planar = nullptr)
{
    // planar (an [n, n] scratch per z, ld == n): the every-ordered-pair and
    // paired sweeps flush real and imaginary planes; *planar says whether the
    // sweep taken did
    if (planar) *planar = false;
    const float* rec = (Q == 6) ? m.rec6 : (Q == 3) ? m.rec3 : m.rec1;
    const int64_t n = m.nv, nb = (n + kSymT - 1) / kSymT;
    constexpr bool sym = OP == SLP || OP == HYP;
    constexpr bool pair = (OP == DLP || OP == ADLP) && Q == 6;
    const int64_t plane = planar ? n * n : 0;
    if (!one_sided && m.tiles.nsym <= 0x7fffffffLL && nb * nb <= 0x7fffffffLL) {
        if constexpr (sym) {
            hipLaunchKernelGGL((p1_regular_kernel<OP, Q, CLAMP, kSweepSym>), dim3((unsigned)m.tiles.nsym, 1, nz),
                               dim3(256), 0, s, rec, m.tris, m.curls, m.tiles, kb, out, ld, vmap, nullptr, nullptr,
                               (int64_t)0);
            hipLaunchKernelGGL(p1_symmetrize_kernel, dim3((unsigned)(nb * (nb + 1) / 2), 1, nz), dim3(256), 0, s, out,
                               n, ld, kb.zstride);
            return hipGetLastError();
        } else if constexpr (pair) {
            if (n >= kP1PairMinVerts) {
                if (planar) *planar = true;
                hipLaunchKernelGGL((p1_regular_kernel<OP, Q, CLAMP, kSweepPair>), dim3((unsigned)m.tiles.nsym, 1, nz),
                                   dim3(256), 0, s, rec, m.tris, m.curls, m.tiles, kb, out, ld, vmap, nullptr, nullptr,
                                   plane);
                return hipGetLastError();
            }
        }
    }
    dim3 grid((unsigned)m.tiles.ncol_tiles, (unsigned)m.tiles.nrow_tiles, (unsigned)nz);
    if (planar) *planar = true;
    if constexpr (OP == BM_LHS) {
        if (rg_out) {
            hipLaunchKernelGGL((p1_regular_kernel<OP, Q, CLAMP, kSweepAllRg>), grid, dim3(256), 0, s, rec, m.tris,
                               m.curls, m.tiles, kb, out, ld, vmap, rg_g, rg_out, plane);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((p1_regular_kernel<OP, Q, CLAMP, kSweepAll>), grid, dim3(256), 0, s, rec, m.tris, m.curls,
                       m.tiles, kb, out, ld, vmap, nullptr, nullptr, plane);
    return hipGetLastError();
}

template <int OP, int NL, int BASIS, bool CLAMP>
static hipError_t singular_t(const MeshDev& m, int64_t p0, int64_t p1, const KBatch& kb, int nz, float2* out,
                             int64_t ld, int64_t row0, hipStream_t s)
{
    if (p1 <= p0) return hipSuccess;
    // pairs per workgroup: 4 wavefronts, 4 pairs each for the 2-point P1 rule
    const int64_t per = (NL == 2 && BASIS == P1) ? 16 : 4;
    unsigned g = (unsigned)((p1 - p0 + per - 1) / per);
    // P1 takes every pair (p0 = 0, p1 = npairs): the list grouped by kind
    const int4* pl = BASIS == P1 ? m.pairs_by_kind : m.pairs;
    hipLaunchKernelGGL((singular_kernel<OP, NL, BASIS, CLAMP>), dim3(g, 1, nz), dim3(256), 0, s, m.geo, m.tris,
                       m.curls, pl, p0, p1, kb, out, ld, row0, 1.0f, nullptr, nullptr);
    return hipGetLastError();
}

// ---- fused DP0 CBIE system -------------------------------------------------
// work: prow [nb, nt] | pcol [nb * 64 / rt, nt] | vpsi [npairs] (complex64)
int64_t cbie_work_bytes(const MeshDev& m, int rt)
{
    const int64_t nb = (m.nt + kSymBlock - 1) / kSymBlock;
    return (int64_t)sizeof(float2) * (nb * m.nt + nb * (kSymBlock / rt) * m.nt + m.npairs);
}

CbieWork cbie_work_split(const MeshDev& m, int rt, void* work)
{
    const int64_t nb = (m.nt + kSymBlock - 1) / kSymBlock;
    CbieWork w;
    w.prow = (float2*)work;
    w.pcol = w.prow + nb * m.nt;
    w.vpsi = w.pcol + nb * (kSymBlock / rt) * m.nt;
    w.nsub = kSymBlock / rt;
    return w;
}

template <int Q, int NL, bool CLAMP>
static hipError_t cbie_t(const MeshDev& m, float k, const float2* psi, float2* lhs, int64_t ld, float2* rhs,
                         const CbieWork& w, int parts, hipStream_t s)
{
    const float* rec = (Q == 6) ? m.rec6 : (Q == 3) ? m.rec3 : m.rec1;
    if (parts & 1) {
        if (w.nsub != 1) return hipErrorInvalidValue;
        const int64_t ntasks = cbie_sym_tasks(m.nt, kSymBlock);
        hipLaunchKernelGGL((dp0_cbie_kernel<Q, CLAMP, kSymBlock>), dim3((unsigned)((ntasks + 3) / 4)), dim3(256), 0,
                           s, rec, m.tris, m.nt, k, psi, lhs, ld, w.prow, w.pcol, ntasks);
    }
    if (!(parts & 2)) return hipGetLastError();
    const unsigned g = (unsigned)((m.npairs + 3) / 4);
    hipLaunchKernelGGL((singular_kernel<DLP, NL, DP0, CLAMP, 0>), dim3(g), dim3(256), 0, s, m.geo, m.tris, m.curls,
                       m.pairs, (int64_t)0, m.npairs, kbatch1(k), lhs, ld, (int64_t)0, -1.0f, nullptr, nullptr);
    hipLaunchKernelGGL((singular_kernel<SLP, NL, DP0, CLAMP, 1>), dim3(g), dim3(256), 0, s, m.geo, m.tris, m.curls,
                       m.pairs, (int64_t)0, m.npairs, kbatch1(k), lhs, ld, (int64_t)0, 1.0f, psi, w.vpsi);
    hipLaunchKernelGGL(cbie_finalize_kernel, dim3(grid1d(m.nt)), dim3(256), 0, s, w.prow, w.pcol, w.nsub, w.vpsi,
                       m.pair_ptr, m.geo, m.nt, lhs, ld, rhs);
    return hipGetLastError();
}

hipError_t launch_cbie_dp0(const MeshDev& m, float k, int quad, int clamp, int parts, const float2* psi, float2* lhs,
                           int64_t ld, float2* rhs, const CbieWork& w, hipStream_t s)
{
    if (quad == 0) return clamp ? cbie_t<6, 4, true>(m, k, psi, lhs, ld, rhs, w, parts, s)
                                : cbie_t<6, 4, false>(m, k, psi, lhs, ld, rhs, w, parts, s);
    if (quad == 1) return clamp ? cbie_t<3, 2, true>(m, k, psi, lhs, ld, rhs, w, parts, s)
                                : cbie_t<3, 2, false>(m, k, psi, lhs, ld, rhs, w, parts, s);
    return clamp ? cbie_t<1, 2, true>(m, k, psi, lhs, ld, rhs, w, parts, s)
                 : cbie_t<1, 2, false>(m, k, psi, lhs, ld, rhs, w, parts, s);
}

// parts: bit 0 = regular pairs (and, for P1, the zero fill), bit 1 = neighbour
// pairs, bit 2 = P1 regular pairs one evaluation per ordered pair; nz
// wavenumbers of kb at out + z * kb.zstride
template <int OP, bool CLAMP>
static hipError_t assemble_op(const MeshDev& m, int basis, int quad, int parts, int64_t row0, int64_t row1,
                              const KBatch& kb, int nz, float2* out, int64_t ld, hipStream_t s)
{
    hipError_t e = hipSuccess;
    if constexpr (is_dk(OP)) {
        if (basis == DP0) return hipErrorInvalidValue;  // derivatives: P1 only
    } else if (basis == DP0) {
        if (parts & 1) {
            if (quad == 0) e = dp0_regular_t<OP, 6, CLAMP>(m, row0, row1, kb, nz, out, ld, s);
            else if (quad == 1) e = dp0_regular_t<OP, 3, CLAMP>(m, row0, row1, kb, nz, out, ld, s);
            else e = dp0_regular_t<OP, 1, CLAMP>(m, row0, row1, kb, nz, out, ld, s);
            if (e != hipSuccess) return e;
        }
        if (!(parts & 2)) return hipSuccess;
        const int64_t p0 = m.h_pair_ptr[row0], p1 = m.h_pair_ptr[row1];
        return (quad == 0) ? singular_t<OP, 4, DP0, CLAMP>(m, p0, p1, kb, nz, out, ld, row0, s)
                           : singular_t<OP, 2, DP0, CLAMP>(m, p0, p1, kb, nz, out, ld, row0, s);
    }
    if (parts & 1) {
        // The regular pairs accumulate in the tiles' P1 numbering, where a
        // tile's flush atomics land in few cache lines (scattered over the
        // caller's numbering the flush took half of the level-5
        // Burton-Miller kernels' 10 ms), in a scratch matrix per wavenumber
        // from the stream-ordered pool; one pass then moves them to out in
        // the caller's numbering.  Without room for the scratch the kernels
        // flush to out through the index map (slower: the slots of a tile
        // are in P1 order, so the mapped addresses are scattered).
        const bool one_sided = (parts & 4) != 0;
        const int64_t n = m.nv;
        auto regular = [&](float2* dst, int64_t ldd, const KBatch& kbd, const int32_t* vmap,
                           bool* planar) -> hipError_t {
            hipError_t er = launch_fill_z(dst, n, n, ldd, kbd.zstride, nz, make_float2(0.f, 0.f), s);
            if (er != hipSuccess) return er;
            if (quad == 0) return p1_regular_t<OP, 6, CLAMP>(m, kbd, nz, dst, ldd, one_sided, vmap, s, nullptr,
                                                             nullptr, planar);
            if (quad == 1) return p1_regular_t<OP, 3, CLAMP>(m, kbd, nz, dst, ldd, one_sided, vmap, s, nullptr,
                                                             nullptr, planar);
            return p1_regular_t<OP, 1, CLAMP>(m, kbd, nz, dst, ldd, one_sided, vmap, s, nullptr, nullptr, planar);
        };
        float2* S = nullptr;
        if (!(parts & 8) &&
            hipMallocAsync((void**)&S, sizeof(float2) * (size_t)n * (size_t)n * (size_t)nz, s) == hipSuccess) {
            KBatch kbs = kb;
            kbs.zstride = n * n;
            bool planar = false;
            e = regular(S, n, kbs, nullptr, &planar);
            if (e == hipSuccess) {
                if (planar)
                    hipLaunchKernelGGL(p1_permute_planar_kernel, dim3((unsigned)n, (unsigned)nz), dim3(256), 0, s,
                                       (const float*)S, n, m.p1_vnew, out, ld, kb.zstride);
                else
                    hipLaunchKernelGGL(p1_permute_kernel, dim3((unsigned)n, (unsigned)nz), dim3(256), 0, s, S, n,
                                       m.p1_vnew, out, ld, kb.zstride);
                e = hipGetLastError();
            }
            const hipError_t ef = hipFreeAsync(S, s);
            if (e == hipSuccess) e = ef;
        } else {
            (void)hipGetLastError();  // no room for the scratch, or none asked for (parts & 8)
            e = regular(out, ld, kb, m.p1_vold, nullptr);
        }
        if (e != hipSuccess) return e;
    }
    if (!(parts & 2)) return hipSuccess;
    return (quad == 0) ? singular_t<OP, 4, P1, CLAMP>(m, 0, m.npairs, kb, nz, out, ld, 0, s)
                       : singular_t<OP, 2, P1, CLAMP>(m, 0, m.npairs, kb, nz, out, ld, 0, s);
}

static hipError_t assemble_dispatch(const MeshDev& m, int op, int basis, int quad, int clamp, int parts, int64_t row0,
                                    int64_t row1, const KBatch& kb, int nz, float2* out, int64_t ld, hipStream_t s)
{
#define BEM_OP_CASE(OPV)                                                                                     \
    case OPV:                                                                                                \
        return clamp ? assemble_op<OPV, true>(m, basis, quad, parts, row0, row1, kb, nz, out, ld, s)       \
                     : assemble_op<OPV, false>(m, basis, quad, parts, row0, row1, kb, nz, out, ld, s);
    switch (op) {
        BEM_OP_CASE(SLP)
        BEM_OP_CASE(DLP)
        BEM_OP_CASE(ADLP)
        BEM_OP_CASE(HYP)
        BEM_OP_CASE(BM_LHS)
        BEM_OP_CASE(BM_RHS)
        BEM_OP_CASE(SLP | kDk)
        BEM_OP_CASE(DLP | kDk)
        BEM_OP_CASE(ADLP | kDk)
        BEM_OP_CASE(BM_LHS | kDk)
        BEM_OP_CASE(BM_RHS | kDk)
    default:
        return hipErrorInvalidValue;
    }
#undef BEM_OP_CASE
}

hipError_t launch_assemble(const MeshDev& m, int op, int basis, int quad, int clamp, int parts, int64_t row0,
                           int64_t row1, float k, float2* out, int64_t ld, hipStream_t s)
{
    return assemble_dispatch(m, op, basis, quad, clamp, parts, row0, row1, kbatch1(k), 1, out, ld, s);
}

hipError_t launch_assemble_batch_k(const MeshDev& m, int op, int basis, int quad, int clamp, const float* ks, int nk,
                                   float2* out, int64_t ld, int64_t stride, hipStream_t s)
{
    const int64_t n = basis == DP0 ? m.nt : m.nv;
    for (int z0 = 0; z0 < nk; z0 += kMaxBatchK) {
        const int nz = nk - z0 < kMaxBatchK ? nk - z0 : kMaxBatchK;
        KBatch kb{};
        for (int z = 0; z < nz; ++z) kb.k[z] = ks[z0 + z];
        kb.zstride = stride;
        const hipError_t e =
            assemble_dispatch(m, op, basis, quad, clamp, 3, 0, n, kb, nz, out + (int64_t)z0 * stride, ld, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_identity(const MeshDev& m, int basis, float2* out, int64_t ld, hipStream_t s)
{
    const int64_t n = basis == DP0 ? m.nt : m.nv;
    hipError_t e = launch_fill(out, n, n, ld, make_float2(0.f, 0.f), s);
    if (e != hipSuccess) return e;
    if (basis == DP0)
        hipLaunchKernelGGL(identity_dp0_kernel, dim3(grid1d(m.nt)), dim3(256), 0, s, m.geo, m.nt, out, ld);
    else
        hipLaunchKernelGGL(identity_p1_kernel, dim3(grid1d(m.nv)), dim3(256), 0, s, m.geo, m.tris, m.vt_ptr, m.vt_idx,
                           m.nv, out, ld);
    return hipGetLastError();
}

// out_z += alpha_z * M_z (mass matrix of geometry z) in place, z = blockIdx.y
// (one geometry: gstride 0): DP0 diagonal, P1 one thread per row

__global__ __launch_bounds__(256) void add_identity_dp0_kernel(const float* __restrict__ geo, int64_t gstride,
                                                               int64_t nt, const AlphaBatch ab,
                                                               float2* __restrict__ out, int64_t ld, int64_t zstride)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nt) return;
    geo += (int64_t)blockIdx.y * gstride;
    out += (int64_t)blockIdx.y * zstride;
    const float2 alpha = ab.a[blockIdx.y];
    const float a = 0.5f * geo[t * 16 + 12];
    float2* d = &out[t * ld + t];
    d->x += alpha.x * a;
    d->y += alpha.y * a;
}

__global__ __launch_bounds__(256) void add_identity_p1_kernel(const float* __restrict__ geo, int64_t gstride,
                                                              const int4* __restrict__ tris,
                                                              const int32_t* __restrict__ vt_ptr,
                                                              const int32_t* __restrict__ vt_idx, int64_t nv,
                                                              const AlphaBatch ab, float2* __restrict__ out,
                                                              int64_t ld, int64_t zstride)
{
    const int64_t va = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (va >= nv) return;
    geo += (int64_t)blockIdx.y * gstride;
    out += (int64_t)blockIdx.y * zstride;
    const float2 alpha = ab.a[blockIdx.y];
    for (int e = vt_ptr[va]; e < vt_ptr[va + 1]; ++e) {
        const int t = vt_idx[e];
        const float area = 0.5f * geo[(int64_t)t * 16 + 12];
        const int4 tv = tris[t];
        const int vs[3] = {tv.x, tv.y, tv.z};
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float m = area * (vs[b] == va ? (2.0f / 12.0f) : (1.0f / 12.0f));
            float2* d = &out[va * ld + vs[b]];
            d->x += alpha.x * m;
            d->y += alpha.y * m;
        }
    }
}

static hipError_t add_identity_z(const MeshDev& m, int64_t gstride, int basis, const AlphaBatch& ab, int nz,
                                 float2* out, int64_t ld, int64_t zstride, hipStream_t s)
{
    if (basis == DP0)
        hipLaunchKernelGGL(add_identity_dp0_kernel, dim3(grid1d(m.nt), (unsigned)nz), dim3(256), 0, s, m.geo, gstride,
                           m.nt, ab, out, ld, zstride);
    else
        hipLaunchKernelGGL(add_identity_p1_kernel, dim3(grid1d(m.nv), (unsigned)nz), dim3(256), 0, s, m.geo, gstride,
                           m.tris, m.vt_ptr, m.vt_idx, m.nv, ab, out, ld, zstride);
    return hipGetLastError();
}

hipError_t launch_add_identity(const MeshDev& m, int basis, float2 alpha, float2* out, int64_t ld, hipStream_t s)
{
    AlphaBatch ab{};
    ab.a[0] = alpha;
    return add_identity_z(m, 0, basis, ab, 1, out, ld, 0, s);
}

// The Burton-Miller system of neumann2dirichlet in the P1 basis:
// lhs = bm_lhs + M/2, rhs = (bm_rhs - (i/2k) M) g.  The RHS operator is never
// stored: its regular pairs are contracted with g inside the LHS tile kernel
// (the same Gauss-point pairs, sharing e^{ikr}/r), its neighbour pairs by the
// contraction mode of the neighbour-pair kernel, its mass term by one pass
// over the triangles.
template <bool CLAMP>
static hipError_t bm_system_p1_t(const MeshDev& m, int64_t gstride, const KBatch& kbin, int nz, int quad,
                                 const float2* g, float2* lhs, int64_t ld, int64_t lstride, float2* rhs,
                                 hipStream_t s)
{
    const int64_t n = m.nv;
    KBatch kb = kbin;
    kb.zstride = lstride;
    kb.gstride = gstride;
    kb.vstride = n;
    hipError_t e = launch_fill(rhs, nz, n, n, make_float2(0.f, 0.f), s);
    if (e != hipSuccess) return e;
    // the right-hand side's regular pairs: one partial row per column tile and system
    const int64_t nslots = m.h_row_slots;
    float2* part = nullptr;
    e = hipMallocAsync((void**)&part, sizeof(float2) * (size_t)m.tiles.ncol_tiles * (size_t)nslots * (size_t)nz, s);
    if (e != hipSuccess) return e;
    auto regular = [&](float2* dst, int64_t ldd, const KBatch& kbd, const int32_t* vmap,
                       bool* planar) -> hipError_t {
        hipError_t er = launch_fill_z(dst, n, n, ldd, kbd.zstride, nz, make_float2(0.f, 0.f), s);
        if (er != hipSuccess) return er;
        if (quad == 0) return p1_regular_t<BM_LHS, 6, CLAMP>(m, kbd, nz, dst, ldd, true, vmap, s, g, part, planar);
        if (quad == 1) return p1_regular_t<BM_LHS, 3, CLAMP>(m, kbd, nz, dst, ldd, true, vmap, s, g, part, planar);
        return p1_regular_t<BM_LHS, 1, CLAMP>(m, kbd, nz, dst, ldd, true, vmap, s, g, part, planar);
    };
    float2* S = nullptr;
    if (hipMallocAsync((void**)&S, sizeof(float2) * (size_t)n * (size_t)n * (size_t)nz, s) == hipSuccess) {
        KBatch kbs = kb;
        kbs.zstride = n * n;
        bool planar = false;
        e = regular(S, n, kbs, nullptr, &planar);
        if (e == hipSuccess) {
            if (planar)
                hipLaunchKernelGGL(p1_permute_planar_kernel, dim3((unsigned)n, (unsigned)nz), dim3(256), 0, s,
                                   (const float*)S, n, m.p1_vnew, lhs, ld, lstride);
            else
                hipLaunchKernelGGL(p1_permute_kernel, dim3((unsigned)n, (unsigned)nz), dim3(256), 0, s, S, n,
                                   m.p1_vnew, lhs, ld, lstride);
            e = hipGetLastError();
        }
        const hipError_t ef = hipFreeAsync(S, s);
        if (e == hipSuccess) e = ef;
    } else {
        (void)hipGetLastError();  // no room for the scratch: flush through the vertex map
        e = regular(lhs, ld, kb, m.p1_vold, nullptr);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(p1_rg_reduce_kernel, dim3(grid1d(nslots), (unsigned)nz), dim3(256), 0, s,
                           (const float2*)part, m.tiles.ncol_tiles, nslots, m.tiles.row_verts, m.p1_vold, rhs, n);
        e = hipGetLastError();
    }
    {
        const hipError_t ef = hipFreeAsync(part, s);
        if (e == hipSuccess) e = ef;
    }
    if (e != hipSuccess) return e;
    // neighbour pairs: the LHS into lhs, the RHS contracted with g into rhs
    if (quad == 0) e = singular_t<BM_LHS, 4, P1, CLAMP>(m, 0, m.npairs, kb, nz, lhs, ld, 0, s);
    else e = singular_t<BM_LHS, 2, P1, CLAMP>(m, 0, m.npairs, kb, nz, lhs, ld, 0, s);
    if (e != hipSuccess) return e;
    if (m.npairs > 0) {
        if (quad == 0) {
            const unsigned gs = (unsigned)((m.npairs + 3) / 4);  // one pair per wavefront
            hipLaunchKernelGGL((singular_kernel<BM_RHS, 4, P1, CLAMP, 2>), dim3(gs, 1, (unsigned)nz), dim3(256), 0, s,
                               m.geo, m.tris, m.curls, m.pairs_by_kind, (int64_t)0, m.npairs, kb, (float2*)nullptr,
                               (int64_t)0, (int64_t)0, 1.0f, g, rhs);
        } else {
            const unsigned gs = (unsigned)((m.npairs + 15) / 16);  // four pairs per wavefront
            hipLaunchKernelGGL((singular_kernel<BM_RHS, 2, P1, CLAMP, 2>), dim3(gs, 1, (unsigned)nz), dim3(256), 0, s,
                               m.geo, m.tris, m.curls, m.pairs_by_kind, (int64_t)0, m.npairs, kb, (float2*)nullptr,
                               (int64_t)0, (int64_t)0, 1.0f, g, rhs);
        }
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    // mass terms: rhs_z -= (i/2k_z) M_z g_z, lhs_z += M_z/2
    AlphaBatch ar{}, al{};
    for (int z = 0; z < nz; ++z) {
        ar.a[z] = make_float2(0.f, -0.5f / kb.k[z]);
        al.a[z] = make_float2(0.5f, 0.f);
    }
    hipLaunchKernelGGL(p1_mass_apply_kernel, dim3(grid1d(m.nt), (unsigned)nz), dim3(256), 0, s, m.geo, gstride, m.tris,
                       m.nt, g, ar, rhs, n);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return add_identity_z(m, gstride, P1, al, nz, lhs, ld, lstride, s);
}

hipError_t launch_bm_system_p1(const MeshDev& m, float k, int quad, int clamp, const float2* g, float2* lhs,
                               int64_t ld, float2* rhs, hipStream_t s)
{
    const KBatch kb = kbatch1(k);
    return clamp ? bm_system_p1_t<true>(m, 0, kb, 1, quad, g, lhs, ld, 0, rhs, s)
                 : bm_system_p1_t<false>(m, 0, kb, 1, quad, g, lhs, ld, 0, rhs, s);
}

// The position-dependent pointers of m moved by off floats (the geometry of
// sample z0 of a geometry set whose region view is m).
static MeshDev geometry_shift(const MeshDev& m, int64_t off)
{
    MeshDev g = m;
    g.verts += off;
    g.rec6 += off;
    g.rec3 += off;
    g.rec1 += off;
    g.curls += off;
    g.geo += off;
    return g;
}

hipError_t launch_bm_systems_p1(const MeshDev& mg, int64_t gstride, const float* ks, int nk, int quad, int clamp,
                                const float2* g, float2* lhs, int64_t ld, int64_t stride, float2* rhs, hipStream_t s)
{
    const int64_t n = mg.nv;
    for (int z0 = 0; z0 < nk; z0 += kMaxBatchK) {
        const int nz = nk - z0 < kMaxBatchK ? nk - z0 : kMaxBatchK;
        KBatch kb{};
        for (int z = 0; z < nz; ++z) kb.k[z] = ks[z0 + z];
        const MeshDev m = geometry_shift(mg, (int64_t)z0 * gstride);
        const float2* gz = g + (int64_t)z0 * n;
        float2* lz = lhs + (int64_t)z0 * stride;
        float2* rz = rhs + (int64_t)z0 * n;
        const hipError_t e = clamp ? bm_system_p1_t<true>(m, gstride, kb, nz, quad, gz, lz, ld, stride, rz, s)
                                   : bm_system_p1_t<false>(m, gstride, kb, nz, quad, gz, lz, ld, stride, rz, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_assemble_samples(const MeshDev& mg, int64_t gstride, const float* ks, int ns, int op, int basis,
                                   int quad, int clamp, float2* out, int64_t ld, int64_t stride, hipStream_t s)
{
    const int64_t n = basis == DP0 ? mg.nt : mg.nv;
    for (int z0 = 0; z0 < ns; z0 += kMaxBatchK) {
        const int nz = ns - z0 < kMaxBatchK ? ns - z0 : kMaxBatchK;
        KBatch kb{};
        for (int z = 0; z < nz; ++z) kb.k[z] = ks[z0 + z];
        kb.zstride = stride;
        kb.gstride = gstride;
        const hipError_t e = assemble_dispatch(geometry_shift(mg, (int64_t)z0 * gstride), op, basis, quad, clamp, 3, 0,
                                               n, kb, nz, out + (int64_t)z0 * stride, ld, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_add_identity_samples(const MeshDev& mg, int64_t gstride, int basis, const float* alphas, int ns,
                                       float2* out, int64_t ld, int64_t stride, hipStream_t s)
{
    for (int z0 = 0; z0 < ns; z0 += kMaxBatchK) {
        const int nz = ns - z0 < kMaxBatchK ? ns - z0 : kMaxBatchK;
        AlphaBatch ab{};
        for (int z = 0; z < nz; ++z) ab.a[z] = make_float2(alphas[2 * (z0 + z)], alphas[2 * (z0 + z) + 1]);
        const hipError_t e = add_identity_z(geometry_shift(mg, (int64_t)z0 * gstride), gstride, basis, ab, nz,
                                            out + (int64_t)z0 * stride, ld, stride, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_tri2vert(const MeshDev& m, const float2* tv, float2* out, hipStream_t s)
{
    hipLaunchKernelGGL(tri2vert_kernel, dim3(grid1d(m.nv)), dim3(256), 0, s, tv, m.vt_ptr, m.vt_idx, m.nv, out);
    return hipGetLastError();
}

hipError_t launch_tri2vert_raw(const int32_t* tris, int64_t nt, int64_t nv, const float2* tv, float2* out,
                               int32_t* work, hipStream_t s)
{
    int32_t* cnt = work;
    int32_t* ptr = cnt + nv;
    int32_t* idx = ptr + nv + 1;
    int32_t* flag = idx + 3 * nt;
    hipLaunchKernelGGL(fill_i32_kernel, dim3(grid1d(nv, 256, 65536)), dim3(256), 0, s, cnt, nv);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(256), 0, s, flag, (int64_t)1);
    hipLaunchKernelGGL(t2v_count_kernel, dim3(grid1d(3 * nt, 256, 65536)), dim3(256), 0, s, tris, nt, nv, cnt, flag);
    hipLaunchKernelGGL(t2v_scan_kernel, dim3(1), dim3(1024), 0, s, cnt, nv, ptr);
    hipLaunchKernelGGL(t2v_fill_kernel, dim3(grid1d(3 * nt, 256, 65536)), dim3(256), 0, s, tris, nt, nv, cnt, idx);
    hipLaunchKernelGGL(t2v_gather_kernel, dim3(grid1d(nv, 256, 65536)), dim3(256), 0, s, tv, ptr, idx, nv, out);
    return hipGetLastError();
}

// out[p, v] = sum of the corner values of v's triangles for point p (the
// P1 potential from the per-(point, triangle) scratch of potential_kernel):
// one thread per (point, vertex), every entry written once (no fill, no
// atomics, a fixed summation order: v's triangles ascending).  vt_corner
// holds each incidence's scratch column (3 t + corner), so a value costs one
// index load and one scratch load.
__global__ __launch_bounds__(256) void p1_point_gather_kernel(const float2* __restrict__ t,
                                                              const int32_t* __restrict__ vt_ptr,
                                                              const int32_t* __restrict__ vt_corner, int64_t nv,
                                                              int64_t nt, int64_t np, int64_t tzstride,
                                                              float2* __restrict__ out, int64_t ld, int64_t zstride)
{
    t += (int64_t)blockIdx.z * tzstride;
    out += (int64_t)blockIdx.z * zstride;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < np * nv; e += (int64_t)gridDim.x * 256) {
        const int64_t p = e / nv, v = e - p * nv;
        const float2* row = t + p * nt * 3;
        float sr = 0.f, si = 0.f;
        for (int q = vt_ptr[v]; q < vt_ptr[v + 1]; ++q) {
            const float2 x = row[vt_corner[q]];
            sr += x.x;
            si += x.y;
        }
        out[p * ld + v] = make_float2(sr, si);
    }
}

// Scratch for the P1 point gather, per launch: at most this many bytes
// (points are taken in chunks that fit).
constexpr int64_t kPotScratchBytes = (int64_t)512 << 20;
constexpr int64_t kPotMaxPoints = (int64_t)65535 * kPotRowTile;  // points per launch (grid y <= 65535)

template <int OP, int BASIS, bool CLAMP>
static hipError_t potential_t(const MeshDev& m, const float* pts, int64_t np, const KBatch& kb, int nz, float2* out,
                              int64_t ld, hipStream_t s)
{
    if constexpr (BASIS == P1) {
        // two passes through a scratch array of point chunks; the atomic
        // kernel only if no scratch can be had
        const int64_t per_point = (int64_t)sizeof(float2) * 3 * m.nt * nz;
        int64_t pc = kPotScratchBytes / per_point;
        pc = pc < np ? pc : np;
        pc = pc < kPotRowTile ? kPotRowTile : pc;
        pc = pc < kPotMaxPoints ? pc : kPotMaxPoints;
        float2* t = nullptr;
        if (np > 0 && hipMallocAsync((void**)&t, (size_t)(per_point * pc), s) == hipSuccess) {
            hipError_t e = hipSuccess;
            for (int64_t c0 = 0; c0 < np && e == hipSuccess; c0 += pc) {
                const int64_t c1 = c0 + pc < np ? c0 + pc : np, cn = c1 - c0;
                KBatch kt = kb;
                kt.zstride = cn * m.nt * 3;
                dim3 grid((unsigned)((m.nt + 255) / 256), (unsigned)((cn + kPotRowTile - 1) / kPotRowTile),
                          (unsigned)nz);
                hipLaunchKernelGGL((potential_kernel<OP, P1, CLAMP, true>), grid, dim3(256), 0, s, m.rec6, m.tris,
                                   m.nt, pts + 3 * c0, cn, kt, t, (int64_t)0);
                hipLaunchKernelGGL(p1_point_gather_kernel, dim3(grid1d(cn * m.nv, 256, 65536), 1, (unsigned)nz),
                                   dim3(256), 0, s, (const float2*)t, m.vt_ptr, m.vt_corner, m.nv, m.nt, cn,
                                   kt.zstride, out + c0 * ld, ld, kb.zstride);
                e = hipGetLastError();
            }
            const hipError_t ef = hipFreeAsync(t, s);
            return e != hipSuccess ? e : ef;
        }
        (void)hipGetLastError();
        hipError_t e = launch_fill_z(out, np, m.nv, ld, kb.zstride, nz, make_float2(0.f, 0.f), s);
        if (e != hipSuccess) return e;
    }
    // grid y covers kPotRowTile points per workgroup: chunks of at most
    // kPotMaxPoints points (the points of sample z stay at pts + z pstride)
    for (int64_t c0 = 0; c0 < np; c0 += kPotMaxPoints) {
        const int64_t cn = np - c0 < kPotMaxPoints ? np - c0 : kPotMaxPoints;
        dim3 grid((unsigned)((m.nt + 255) / 256), (unsigned)((cn + kPotRowTile - 1) / kPotRowTile), (unsigned)nz);
        hipLaunchKernelGGL((potential_kernel<OP, BASIS, CLAMP>), grid, dim3(256), 0, s, m.rec6, m.tris, m.nt,
                           pts + 3 * c0, cn, kb, out + c0 * ld, ld);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

static hipError_t potential_dispatch(const MeshDev& m, const float* pts, int64_t np, const KBatch& kb, int nz, int op,
                                     int basis, int clamp, float2* out, int64_t ld, hipStream_t s)
{
    if (op == SLP) {
        if (basis == DP0) return clamp ? potential_t<SLP, DP0, true>(m, pts, np, kb, nz, out, ld, s)
                                       : potential_t<SLP, DP0, false>(m, pts, np, kb, nz, out, ld, s);
        return clamp ? potential_t<SLP, P1, true>(m, pts, np, kb, nz, out, ld, s)
                     : potential_t<SLP, P1, false>(m, pts, np, kb, nz, out, ld, s);
    }
    if (op == DLP) {
        if (basis == DP0) return clamp ? potential_t<DLP, DP0, true>(m, pts, np, kb, nz, out, ld, s)
                                       : potential_t<DLP, DP0, false>(m, pts, np, kb, nz, out, ld, s);
        return clamp ? potential_t<DLP, P1, true>(m, pts, np, kb, nz, out, ld, s)
                     : potential_t<DLP, P1, false>(m, pts, np, kb, nz, out, ld, s);
    }
    if (op == (SLP | kDk)) {
        if (basis == DP0) return clamp ? potential_t<SLP | kDk, DP0, true>(m, pts, np, kb, nz, out, ld, s)
                                       : potential_t<SLP | kDk, DP0, false>(m, pts, np, kb, nz, out, ld, s);
        return clamp ? potential_t<SLP | kDk, P1, true>(m, pts, np, kb, nz, out, ld, s)
                     : potential_t<SLP | kDk, P1, false>(m, pts, np, kb, nz, out, ld, s);
    }
    if (op == (DLP | kDk)) {
        if (basis == DP0) return clamp ? potential_t<DLP | kDk, DP0, true>(m, pts, np, kb, nz, out, ld, s)
                                       : potential_t<DLP | kDk, DP0, false>(m, pts, np, kb, nz, out, ld, s);
        return clamp ? potential_t<DLP | kDk, P1, true>(m, pts, np, kb, nz, out, ld, s)
                     : potential_t<DLP | kDk, P1, false>(m, pts, np, kb, nz, out, ld, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_potential(const MeshDev& m, const float* pts, int64_t np, float k, int op, int basis, int clamp,
                            float2* out, int64_t ld, hipStream_t s)
{
    return potential_dispatch(m, pts, np, kbatch1(k), 1, op, basis, clamp, out, ld, s);
}

hipError_t launch_potential_samples(const MeshDev& mg, int64_t gstride, const float* pts, int64_t np, const float* ks,
                                    int ns, int op, int basis, int clamp, float2* out, int64_t ld, int64_t stride,
                                    hipStream_t s)
{
    for (int z0 = 0; z0 < ns; z0 += kMaxBatchK) {
        const int nz = ns - z0 < kMaxBatchK ? ns - z0 : kMaxBatchK;
        KBatch kb{};
        for (int z = 0; z < nz; ++z) kb.k[z] = ks[z0 + z];
        kb.zstride = stride;
        kb.gstride = gstride;
        kb.pstride = 3 * np;
        const hipError_t e = potential_dispatch(geometry_shift(mg, (int64_t)z0 * gstride), pts + 3 * np * z0, np, kb,
                                                nz, op, basis, clamp, out + (int64_t)z0 * stride, ld, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_potential_batch_k(const MeshDev& m, const float* pts, int64_t np, const float* ks, int nk, int op,
                                    int basis, int clamp, float2* out, int64_t ld, int64_t stride, hipStream_t s)
{
    for (int z0 = 0; z0 < nk; z0 += kMaxBatchK) {
        const int nz = nk - z0 < kMaxBatchK ? nk - z0 : kMaxBatchK;
        KBatch kb{};
        for (int z = 0; z < nz; ++z) kb.k[z] = ks[z0 + z];
        kb.zstride = stride;
        const hipError_t e =
            potential_dispatch(m, pts, np, kb, nz, op, basis, clamp, out + (int64_t)z0 * stride, ld, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace bem
