// Batched solves over many (mesh, k) items of uneven size.
//
//  * bem_batch_cbie: the fused DP0 CBIE systems of all items in one
//    persistent launch (a device-side queue of tiles, largest items first),
//    then the neighbour pairs and the row reduction of each item.
//  * bem_batch_gmres: restarted GMRES on all items in lockstep.  Each item has
//    its own matrix and size; vectors of all items live in one concatenated
//    space (item s at offset off_s), and every kernel of an Arnoldi step covers
//    all items (grid y = item).  Per-item Givens rotations and convergence
//    flags stay on the device; a converged item skips its work.  One restart
//    cycle (residual, set-up, m Arnoldi steps, update) is captured once as a
//    hipGraph and replayed; the host reads the items' flags once per cycle.
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

struct bem_mesh {
    bem::MeshDev dev;
};

namespace bem {

struct BSys {
    const float2* A;
    int64_t lda;
    int64_t n;
    int64_t off;
};

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

// x_s += sum_{q < nk_s} y_s[q] V_q ; grid (element blocks, items)
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

namespace {

size_t al(size_t b) { return (b + 255) / 256 * 256; }

struct XLayout {
    size_t sys, V, w, r, b, x, H, h, h2, cs, sn, g, y, res, rn2, wn2, tgt, done, nk, conv, total;
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

int bem_batch_gmres(const bem_batch_system* sys, int ns, float tol, int restart, int max_cycles, int use_graph,
                    void* work, bem_solve_info* info, void* stream)
{
    if (!sys || ns <= 0 || restart <= 0 || max_cycles < 0 || !work || !(tol > 0.f)) return BEM_ERR_BAD_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int m = restart;
    std::vector<BSys> hs(ns);
    int64_t ntot = 0, nmax = 0;
    for (int q = 0; q < ns; ++q) {
        if (!sys[q].A || !sys[q].b || !sys[q].x || sys[q].n <= 0 || sys[q].lda < sys[q].n) return BEM_ERR_BAD_ARG;
        hs[q] = BSys{(const float2*)sys[q].A, sys[q].lda, sys[q].n, ntot};
        ntot += padded(sys[q].n);
        nmax = std::max(nmax, sys[q].n);
    }
    const XLayout L = xlayout(ns, ntot, m);
    char* W = (char*)work;
    BSys* dsys = (BSys*)(W + L.sys);
    float2* V = (float2*)(W + L.V);
    float2* w = (float2*)(W + L.w);
    float2* r = (float2*)(W + L.r);
    float2* B = (float2*)(W + L.b);
    float2* X = (float2*)(W + L.x);
    double2* H = (double2*)(W + L.H);
    double2* h = (double2*)(W + L.h);
    double2* h2 = (double2*)(W + L.h2);
    double2* cs = (double2*)(W + L.cs);
    double2* sn = (double2*)(W + L.sn);
    double2* g = (double2*)(W + L.g);
    double2* y = (double2*)(W + L.y);
    double* res = (double*)(W + L.res);
    double* rn2 = (double*)(W + L.rn2);
    double* wn2 = (double*)(W + L.wn2);
    double* tgt = (double*)(W + L.tgt);
    int* done = (int*)(W + L.done);
    int* nk = (int*)(W + L.nk);
    int* conv = (int*)(W + L.conv);
    XTRY(hipMemcpyAsync(dsys, hs.data(), sizeof(BSys) * ns, hipMemcpyHostToDevice, s));
    // gather b and the initial x into the concatenated space
    for (int q = 0; q < ns; ++q) {
        XTRY(hipMemcpyAsync(B + hs[q].off, sys[q].b, sizeof(float2) * hs[q].n, hipMemcpyDeviceToDevice, s));
        XTRY(hipMemcpyAsync(X + hs[q].off, sys[q].x, sizeof(float2) * hs[q].n, hipMemcpyDeviceToDevice, s));
    }
    std::vector<double> hb(ns), ht(ns), hr(ns);
    hipLaunchKernelGGL(xnorm2_kernel, dim3(ns), dim3(256), 0, s, dsys, B, rn2);
    XTRY(hipGetLastError());
    XTRY(hipMemcpyAsync(hb.data(), rn2, sizeof(double) * ns, hipMemcpyDeviceToHost, s));
    XTRY(hipStreamSynchronize(s));
    for (int q = 0; q < ns; ++q) ht[q] = (double)tol * std::sqrt(hb[q]);
    XTRY(hipMemcpyAsync(tgt, ht.data(), sizeof(double) * ns, hipMemcpyHostToDevice, s));

    const dim3 gx((unsigned)((nmax + 255) / 256), (unsigned)ns);
    const dim3 gmv((unsigned)((nmax + 15) / 16), (unsigned)ns);
    const dim3 gc((unsigned)((nmax + 63) / 64), (unsigned)ns);
    const float2 one = make_float2(1.f, 0.f), zero = make_float2(0.f, 0.f), mone = make_float2(-1.f, 0.f);
    // one restart cycle: residual, set-up, m Arnoldi steps, back-solve, update,
    // then the true residual and the convergence flags
    auto enqueue_cycle = [&](hipStream_t st) -> hipError_t {
        hipLaunchKernelGGL(xcopy_kernel, dim3((unsigned)((ntot + 255) / 256)), dim3(256), 0, st, B, r, ntot);
        hipLaunchKernelGGL(xgemv_kernel, gmv, dim3(256), 0, st, dsys, X, r, mone, one, 0, (const int*)nullptr);
        hipLaunchKernelGGL(xnorm2_kernel, dim3(ns), dim3(256), 0, st, dsys, r, rn2);
        hipError_t e = launch_binit(g, m, rn2, tgt, done, nk, ns, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(xnormalize_kernel, gx, dim3(256), 0, st, dsys, r, rn2, V, (const int*)nullptr);
        for (int j = 0; j < m; ++j) {
            hipLaunchKernelGGL(xgemv_kernel, gmv, dim3(256), 0, st, dsys, V + (int64_t)j * ntot, w, one, zero, 1,
                               (const int*)done);
            hipLaunchKernelGGL(xmultidot_kernel, dim3(j + 1, ns), dim3(256), 0, st, dsys, V, ntot, w, h, m + 1,
                               (const int*)done);
            hipLaunchKernelGGL(xcombine_kernel, gc, dim3(64 * kXWaves), 0, st, dsys, V, ntot, j + 1,
                               (const int*)nullptr, h, m + 1, -1.0, w,
                               (const int*)done);
            hipLaunchKernelGGL(xmultidot_kernel, dim3(j + 1, ns), dim3(256), 0, st, dsys, V, ntot, w, h2, m + 1,
                               (const int*)done);
            hipLaunchKernelGGL(xcombine_kernel, gc, dim3(64 * kXWaves), 0, st, dsys, V, ntot, j + 1,
                               (const int*)nullptr, h2, m + 1, -1.0, w,
                               (const int*)done);
            hipLaunchKernelGGL(xnorm2_kernel, dim3(ns), dim3(256), 0, st, dsys, w, wn2);
            hipLaunchKernelGGL(xnormalize_kernel, gx, dim3(256), 0, st, dsys, w, wn2, V + (int64_t)(j + 1) * ntot,
                               (const int*)done);
            e = launch_bgivens(H, m, j, h, h2, m + 1, wn2, cs, sn, g, res, tgt, done, nk, ns, st);
            if (e != hipSuccess) return e;
        }
        e = launch_bbacksolve(H, m, nk, g, y, ns, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(xcombine_kernel, gc, dim3(64 * kXWaves), 0, st, dsys, V, ntot, 0, (const int*)nk, y, m,
                           1.0, X, (const int*)nullptr);
        hipLaunchKernelGGL(xcopy_kernel, dim3((unsigned)((ntot + 255) / 256)), dim3(256), 0, st, B, r, ntot);
        hipLaunchKernelGGL(xgemv_kernel, gmv, dim3(256), 0, st, dsys, X, r, mone, one, 0, (const int*)nullptr);
        hipLaunchKernelGGL(xnorm2_kernel, dim3(ns), dim3(256), 0, st, dsys, r, rn2);
        hipLaunchKernelGGL(xconv_kernel, dim3((ns + 63) / 64), dim3(64), 0, st, rn2, tgt, conv, ns);
        return hipGetLastError();
    };

    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    hipStream_t cap = nullptr;
    if (use_graph) {
        XTRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        XTRY(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        hipError_t e = enqueue_cycle(cap);
        hipError_t e2 = hipStreamEndCapture(cap, &graph);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipStreamDestroy(cap);
            g_batch_err = "graph capture failed";
            return BEM_ERR_HIP;
        }
        XTRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    std::vector<int> hc(ns, 0), hk(ns, 0), its(ns, 0);
    int cycles = 0;
    bool all = false;
    for (; cycles < max_cycles; ++cycles) {
        if (use_graph) {
            XTRY(hipGraphLaunch(exec, s));
        } else {
            XTRY(enqueue_cycle(s));
        }
        XTRY(hipMemcpyAsync(hc.data(), conv, sizeof(int) * ns, hipMemcpyDeviceToHost, s));
        XTRY(hipMemcpyAsync(hr.data(), rn2, sizeof(double) * ns, hipMemcpyDeviceToHost, s));
        XTRY(hipMemcpyAsync(hk.data(), nk, sizeof(int) * ns, hipMemcpyDeviceToHost, s));
        XTRY(hipStreamSynchronize(s));
        all = true;
        for (int q = 0; q < ns; ++q) {
            all = all && hc[q];
            its[q] += hk[q];  // Arnoldi steps of this item in the cycle (0 once converged)
        }
        if (all) {
            ++cycles;
            break;
        }
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
    for (int q = 0; q < ns; ++q)
        XTRY(hipMemcpyAsync(sys[q].x, X + hs[q].off, sizeof(float2) * hs[q].n, hipMemcpyDeviceToDevice, s));
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

}  // extern "C"
