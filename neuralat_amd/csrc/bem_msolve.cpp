// Multi-right-hand-side solves on one dense matrix: lockstep GMRES over up to
// any number of systems (one multi-RHS product per Arnoldi step, per-system
// Gram-Schmidt and Givens on the device) and mixed-precision iterative
// refinement (inner GMRES on the 16-bit matrix with matrix cores, outer FP32
// residual on the complex64 matrix).  No reference counterpart: the
// reference solves one right-hand side at a time with BiCGSTAB
// (src/bem/solver.py:61-66).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "bem_linalg.hpp"
#include "bem_lowp.hpp"
#include "bem_msolve.hpp"

namespace bem {

namespace {
size_t al(size_t b) { return (b + 255) / 256 * 256; }
}  // namespace

hipError_t MultiOp::apply(const float2* X, int64_t ldx, int ns, float2* Y, int64_t ldy, float2 alpha, float2 beta,
                          hipStream_t s) const
{
    for (int s0 = 0; s0 < ns; s0 += 16) {
        const int cnt = std::min(16, ns - s0);
        const float2* Xs = X + (int64_t)s0 * ldx;
        float2* Ys = Y + (int64_t)s0 * ldy;
        hipError_t e;
        if (dtype == 0) {
            if (cnt == 1)
                e = launch_gemv(A, n, n, lda, 0, Xs, Ys, alpha, beta, nullptr, s);
            else
                e = launch_gemm_c64(A, n, n, lda, Xs, ldx, cnt, Ys, ldy, alpha, beta, part, s);
        } else {
            e = launch_vec_to_lowp(dtype, Xs, n, ldx, cnt, Xr, Xi, ldlp, s);
            if (e == hipSuccess)
                e = launch_gemm_lowp(dtype, Ar, Ai, n, ldlp, inv_scale, Xr, Xi, ldlp, cnt, Ys, ldy, alpha, beta, part, s);
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Xr, Xi (hi + lo planes of 16 right-hand sides, 32 x ldlp each), then the
// split-K partial sums of the multi-RHS products (rows <= ldlp)
size_t lowp_scratch_bytes(int64_t ldlp)
{
    return 2 * al(sizeof(uint16_t) * 32 * (size_t)ldlp) + al(sizeof(float2) * (size_t)splitk_part_elems(ldlp));
}
float2* lowp_scratch_part(void* scratch, int64_t ldlp)
{
    return (float2*)((char*)scratch + 2 * al(sizeof(uint16_t) * 32 * (size_t)ldlp));
}

struct MLayout {
    size_t V, w, r, H, h, h2, cs, sn, g, y, res, rn2, wn2, tgt, done, nk, total;
};

static MLayout mlayout(int64_t n, int ns, int m)
{
    MLayout L{};
    size_t o = 0;
    auto take = [&](size_t b) {
        size_t at = o;
        o += al(b);
        return at;
    };
    const size_t ldn = (size_t)((n + 63) / 64 * 64);
    L.V = take(sizeof(float2) * ldn * (size_t)ns * (size_t)(m + 1));
    L.w = take(sizeof(float2) * ldn * ns);
    L.r = take(sizeof(float2) * ldn * ns);
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
    L.total = o;
    return L;
}

size_t gmres_multi_bytes(int64_t n, int ns, int m) { return mlayout(n, ns, m).total; }

#define TRY(expr)                                   \
    do {                                            \
        hipError_t _e = (expr);                     \
        if (_e != hipSuccess) return _e;            \
    } while (0)

hipError_t gmres_multi(const MultiOp& op, int64_t n, int ns, const float2* B, int64_t ldb, float2* X, int64_t ldx,
                       double tol, int m, int maxiter, void* work, MultiInfo* info, hipStream_t s, FlexPrecond* prec,
                       const double* rel)
{
    const MLayout L = mlayout(n, ns, m);
    char* W = (char*)work;
    const int64_t ldn = (n + 63) / 64 * 64;
    float2* V = (float2*)(W + L.V);
    float2* w = (float2*)(W + L.w);
    float2* r = (float2*)(W + L.r);
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
    const float2 one = make_float2(1.f, 0.f), zero = make_float2(0.f, 0.f), mone = make_float2(-1.f, 0.f);

    std::vector<double> hb(ns), ht(ns), hr(ns), hres(ns), irel(ns);
    std::vector<int> hd(ns);
    TRY(launch_bnorm2(B, ldb, n, ns, rn2, s));
    TRY(hipMemcpyAsync(hb.data(), rn2, sizeof(double) * ns, hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    for (int q = 0; q < ns; ++q) ht[q] = (rel ? rel[q] : tol) * std::sqrt(hb[q]);
    TRY(hipMemcpyAsync(tgt, ht.data(), sizeof(double) * ns, hipMemcpyHostToDevice, s));
    int its = 0;
    bool all_conv = false;
    while (true) {
        TRY(launch_bcopy(B, ldb, r, ldn, n, ns, s));
        TRY(op.apply(X, ldx, ns, r, ldn, mone, one, s));
        TRY(launch_bnorm2(r, ldn, n, ns, rn2, s));
        TRY(hipMemcpyAsync(hr.data(), rn2, sizeof(double) * ns, hipMemcpyDeviceToHost, s));
        TRY(hipStreamSynchronize(s));
        all_conv = true;
        for (int q = 0; q < ns; ++q) all_conv = all_conv && (std::sqrt(hr[q]) <= ht[q]);
        if (all_conv || its >= maxiter) break;
        TRY(launch_binit(g, m, rn2, tgt, done, nk, ns, s));
        TRY(launch_bnormalize(r, ldn, n, ns, rn2, V, ldn, nullptr, s));
        for (int j = 0; j < m && its < maxiter; ++j) {
            const float2* Vj = V + (int64_t)j * ns * ldn;
            const float2* dir = Vj;
            if (prec) {
                // flexible right preconditioning: z_j ~ M^{-1} v_j by an inner
                // solve whose target relaxes as the outer residual falls
                // (inexact Krylov: step j only needs the inner error below
                // target / |r_j|)
                float2* Zj = prec->Z + (int64_t)j * ns * ldn;
                TRY(hipMemsetAsync(Zj, 0, sizeof(float2) * (size_t)ldn * ns, s));
                for (int q = 0; q < ns; ++q) {
                    const double cur = j == 0 ? std::sqrt(hr[q]) : hres[q];
                    irel[q] = cur > 0.0 ? std::min(0.5, std::max(prec->tol, ht[q] / cur)) : 0.5;
                }
                MultiInfo mi;
                TRY(gmres_multi(*prec->op, n, ns, Vj, ldn, Zj, ldn, prec->tol, prec->m, prec->maxiter, prec->work, &mi,
                                s, nullptr, irel.data()));
                prec->inner_iterations += mi.iterations;
                dir = Zj;
            }
            TRY(op.apply(dir, ldn, ns, w, ldn, one, zero, s));
            TRY(launch_bmultidot(V, ldn, ns, n, j + 1, w, h, m + 1, done, s));
            TRY(launch_bgs_update(V, ldn, ns, n, j + 1, h, m + 1, w, done, s));
            TRY(launch_bmultidot(V, ldn, ns, n, j + 1, w, h2, m + 1, done, s));
            TRY(launch_bgs_update(V, ldn, ns, n, j + 1, h2, m + 1, w, done, s));
            TRY(launch_bnorm2(w, ldn, n, ns, wn2, s));
            TRY(launch_bnormalize(w, ldn, n, ns, wn2, V + (int64_t)(j + 1) * ns * ldn, ldn, done, s));
            TRY(launch_bgivens(H, m, j, h, h2, m + 1, wn2, cs, sn, g, res, tgt, done, nk, ns, s));
            ++its;
            TRY(hipMemcpyAsync(hd.data(), done, sizeof(int) * ns, hipMemcpyDeviceToHost, s));
            if (prec) TRY(hipMemcpyAsync(hres.data(), res, sizeof(double) * ns, hipMemcpyDeviceToHost, s));
            TRY(hipStreamSynchronize(s));
            bool all = true;
            for (int q = 0; q < ns; ++q) all = all && hd[q];
            if (all) break;
        }
        TRY(launch_bbacksolve(H, m, nk, g, y, ns, s));
        TRY(launch_bupdate_x(prec ? prec->Z : V, ldn, ns, n, nk, y, m, X, ldx, s));
    }
    if (info) {
        info->iterations = its;
        info->converged = all_conv ? 1 : 0;
        info->rel_residual.resize(ns);
        for (int q = 0; q < ns; ++q) info->rel_residual[q] = hb[q] > 0 ? std::sqrt(hr[q] / hb[q]) : 0.0;
    }
    return hipSuccess;
}

// Mixed-precision refinement with a Krylov outer loop: flexible GMRES in FP32
// on the complex64 matrix (the FP32 residual and update of iterative
// refinement), each outer direction preconditioned by an inner multi-RHS GMRES
// solve on the 16-bit matrix to inner_tol.  Plain refinement (x += A16^{-1} r)
// converges only while ||I - A16^{-1} A|| < 1, i.e. about kappa(A) * 2^-8 < 1
// for bf16, which the CBIE operator near interior resonances does not meet;
// the flexible outer loop converges whatever the quality of the inner solve,
// and needs few outer steps when it is good.
constexpr int kIrOuter = 20;  // outer Krylov basis (restarted beyond it)

size_t solve_ir_bytes(int64_t n, int ns, int m)
{
    const size_t ldn = (size_t)((n + 63) / 64 * 64);
    return gmres_multi_bytes(n, ns, kIrOuter) + al(sizeof(float2) * ldn * ns * kIrOuter) + gmres_multi_bytes(n, ns, m);
}

hipError_t solve_ir(const MultiOp& hi, const MultiOp& lo, int64_t n, int ns, const float2* B, int64_t ldb, float2* X,
                    int64_t ldx, double tol, double inner_tol, int m, int max_inner, int max_outer, void* work,
                    IrInfo* info, hipStream_t s)
{
    const int64_t ldn = (n + 63) / 64 * 64;
    char* W = (char*)work;
    const size_t ob = gmres_multi_bytes(n, ns, kIrOuter);
    FlexPrecond prec;
    prec.op = &lo;
    prec.tol = inner_tol;
    prec.m = m;
    prec.maxiter = max_inner;
    prec.Z = (float2*)(W + ob);
    prec.work = W + ob + al(sizeof(float2) * (size_t)ldn * ns * kIrOuter);
    for (int q = 0; q < ns; ++q) TRY(hipMemsetAsync(X + (int64_t)q * ldx, 0, sizeof(float2) * n, s));
    MultiInfo mi;
    TRY(gmres_multi(hi, n, ns, B, ldb, X, ldx, tol, kIrOuter, max_outer, W, &mi, s, &prec));
    if (info) {
        info->outer = mi.iterations;
        info->inner = prec.inner_iterations;
        info->converged = mi.converged;
        info->rel_residual = mi.rel_residual;
    }
    return hipSuccess;
}

}  // namespace bem
