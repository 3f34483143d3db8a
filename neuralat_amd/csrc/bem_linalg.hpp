// Launchers of the dense linear-algebra kernels (bem_linalg.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bem {

constexpr int kGemvRows = 4;  // rows per wavefront in the streaming GEMV

// One system of a lockstep batch: its matrix A [n, lda] and its vectors at
// offset off of the batch's concatenated vector space.
struct BSys {
    const float2* A;
    int64_t lda;
    int64_t n;
    int64_t off;
};

// New Hessenberg column j = h + h2 (the two Gram-Schmidt passes), rotated by
// the j previous Givens rotations, then the new rotation and the residual
// estimate.  The rotations are a serial chain: the workgroup stages them (and
// the column) in LDS 256 at a time so that thread 0 never waits on a global
// load inside the chain; the value carried from one rotation to the next
// stays in registers.  Returns, in thread 0, rotated entry j (the value
// entering the new rotation); called by every thread of the workgroup
// (at least kGivensChunk threads).
constexpr int kGivensChunk = 256;
__device__ inline double2 rotate_column_wg(double2* __restrict__ col, int j, const double2* __restrict__ h,
                                           const double2* __restrict__ h2, const double2* __restrict__ cs,
                                           const double2* __restrict__ sn)
{
    __shared__ double2 scs[kGivensChunk], ssn[kGivensChunk], sb[kGivensChunk];
    const int t = threadIdx.x;
    double2 carry = make_double2(h[0].x + h2[0].x, h[0].y + h2[0].y);  // col[q] entering rotation q
    for (int q0 = 0; q0 < j; q0 += kGivensChunk) {
        const int cnt = min(kGivensChunk, j - q0);
        if (t < cnt) {
            scs[t] = cs[q0 + t];
            ssn[t] = sn[q0 + t];
            const int e = q0 + t + 1;
            sb[t] = make_double2(h[e].x + h2[e].x, h[e].y + h2[e].y);
        }
        __syncthreads();
        if (t == 0) {
            for (int u = 0; u < cnt; ++u) {
                // [c s; -conj(s) conj(c)] applied to (col[q], col[q+1])
                const double2 c = scs[u], sv = ssn[u], a = carry, b = sb[u];
                const double2 top = make_double2(c.x * a.x - c.y * a.y + sv.x * b.x - sv.y * b.y,
                                                 c.x * a.y + c.y * a.x + sv.x * b.y + sv.y * b.x);
                carry = make_double2(-(sv.x * a.x + sv.y * a.y) + (c.x * b.x + c.y * b.y),
                                     -(sv.x * a.y - sv.y * a.x) + (c.x * b.y - c.y * b.x));
                col[q0 + u] = top;
            }
        }
        __syncthreads();
    }
    return carry;
}

// y_z = alpha A_z x_z + beta y_z for nz matrices [rows, cols] (A_z at A + z sa,
// x_z at x + z sx, y_z at y + z sy; elements)
hipError_t launch_gemv_strided(const float2* A, int64_t rows, int64_t cols, int64_t lda, int64_t sa, const float2* x,
                               int64_t sx, float2* y, int64_t sy, int nz, float2 alpha, float2 beta, hipStream_t s);
hipError_t launch_gemv(const float2* A, int64_t rows, int64_t cols, int64_t lda, int trans, const float2* x, float2* y,
                       float2 alpha, float2 beta, float2* work, hipStream_t s);
int64_t gemv_work_elems(int64_t rows, int64_t cols, int trans);
hipError_t launch_combine(const float2* A, int64_t lda, const float2* B, int64_t ldb, float2* C, int64_t ldc,
                          int64_t rows, int64_t cols, float2 a, float2 b, hipStream_t s);
hipError_t launch_multidot(const float2* V, int64_t ldv, int64_t n, int nq, const float2* w, double2* h,
                           double2* part, hipStream_t s);
// scratch (double2 entries) of launch_multidot's split form for nq basis vectors
int64_t multidot_part_elems(int nq);
hipError_t launch_gs_update(const float2* V, int64_t ldv, int64_t n, int nq, const double2* h, float2* w,
                            hipStream_t s);
hipError_t launch_norm2(const float2* w, int64_t n, double* out, hipStream_t s);
hipError_t launch_dot(const float2* a, const float2* b, int64_t n, int conj, double2* out, hipStream_t s);
hipError_t launch_normalize(const float2* src, int64_t n, const double* nrm2, float2* out, hipStream_t s);
hipError_t launch_givens(double2* H, int m, int j, const double2* h, const double2* h2, const double* wn2,
                         double2* cs, double2* sn, double2* g, double* res, hipStream_t s);
// one Arnoldi step after w = A v_j (Gram-Schmidt twice, norm into *wn2,
// v_{j+1}, Givens) in one workgroup, for systems of at most kArnoldiSmallN
constexpr int64_t kArnoldiSmallN = 4096;
hipError_t launch_arnoldi_small(float2* V, int64_t n, int j, float2* w, double2* h, double2* h2, double2* H, int m,
                                double2* cs, double2* sn, double2* g, double* res, double* wn2, hipStream_t s);
hipError_t launch_backsolve(const double2* H, int m, int nk, const double2* g, double2* y, hipStream_t s);
hipError_t launch_update_x(const float2* V, int64_t ldv, int64_t n, int nk, const double2* y, float2* x,
                           hipStream_t s);
hipError_t launch_init_g(double2* g, int m, const double* nrm2, hipStream_t s);
hipError_t launch_bicg_scalar(int which, double2* sc, hipStream_t s);
hipError_t launch_bicg_p(const float2* r, float2* p, const float2* v, const double2* sc, int64_t n, hipStream_t s);
hipError_t launch_bicg_axpy(const float2* a, const float2* b, const double2* sc, int idx, double sign, float2* out,
                            int64_t n, hipStream_t s);
hipError_t launch_bicg_x(float2* x, const float2* p, const float2* sv, const double2* sc, int64_t n, hipStream_t s);

// batched Krylov kernels (ns systems, system-major vectors)
hipError_t launch_bmultidot(const float2* V, int64_t ldn, int ns, int64_t n, int nq, const float2* w, double2* h,
                            int hs, const int* done, hipStream_t s);
hipError_t launch_bgs_update(const float2* V, int64_t ldn, int ns, int64_t n, int nq, const double2* h, int hs,
                             float2* w, const int* done, hipStream_t s);
hipError_t launch_bnorm2(const float2* w, int64_t ldn, int64_t n, int ns, double* out, hipStream_t s);
hipError_t launch_bnormalize(const float2* src, int64_t lds, int64_t n, int ns, const double* nrm2, float2* dst,
                             int64_t ldd, const int* done, hipStream_t s);
hipError_t launch_binit(double2* g, int m, const double* nrm2, const double* target, int* done, int* nk, int ns,
                        hipStream_t s);
hipError_t launch_bgivens(double2* H, int m, int j, const double2* h, const double2* h2, int hs, const double* wn2,
                          double2* cs, double2* sn, double2* g, double* res, const double* target, int* done, int* nk,
                          int ns, hipStream_t s);
hipError_t launch_bbacksolve(const double2* H, int m, const int* nk, const double2* g, double2* y, int ns,
                             hipStream_t s);
// One lockstep Arnoldi step after w = A v_j, one workgroup per system (meant
// for small systems, whose steps are launch-bound): Gram-Schmidt (passes = 1
// or 2; with 1, h2 is left as it is), |w|^2 into wn2, v_{j+1}, and the
// Givens bookkeeping of launch_bgivens.  Vector q of system s at
// V + q ntot + sys[s].off; h, h2 [ns, hs]; H [ns, (m + 1) m].
hipError_t launch_barnoldi_small(const BSys* sys, int ns, float2* V, int64_t ntot, int j, float2* w, double2* h,
                                 double2* h2, int hs, double2* H, int m, double2* cs, double2* sn, double2* g,
                                 double* res, const double* target, int* done, int* nk, double* wn2, int passes,
                                 hipStream_t s);
hipError_t launch_bupdate_x(const float2* V, int64_t ldn, int ns, int64_t n, const int* nk, const double2* y, int m,
                            float2* x, int64_t ldx, hipStream_t s);
hipError_t launch_bcopy(const float2* src, int64_t lds, float2* dst, int64_t ldd, int64_t n, int ns, hipStream_t s);
hipError_t launch_badd(float2* x, int64_t ldx, const float2* d, int64_t ldd, int64_t n, int ns, hipStream_t s);

}  // namespace bem
