// Multi-right-hand-side GMRES and mixed-precision iterative refinement.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace bem {

// A dense operator applied to blocks of right-hand sides: complex64
// (dtype 0) or 16-bit planes (dtype kDtBF16 / kDtFP16, holding scale * A).
struct MultiOp {
    int dtype = 0;
    int64_t n = 0;
    const float2* A = nullptr;
    int64_t lda = 0;
    const uint16_t* Ar = nullptr;
    const uint16_t* Ai = nullptr;
    int64_t ldlp = 0;
    float inv_scale = 1.f;
    uint16_t* Xr = nullptr;  // scratch for the 16-bit right-hand sides (16 x ldlp each)
    uint16_t* Xi = nullptr;
    float2* part = nullptr;  // split-K partial sums of the products (null: single pass)
    hipError_t apply(const float2* X, int64_t ldx, int ns, float2* Y, int64_t ldy, float2 alpha, float2 beta,
                     hipStream_t s) const;
};

struct MultiInfo {
    int iterations = 0;
    int converged = 0;
    std::vector<double> rel_residual;
};

struct IrInfo {
    int outer = 0;
    int inner = 0;
    int converged = 0;
    std::vector<double> rel_residual;
};

// Flexible right preconditioner of gmres_multi: an inner multi-RHS GMRES on op
// to tol (restart m, at most maxiter iterations) with its own workspace; the
// preconditioned directions go to Z ((restart of the outer solve) x ns x ldn).
struct FlexPrecond {
    const MultiOp* op = nullptr;
    double tol = 1e-2;
    int m = 100;
    int maxiter = 1000;
    void* work = nullptr;
    float2* Z = nullptr;
    int inner_iterations = 0;
};

size_t lowp_scratch_bytes(int64_t ldlp);
float2* lowp_scratch_part(void* scratch, int64_t ldlp);
size_t gmres_multi_bytes(int64_t n, int ns, int m);
// rel (optional, ns values): per-system relative tolerances replacing tol.
// With prec, the inner target of outer step j of system q is relaxed to
// min(0.5, max(prec->tol, tol |b_q| / |r_j,q|)) (r_j: the outer residual).
hipError_t gmres_multi(const MultiOp& op, int64_t n, int ns, const float2* B, int64_t ldb, float2* X, int64_t ldx,
                       double tol, int m, int maxiter, void* work, MultiInfo* info, hipStream_t s,
                       FlexPrecond* prec = nullptr, const double* rel = nullptr);
size_t solve_ir_bytes(int64_t n, int ns, int m);
hipError_t solve_ir(const MultiOp& hi, const MultiOp& lo, int64_t n, int ns, const float2* B, int64_t ldb, float2* X,
                    int64_t ldx, double tol, double inner_tol, int m, int max_inner, int max_outer, void* work,
                    IrInfo* info, hipStream_t s);

}  // namespace bem
