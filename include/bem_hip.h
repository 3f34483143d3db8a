/*
 * bem_hip.h -- C ABI of the MI355X-native BEM Helmholtz library (libbem_hip.so).
 *
 * Drop-in for the reference's JIT torch extension CUDA_MODULE
 * (src/bem/cuda/bind.cu:178-223, loaded at src/bem/solver.py:9-31) and the
 * torch matvec/BiCGSTAB it is used with (src/bem/solver.py:61-66,
 * src/bem/BiCGSTAB.py:5-107).  Differences by design:
 *   - plain pointers and sizes, no torch types; all device pointers are HIP
 *     device memory of the current device; outputs are caller-allocated;
 *   - every call takes the HIP stream to launch on (hipStream_t as void*;
 *     NULL = the null stream) and returns a bem_status code; a message for
 *     the last failure on the calling thread is in bem_last_error();
 *   - the mesh is preprocessed once into a bem_mesh handle (normals, curls,
 *     quadrature points, neighbour pairs), which the reference recomputes in
 *     torch and in every singular launch (solver.py:40-58, bind.cu:90-97);
 *   - both bases: P1 (vertex DOFs, the reference's basis, V x V matrices)
 *     and DP0 (one DOF per triangle, N x N);
 *   - complex matrices are complex64 (float re, float im) row-major with a
 *     leading dimension ld (in complex elements); row = test function (x),
 *     column = trial function (y), as the reference.
 * No entry point raises C++ exceptions.  Calls on one stream are ordered;
 * a bem_mesh may be shared by threads (it is read-only after creation).
 */
#ifndef BEM_HIP_H
#define BEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    BEM_OK = 0,
    BEM_ERR_BAD_ARG = 1,      /* invalid argument (null pointer, size, enum) */
    BEM_ERR_BAD_MESH = 2,     /* index out of range, degenerate triangle, ... */
    BEM_ERR_HIP = 3,          /* a HIP runtime call failed */
    BEM_ERR_NOT_CONVERGED = 4,/* solver stopped before reaching the tolerance */
    BEM_ERR_NO_DEVICE = 5,    /* no HIP device */
    BEM_ERR_OOM = 6           /* device allocation failed */
} bem_status;

/* operator kinds (potential.h:7-15) */
typedef enum {
    BEM_OP_SINGLE = 0,         /* single layer  V   */
    BEM_OP_DOUBLE = 1,         /* double layer  K   */
    BEM_OP_ADJOINT_DOUBLE = 2, /* adjoint dbl.  K'  */
    BEM_OP_HYPERSINGULAR = 3,  /* hypersingular W (Maue form, P1 only) */
    BEM_OP_BM_LHS = 4,         /* -K + (i/k) W       (P1 only) */
    BEM_OP_BM_RHS = 5          /* -V - (i/k) K'      */
} bem_op;

typedef enum { BEM_BASIS_P1 = 0, BEM_BASIS_DP0 = 1 } bem_basis;

/* quadrature (bind.cu:180-215): full = 4-pt line / 6-pt triangle,
 * approx = 2 / 3, approx_1 = 2 / 1 */
typedef enum { BEM_QUAD_FULL = 0, BEM_QUAD_APPROX = 1, BEM_QUAD_APPROX_1 = 2 } bem_quad;

typedef struct bem_mesh bem_mesh;

typedef struct {
    int32_t iterations;          /* Krylov iterations (GMRES) or steps (BiCGSTAB) */
    int32_t converged;           /* 1 if the tolerance was reached */
    double rel_residual;         /* ||b - A x|| / ||b|| recomputed at exit */
    double implicit_rel_residual;/* GMRES least-squares estimate at exit */
} bem_solve_info;

const char* bem_version(void);
const char* bem_last_error(void);

/* Mesh handle.  verts: [nv,3] float32, tris: [nt,3] int32 (vertex order
 * gives the outward normal (v1-v0) x (v2-v0)).  on_device != 0: the
 * pointers are device memory (e.g. torch CUDA tensors), else host memory.
 * Replaces BEM_Solver.__init__ + preprocess (solver.py:40-58,70-81). */
int bem_mesh_create(const float* verts, int64_t nv, const int32_t* tris, int64_t nt, int on_device, void* stream,
                    bem_mesh** out);
int bem_mesh_destroy(bem_mesh* mesh);
/* New vertex positions [nv,3] for a mesh handle with the same triangles
 * (rigid or deformed objects in data generation: scene.py:218-291 builds a
 * new BEM_Solver per sample): the quadrature records, normals, curls and
 * geometry are recomputed and uploaded in one copy on `stream`, which is
 * waited for; neighbour pairs, vertex lists and the P1 tiling are kept.
 * Kernels still reading the old positions on other streams must be done
 * first.  Results equal those of a new handle built from the positions (P1
 * sums in another order: the tiling is the first positions'). */
int bem_mesh_set_vertices(bem_mesh* mesh, const float* verts, int on_device, void* stream);
int bem_mesh_info(const bem_mesh* mesh, int64_t* nv, int64_t* nt, int64_t* n_neighbor_pairs,
                  int32_t* max_neighbors);
/* unit normals [nt,3] and P1 surface curls [nt,3,3] (row a = n x grad phi_a)
 * into device memory (solver.py:40-58). */
int bem_mesh_normals_curls(const bem_mesh* mesh, float* normals, float* curls, void* stream);

/* Boundary operator matrix.  Replaces <op>_boundary_matrix[_approx[_1]]
 * (bind.cu:49-154, exports bind.cu:180-215).  P1: out is [nv, nv]; DP0:
 * [nt, nt].  clamp != 0 applies the reference's r >= 1e-3 clamp
 * (potential.h:16-23).  P1 accumulates the regular pairs in an [nv, nv]
 * complex scratch matrix (vertices renumbered along the tiles) taken from
 * the stream-ordered pool and released before return, then moves it to out;
 * without room for it, directly in out (see bem_assemble_ex bit 8). */
int bem_assemble(const bem_mesh* mesh, float k, int op, int basis, int quad, int clamp, void* out, int64_t ld,
                 void* stream);
/* bem_assemble restricted to parts of the pair set: parts = 1 regular pairs
 * only (P1: also zero-fills out first), 2 neighbour (singular) pairs only
 * (added / written on top of out), 3 both (= bem_assemble).  Lets a caller
 * time or overlap the two kernels separately.  Adding 4 makes the P1 regular
 * pairs one evaluation per ordered triangle pair; without it each unordered
 * pair is evaluated once where that is faster (single layer and hypersingular
 * by symmetry; double layer and adjoint double layer with quad 0, on meshes of
 * at least 4096 vertices, for both orders at once).  Adding 8 makes the P1
 * regular pairs accumulate
 * directly in out (no [nv, nv] scratch matrix per wavenumber: the kernels then
 * flush to scattered addresses, slower).  The results agree to rounding. */
int bem_assemble_ex(const bem_mesh* mesh, float k, int op, int basis, int quad, int clamp, int parts, void* out,
                    int64_t ld, void* stream);
/* The system neumann2dirichlet solves (solver.py:118-126) in the P1 basis:
 * lhs = bm_lhs + M/2 ([nv, ld]) and rhs = (bm_rhs - (i/2k) M) g ([nv], g the
 * Neumann data per vertex), quad 0 / 1 / 2 as bem_assemble (the reference's
 * neumann2dirichlet uses the approximate rules, quad 1).  The reference
 * assembles bm_rhs as a matrix and multiplies; here its entries are
 * contracted with g inside the bm_lhs kernels (never stored), so the system
 * costs about one operator assembly instead of two plus a product.  Results
 * agree with that route to rounding. */
int bem_burton_miller_system(const bem_mesh* mesh, float k, int quad, int clamp, const void* g, void* lhs, int64_t ld,
                             void* rhs, void* stream);
/* The same at nk wavenumbers on one mesh (the modes of scene.py:384-397):
 * ks host [nk], g [nk, nv] contiguous, system z at lhs + z * stride complex
 * elements ([nv, ld] each) and rhs + z * nv; one launch per kernel for up to
 * 64 systems. */
int bem_burton_miller_systems(const bem_mesh* mesh, const float* ks, int nk, int quad, int clamp, const void* g,
                              void* lhs, int64_t ld, int64_t stride, void* rhs, void* stream);
/* The derivative with respect to k of the P1 operator op (SINGLE, DOUBLE,
 * ADJOINT_DOUBLE, BM_LHS, BM_RHS; regular and neighbour pairs, no identity
 * terms), [nv, nv] at out, same rules and clamp as bem_assemble.  No
 * counterpart in the reference (it has no gradients): the wavenumber
 * derivative the adjoint of the Burton-Miller pipeline (neumann2dirichlet,
 * solver.py:118-126, operators bind.cu:180-215) needs. */
int bem_assemble_dk(const bem_mesh* mesh, float k, int op, int quad, int clamp, void* out, int64_t ld, void* stream);
/* bem_assemble at nk wavenumbers on one mesh (e.g. the modes of a vibrating
 * object, each solved by neumann2dirichlet: scene.py:384-397 loops over them
 * one call per operator): one launch per kernel for up to 64 wavenumbers.
 * ks: host array [nk]; operator z is written at out + z * stride complex
 * elements, [n, ld] each (stride >= n * ld).  Same results as nk calls of
 * bem_assemble (the P1 double layer and its adjoint use the ordered-pair
 * sweep here). */
int bem_assemble_batch_k(const bem_mesh* mesh, const float* ks, int nk, int op, int basis, int quad, int clamp,
                         void* out, int64_t ld, int64_t stride, void* stream);
/* Geometry sets: ns placements of one mesh's triangles, e.g. the samples of
 * a scene whose objects are resized, rotated and moved (scene.py:218-291,
 * where the reference builds a new BEM_Solver and assembles each sample on
 * its own).  verts: host [ns, nv, 3] float32.  Each sample's
 * position-dependent data (quadrature records, normals, curls, singular-pair
 * geometry) is built on the host and uploaded in one copy (the stream is
 * waited for); the mesh's connectivity (P1 tiling, neighbour pairs, vertex
 * lists) serves every sample, so the mesh must outlive the set.  Errors:
 * BAD_MESH for a degenerate triangle in any sample. */
typedef struct bem_geomset bem_geomset;
int bem_geomset_create(const bem_mesh* mesh, const float* verts, int ns, void* stream, bem_geomset** out);
int bem_geomset_destroy(bem_geomset* set);
int bem_geomset_count(const bem_geomset* set);
/* bem_assemble for every sample z of the set at its own wavenumber ks[z]
 * (host [ns]), in one launch per kernel for up to 64 samples: operator z at
 * out + z * stride complex elements, [n, ld] each (stride >= n * ld). */
int bem_assemble_samples(const bem_geomset* set, const float* ks, int op, int basis, int quad, int clamp, void* out,
                         int64_t ld, int64_t stride, void* stream);
/* out_z += alpha_z M_z (M_z the mass matrix of sample z), alphas host [ns]
 * complex as (re, im) pairs. */
int bem_add_identity_samples(const bem_geomset* set, int basis, const float* alphas, void* out, int64_t ld,
                             int64_t stride, void* stream);
/* bem_potential for every sample z at its own points (device [ns, np, 3])
 * and wavenumber ks[z]: matrix z at out + z * stride ([np, ld] each). */
/* The fused Burton-Miller systems (bem_burton_miller_system) of every sample
 * of a geometry set, each at its wavenumber (ks host [ns]): g [ns, nv]
 * contiguous, system z at lhs + z * stride and rhs + z * nv. */
int bem_burton_miller_samples(const bem_geomset* set, const float* ks, int quad, int clamp, const void* g, void* lhs,
                              int64_t ld, int64_t stride, void* rhs, void* stream);
int bem_potential_samples(const bem_geomset* set, const float* points, int64_t np, const float* ks, int op, int basis,
                          int clamp, void* out, int64_t ld, int64_t stride, void* stream);
/* y_z = alpha A_z x_z + beta y_z for nz complex64 matrices [rows, cols] (A_z at
 * A + z * stride_a, x_z at x + z * stride_x, y_z at y + z * stride_y, complex
 * elements), in one launch; y is not read when beta = 0 (e.g. the right-hand
 * sides and pressures of a batch of samples, solver.py:125,147). */
int bem_matvec_strided(const void* A, int64_t rows, int64_t cols, int64_t lda, int64_t stride_a, const void* x,
                       int64_t stride_x, void* y, int64_t stride_y, int nz, float alpha_re, float alpha_im,
                       float beta_re, float beta_im, void* stream);
/* DP0 rows [row0, row1) only (out has row1-row0 rows); rows are independent,
 * so a large operator can be assembled block by block. */
int bem_assemble_rows(const bem_mesh* mesh, float k, int op, int quad, int clamp, int64_t row0, int64_t row1,
                      void* out, int64_t ld, void* stream);
/* Stateless form of bem_assemble on device pointers (P1, the reference's
 * single_boundary_matrix(vertices, triangles, ...) call shape). */
int bem_boundary_matrix(const float* verts, int64_t nv, const int32_t* tris, int64_t nt, float k, int op, int quad,
                        void* out, void* stream);

/* The DP0 CBIE system of solver.py:97-104 in one pass: lhs = M/2 - K
 * ([nt, nt]) and rhs = -V psi ([nt]) for psi [nt] (complex64, device).  V
 * is evaluated alongside K (shared distances and e^{ikr}) and contracted with
 * psi on the fly, never stored.  work: bem_cbie_work_size() bytes of device
 * scratch.  Deterministic (fixed summation order). */
int64_t bem_cbie_work_size(const bem_mesh* mesh);
int bem_cbie_dp0(const bem_mesh* mesh, float k, int quad, int clamp, const void* psi, void* lhs, int64_t ld,
                 void* rhs, void* work, void* stream);
/* bem_cbie_dp0 in two halves: parts = 1 the fused regular-pair kernel,
 * 2 the neighbour pairs and the final reduction (after 1), 3 both. */
int bem_cbie_dp0_ex(const bem_mesh* mesh, float k, int quad, int clamp, int parts, const void* psi, void* lhs,
                    int64_t ld, void* rhs, void* work, void* stream);

/* Mass matrix: P1 [nv,nv] (bind.cu:10-21), DP0 diag(area) [nt,nt]. */
int bem_identity(const bem_mesh* mesh, int basis, void* out, int64_t ld, void* stream);
/* A += alpha * M in place (M the mass matrix of the basis), e.g. the 0.5 M
 * of the CBIE left-hand side (solver.py:102) without a dense identity. */
int bem_add_identity(const bem_mesh* mesh, int basis, float alpha_re, float alpha_im, void* A, int64_t ld,
                     void* stream);
/* Vertex average of per-triangle complex values (bind.cu:23-48). */
int bem_triangle2vertex(const bem_mesh* mesh, const void* tri_values, void* vert_values, void* stream);
/* The same without a mesh handle, on device triangles [nt,3] (the reference's
 * triangle2vertex(vertices, triangles, features) call shape, bind.cu:23-25,217,
 * which reads no positions): the vertex -> triangle lists are built on the
 * device in work (bem_triangle2vertex_work_size bytes) and each vertex sums
 * its triangles in ascending order (deterministic).  Waits for the stream;
 * returns BEM_ERR_BAD_MESH for an index outside [0, nv).  A vertex in no
 * triangle gets NaN (0/0), as the reference. */
int64_t bem_triangle2vertex_work_size(int64_t nv, int64_t nt);
int bem_triangle2vertex_raw(const int32_t* tris, int64_t nt, int64_t nv, const void* tri_values, void* vert_values,
                            void* work, void* stream);
/* Potential operator from the surface to np points ([np,3] float32, device):
 * out [np, nv] (P1) or [np, nt] (DP0); op = SINGLE or DOUBLE
 * (single/double_boundary_potential, bind.cu:156-176,219-222). */
int bem_potential(const bem_mesh* mesh, const float* points, int64_t np, float k, int op, int basis, int clamp,
                  void* out, int64_t ld, void* stream);
/* The derivative with respect to k of bem_potential's matrix (SINGLE or
 * DOUBLE, either basis).  No counterpart in the reference: the wavenumber
 * derivative of boundary2potential (solver.py:140-147) for its adjoint. */
int bem_potential_dk(const bem_mesh* mesh, const float* points, int64_t np, float k, int op, int basis, int clamp,
                     void* out, int64_t ld, void* stream);
/* bem_potential at nk wavenumbers (ks: host array [nk]) in one launch per
 * kernel for up to 64 of them: matrix z at out + z * stride complex elements,
 * [np, ld] each (stride >= np * ld). */
int bem_potential_batch_k(const bem_mesh* mesh, const float* points, int64_t np, const float* ks, int nk, int op,
                          int basis, int clamp, void* out, int64_t ld, int64_t stride, void* stream);

/* y = alpha op(A) x + beta y, A complex64 [rows, cols] row-major, trans:
 * 0 = A, 1 = A^T, 2 = A^H.  work: device scratch of
 * bem_matvec_work_size() bytes (may be NULL when that is 0).  Replaces the
 * torch `@` (solver.py:65,103,131,137,147). */
int64_t bem_matvec_work_size(int64_t rows, int64_t cols, int trans);
int bem_matvec(const void* A, int64_t rows, int64_t cols, int64_t lda, int trans, const void* x, void* y,
               float alpha_re, float alpha_im, float beta_re, float beta_im, void* work, void* stream);
/* C = a A + b B elementwise (A or B may be NULL); e.g. 0.5 M - K
 * (solver.py:102). */
int bem_combine(const void* A, int64_t lda, const void* B, int64_t ldb, void* C, int64_t ldc, int64_t rows,
                int64_t cols, float a_re, float a_im, float b_re, float b_im, void* stream);

/* Restarted GMRES(restart) on a dense complex64 [n, n] matrix, device
 * resident (classical Gram-Schmidt + re-orthogonalisation, Givens on the
 * device); stops when ||b - A x|| <= tol ||b||.  x: initial guess in,
 * solution out.  Returns BEM_ERR_NOT_CONVERGED (x still the last iterate)
 * when maxiter runs out.  No reference counterpart (the reference solves
 * with BiCGSTAB, solver.py:61-66). */
int64_t bem_gmres_work_size(int64_t n, int restart);
int bem_gmres(const void* A, int64_t n, int64_t lda, const void* b, void* x, float tol, int restart, int maxiter,
              void* work, bem_solve_info* info, void* stream);
/* bem_gmres on op(A): trans 0 = A, 1 = A^T, 2 = A^H (e.g. the adjoint
 * solve of a gradient, reusing the assembled matrix). */
int bem_gmres_ex(const void* A, int64_t n, int64_t lda, int trans, const void* b, void* x, float tol, int restart,
                 int maxiter, void* work, bem_solve_info* info, void* stream);
/* bem_gmres_ex that also returns the summed HIP-event time (ms) and the
 * number of its matrix-vector products (either pointer may be NULL), for
 * measuring the product kernel inside a solve.  No reference counterpart. */
int bem_gmres_timed(const void* A, int64_t n, int64_t lda, int trans, const void* b, void* x, float tol, int restart,
                    int maxiter, void* work, bem_solve_info* info, double* gemv_ms, int64_t* gemv_count,
                    void* stream);
/* bem_gmres_ex that allocates nothing: no workspace argument; the workspace
 * is a static device array of BEM_GMRES_NOALLOC_SCRATCH_BYTES (64 MiB) that
 * each GPU running the library holds until the process ends.  No counterpart
 * in the reference: for callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc).  Same results as bem_gmres_ex.
 * Returns BEM_ERR_BAD_ARG when bem_gmres_work_size(n, restart) exceeds the
 * array (e.g. n = 20480 takes restart <= 240).  Calls that overlap on one GPU
 * (other streams or threads) share the array and race: serialise them. */
#define BEM_GMRES_NOALLOC_SCRATCH_BYTES (64ll << 20)
int64_t bem_gmres_noalloc_capacity(void);
int bem_gmres_noalloc(const void* A, int64_t n, int64_t lda, int trans, const void* b, void* x, float tol,
                      int restart, int maxiter, bem_solve_info* info, void* stream);
/* Reference-compatible BiCGSTAB (BiCGSTAB.py:5-107): bilinear rho, alpha,
 * omega; stops when ||b - A x||^2 < tol ||b||^2 or < atol; stops silently
 * (returns BEM_OK, info->converged = 0) when nsteps run out or rho == 0. */
int64_t bem_bicgstab_work_size(int64_t n);
int bem_bicgstab(const void* A, int64_t n, int64_t lda, const void* b, void* x, int nsteps, double tol, double atol,
                 void* work, bem_solve_info* info, void* stream);

/* ---- mixed precision (no reference counterpart; the reference is FP32 only) ----
 * A complex64 matrix is stored as two 16-bit planes Ar, Ai (dtype 1 = bf16,
 * 2 = fp16) holding round(scale * A), with a column pitch ldlp that is a
 * multiple of 32 (bem_lowp_ld).  The planes are tiled for the matrix cores
 * (16 x 32 tiles, each the 1 KB a wavefront reads in one step; zero padding to
 * 16 rows and ldlp columns): each plane holds bem_lowp_plane_elems(rows, cols)
 * 16-bit entries.  The products (any number of right-hand sides, 16 per pass)
 * run on the matrix cores with FP32 accumulation. */
int64_t bem_lowp_ld(int64_t cols);
int64_t bem_lowp_plane_elems(int64_t rows, int64_t cols);
/* scale <= 0 picks it: 1 for bf16, the power of two putting max|A| at 2^14
 * for fp16; the value used is returned in *scale_used. */
int bem_convert_lowp(const void* A, int64_t rows, int64_t cols, int64_t lda, int dtype, float scale, void* Ar,
                     void* Ai, int64_t ldlp, float* scale_used, void* stream);
/* Y = alpha A X + beta Y for nrhs right-hand sides (X: nrhs systems of cols
 * entries at stride ldx; Y likewise with ldy), complex64 A on FP32 matrix
 * cores.  work (bem_matmul_work_size(rows) bytes, or NULL) holds the partial
 * sums of a column split, which fills the GPU when the row tiles alone do
 * not (summed in a fixed order: the result does not depend on the split);
 * with NULL every product is one pass over the columns. */
int64_t bem_matmul_work_size(int64_t rows);
int bem_matmul(const void* A, int64_t rows, int64_t cols, int64_t lda, const void* X, int64_t ldx, int nrhs, void* Y,
               int64_t ldy, float alpha_re, float alpha_im, float beta_re, float beta_im, void* work, void* stream);
/* Same on the 16-bit planes (result multiplied by inv_scale = 1/scale). */
int64_t bem_matmul_lowp_work_size(int64_t ldlp);
int bem_matmul_lowp(const void* Ar, const void* Ai, int dtype, int64_t rows, int64_t cols, int64_t ldlp,
                    float inv_scale, const void* X, int64_t ldx, int nrhs, void* Y, int64_t ldy, float alpha_re,
                    float alpha_im, float beta_re, float beta_im, void* work, void* stream);
/* GMRES on nrhs systems sharing one matrix (complex64 A when dtype = 0, else
 * the 16-bit planes), in lockstep: one multi-RHS product per Arnoldi step,
 * per-system Gram-Schmidt, Givens and convergence on the device.  info:
 * nrhs entries. */
int64_t bem_gmres_multi_work_size(int64_t n, int nrhs, int restart, int64_t ldlp);
int bem_gmres_multi(const void* A, int64_t lda, const void* Ar, const void* Ai, int dtype, int64_t ldlp,
                    float inv_scale, int64_t n, int nrhs, const void* B, int64_t ldb, void* X, int64_t ldx, float tol,
                    int restart, int maxiter, void* work, bem_solve_info* info, void* stream);
/* Mixed-precision refinement with a Krylov outer loop: x = 0; flexible GMRES
 * in FP32 on the complex64 A (residuals and updates in FP32, basis of 20,
 * restarted beyond it, at most max_outer outer steps), each outer direction
 * preconditioned by a 16-bit multi-RHS GMRES solve (restart, max_inner) to
 * max(inner_tol, tol ||b|| / ||r_j||) (r_j: the outer residual; capped at
 * 0.5); until ||b - A x|| <= tol ||b|| for every system.
 * info[q].iterations = total inner iterations, *outer_iterations = outer
 * steps. */
int64_t bem_solve_ir_work_size(int64_t n, int nrhs, int restart, int64_t ldlp);
int bem_solve_ir(const void* A, int64_t lda, const void* Ar, const void* Ai, int dtype, int64_t ldlp, float inv_scale,
                 int64_t n, int nrhs, const void* B, int64_t ldb, void* X, int64_t ldx, float tol, float inner_tol,
                 int restart, int max_inner, int max_outer, void* work, bem_solve_info* info,
                 int32_t* outer_iterations, void* stream);

/* ---- differentiation (no reference counterpart; the reference has no
 * gradients).  For the DP0 CBIE pressure p = D_pot phi - S_pot psi with
 * (M/2 - K) phi = -V psi and a real loss L with pbar = dL/dp* (half of torch's
 * complex gradient) and lambda solving (M/2 - K)^H lambda = D_pot^H pbar, these
 * ADD into FP64 accumulators:
 *   grad_vertices [nv,3], grad_k [1]: the parts of dL/dv, dL/dk through K, V,
 *     M (bem_adjoint_cbie_dp0) and through S_pot, D_pot (bem_adjoint_potential_dp0);
 *   vhl [nt] (complex, as 2 doubles): V^H lambda (needed for dL/dpsi).
 * Derivative blocks are recomputed per Gauss-point pair, never stored. */
int64_t bem_adjoint_work_size(const bem_mesh* mesh);
int bem_adjoint_cbie_dp0(const bem_mesh* mesh, float k, int clamp, const void* lam, const void* psi, const void* phi,
                         double* grad_vertices, double* grad_k, double* vhl, void* work, void* stream);
int bem_adjoint_potential_dp0(const bem_mesh* mesh, const float* points, int64_t np, float k, int clamp,
                              const void* pbar, const void* psi, const void* phi, double* grad_vertices,
                              double* grad_k, void* stream);
/* The same for the P1 Burton-Miller pressure (neumann2dirichlet +
 * boundary2potential with the approximate rules, solver.py:118-126,140-147):
 * (bm_lhs + M/2) phi = (bm_rhs - (i/2k) M) g, p = D_pot phi - S_pot g, lambda
 * solving (bm_lhs + M/2)^H lambda = D_pot^H pbar; all vectors P1 (one complex
 * value per vertex).  bem_adjoint_bm_p1 adds 2 Re lambda^H (dR g - dL phi)
 * (regular and neighbour pairs, curl terms, mass terms) to grad_vertices;
 * bem_adjoint_potential_p1 adds 2 Re pbar^H (dD_pot phi - dS_pot g) to
 * grad_vertices and, when grad_points [np,3] is not NULL, the listener
 * gradients to it.  The k gradient comes from bem_assemble_dk /
 * bem_potential_dk. */
int bem_adjoint_bm_p1(const bem_mesh* mesh, float k, int clamp, const void* lam, const void* g, const void* phi,
                      double* grad_vertices, void* stream);
int bem_adjoint_potential_p1(const bem_mesh* mesh, const float* points, int64_t np, float k, int clamp,
                             const void* pbar, const void* g, const void* phi, double* grad_vertices,
                             double* grad_points, void* stream);

/* ---- batched items (no reference counterpart; the reference solves one
 * (mesh, k) item at a time, src/scene.py:270-291) ---- */
typedef struct {
    const bem_mesh* mesh;
    float k;
    const void* psi; /* [nt] complex64 Neumann data */
    void* lhs;       /* [nt, ld] complex64 out: M/2 - K */
    int64_t ld;
    void* rhs;       /* [nt] complex64 out: -V psi */
} bem_batch_item;
/* The DP0 CBIE systems of n_items items (any sizes, any wavenumbers) with the
 * regular pairs of all items in ONE persistent launch over a device-side
 * queue of (item, 64-column block, row tile) wavefront tasks, largest items
 * first.  row_tile: rows per task (16/32/64) or 0 for the value
 * bem_batch_autotune stored for this size class (64 if none).  The result
 * does not depend on row_tile (bit for bit). */
int64_t bem_batch_cbie_work_size(const bem_batch_item* items, int n_items);
int bem_batch_cbie(const bem_batch_item* items, int n_items, int clamp, int row_tile, void* work, void* stream);
/* Times the three row tiles on the largest item, keeps the fastest for its
 * size class (process-wide, mutex-guarded), returns it and the times
 * (ms3[0..2] for 16, 32, 64 rows). */
int bem_batch_autotune(const bem_batch_item* items, int n_items, int clamp, void* work, int* best_row_tile,
                       float* ms3, void* stream);
typedef struct {
    const void* A; /* [n, lda] complex64 */
    int64_t lda;
    int64_t n;
    const void* b; /* [n] */
    void* x;       /* [n] initial guess in, solution out */
} bem_batch_system;
/* GMRES(restart) on all systems in lockstep (different matrices and sizes);
 * with use_graph != 0 the parts of a restart cycle (start, segments of 4,
 * then 20 Arnoldi steps, end) are replayed from hipGraphs, each captured on
 * its third use (a solve done in one or two cycles captures nothing), and the
 * cycle ends early once every
 * system meets its target; stops when every system has
 * ||b - A x|| <= tol ||b|| or after max_cycles cycles.  info[q].iterations =
 * Arnoldi steps taken by system q. */
int64_t bem_batch_gmres_work_size(const bem_batch_system* sys, int n_sys, int restart);
int bem_batch_gmres(const bem_batch_system* sys, int n_sys, float tol, int restart, int max_cycles, int use_graph,
                    void* work, bem_solve_info* info, void* stream);
/* Batched mixed-precision solve: flexible GMRES(outer_restart) in FP32 on
 * every system's complex64 matrix (products, residuals and updates in FP32),
 * whose directions z_j ~ A^-1 v_j come from lockstep GMRES(inner_restart) on
 * a 16-bit copy of the matrix (dtype 1 = bf16, 2 = fp16 with a per-system
 * power-of-two scale) or an 8-bit one (dtype 3 = OCP e4m3 with a
 * power-of-two scale per row: half the bytes of the 16-bit copies, about 1 %
 * from A, so more outer steps), the diagonal kept in FP32, made inside the
 * call in the workspace, run per system to ||v_j - Alow z_j|| <=
 * max(inner_tol, tol ||b|| / ||r_j||) (r_j: the current outer residual;
 * capped at 0.5) or for max_inner_cycles cycles.  Stops when every system has
 * ||b - A x|| <= tol ||b|| (FP32 residual) or after max_outer_cycles cycles.
 * info[q].iterations = outer steps; inner_iterations[q] (optional) =
 * low-precision Arnoldi steps. */
int64_t bem_batch_fgmres_work_size(const bem_batch_system* sys, int n_sys, int outer_restart, int inner_restart);
int bem_batch_fgmres(const bem_batch_system* sys, int n_sys, int dtype, float tol, float inner_tol, int outer_restart,
                     int inner_restart, int max_inner_cycles, int max_outer_cycles, int use_graph, void* work,
                     bem_solve_info* info, int* inner_iterations, void* stream);
const char* bem_batch_last_error(void);

/* ---- FEM modal analysis (linear tetrahedra, 3 degrees of freedom per vertex,
 * dof 3 v + c): the mass and stiffness matrices of the solid whose modes give
 * the BEM's Neumann data.  Replaces the reference's assemble_mass_matrix /
 * assemble_stiffness_matrix (src/modalobj/cuda/computeMatrix.cu:130-166,
 * called by FEMmodel, src/modalobj/fem.py:95-166), which write 144 COO
 * triplets per tetrahedron; here both matrices are 3 x 3 block CSR over the
 * vertex adjacency (rows and columns sorted), assembled deterministically. */
typedef struct bem_fem bem_fem;
/* verts [nv,3] f32, tets [ntet,4] i32, host memory.  Errors: BAD_MESH for an
 * index out of range, a repeated vertex or a zero-volume tetrahedron. */
int bem_fem_create(const float* verts, int64_t nv, const int32_t* tets, int64_t ntet, bem_fem** out);
void bem_fem_free(bem_fem* fem);
int64_t bem_fem_num_blocks(const bem_fem* fem);
/* device int32 outputs: row_ptr [nv+1], block_rows [nblocks], cols [nblocks] */
int bem_fem_pattern(const bem_fem* fem, void* row_ptr, void* block_rows, void* cols, void* stream);
/* values: device [nblocks, 3, 3], f32 (fp64 = 0, the reference's type) or
 * f64 (fp64 = 1; summed in FP64 either way); which 0 = mass (consistent,
 * density), 1 = stiffness (isotropic linear elasticity: Young's modulus,
 * Poisson's ratio). */
int bem_fem_assemble(const bem_fem* fem, int which, float density, float youngs, float poisson, int fp64,
                     void* values, void* stream);
/* Y = A X for A = values (block CSR of fem), X and Y row-major [3 nv, m]
 * (leading dimensions ldx, ldy >= m), device, all f32 (fp64 = 0) or f64. */
int bem_fem_spmm(const bem_fem* fem, int fp64, const void* values, const void* X, int64_t ldx, int m, void* Y,
                 int64_t ldy, void* stream);
const char* bem_fem_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* BEM_HIP_H */
