// Library-internal types shared by the host code and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace bem {

// DP0 regular kernel: rows per workgroup (each lane computes this many
// entries of its column).
constexpr int kDp0RowTile = 64;
// Potential kernel: points per workgroup.
constexpr int kPotRowTile = 16;
// P1 tiles: at most kP1RowTris row triangles touching at most kP1RowVerts
// vertices, at most kP1ColTris column triangles touching at most kP1ColVerts
// vertices (the LDS accumulator is kP1RowVerts x kP1ColVerts complex, the
// per-row block stage 9 x kP1ColTris complex; a 256-thread workgroup has one
// lane per column triangle and one per column vertex).
constexpr int kP1RowTris = 16;
constexpr int kP1RowVerts = 20;
constexpr int kP1ColTris = 256;
constexpr int kP1ColVerts = 200;

// Locality-ordered tiling of the triangle list for the P1 scatter.
struct P1Tiles {
    const int32_t* perm;       // [nt] triangle id at ordered position
    const int32_t* row_ptr;    // [nrow_tiles+1] position ranges
    const int32_t* row_vptr;   // [nrow_tiles+1] ranges into row_verts
    const int32_t* row_verts;  // global vertex ids of each row tile
    const int4* row_slot;      // [nt] per position: tile-local slots of the 3 vertices
    const int32_t* col_ptr;
    const int32_t* col_vptr;
    const int32_t* col_verts;
    const int4* col_slot;
    // column-vertex incidences: for tile vertex slot g (col_vptr[tile] + w),
    // col_inc[col_inc_ptr[g] .. col_inc_ptr[g + 1]) = 4 * (position - col_ptr[tile]) + b
    // for each triangle of the tile whose local vertex b is that vertex
    const int32_t* col_inc_ptr;
    const int32_t* col_inc;
    // row tiles nest in column tiles: row_ctile[r] = the column tile holding
    // row tile r; the symmetric sweep pairs row tile r with column tiles
    // row_ctile[r] .. ncol_tiles - 1, work items sym_ptr[r] .. sym_ptr[r + 1)
    const int32_t* row_ctile;
    const int64_t* sym_ptr;
    int64_t nsym;
    int nrow_tiles, ncol_tiles;
};

// Device-resident mesh data (one allocation).
struct MeshDev {
    int device = 0;
    int64_t nv = 0, nt = 0;
    float* verts = nullptr;  // [nv*3]
    int4* tris = nullptr;    // [nt] (v0, v1, v2, 0)
    float* rec6 = nullptr;   // [nt * Rec<6>::kFloats]
    float* rec3 = nullptr;
    float* rec1 = nullptr;
    float* curls = nullptr;  // [nt*12]: rows n x grad(phi_a), a = 0..2 (9 used)
    float* geo = nullptr;    // [nt*16]: v0 v1 v2 (9), unit normal (3), |J| (1)
    int4* pairs = nullptr;   // [npairs]: (i, j, shared vertices, 0), sorted by (i, j)
    int4* pairs_by_kind = nullptr;  // the same pairs, coincident, edge, vertex (P1 launches)
    int64_t* pair_ptr = nullptr;  // [nt+1] pair range of each row triangle
    int64_t npairs = 0;
    int32_t* vt_ptr = nullptr;  // [nv+1] vertex -> triangle CSR (ascending)
    int32_t* vt_idx = nullptr;
    int32_t* vt_corner = nullptr;  // [3nt] in vt_idx order: 3 t + the vertex's corner in triangle t
    P1Tiles tiles{};
    // the P1 numbering of the tiles (vertices in order of first appearance
    // along the tiles' triangle order): p1_vnew [nv] caller id -> P1 id,
    // p1_vold [nv] P1 id -> caller id; tiles.row_verts / col_verts hold P1 ids
    int32_t* p1_vnew = nullptr;
    int32_t* p1_vold = nullptr;
    std::vector<int64_t> h_pair_ptr;  // [nt+1] host copy: pair range of each row triangle
    int64_t h_row_slots = 0;          // total vertex slots of the row tiles (tiles.row_vptr[nrow_tiles])
    int32_t max_neighbors = 0;
    double diameter = 0.0;  // bounding-box diagonal
    std::vector<int32_t> h_tris;  // [nt*3] host copy (for mesh_set_vertices)
    size_t geom_bytes = 0;  // the position-dependent arrays: the first geom_bytes of block
    void* block = nullptr;  // the single device allocation
};

// Wavenumbers of one launch over blockIdx.z: operator z of a batch on one
// mesh uses k[z] and is written at out + z * zstride (elements).  Passed by
// value (kernel arguments), at most kMaxBatchK per launch.
// With a geometry set (several placements of one mesh's triangles) launch z
// also reads the position-dependent arrays (records, curls, singular-pair
// geometry) gstride floats further on, and the evaluation points pstride
// floats further on; both are 0 for one geometry and one point set.
constexpr int kMaxBatchK = 64;
struct KBatch {
    float k[kMaxBatchK];
    int64_t zstride;
    int64_t gstride = 0;
    int64_t pstride = 0;
    int64_t vstride = 0;  // per-launch vertex vectors (the fused Burton-Miller right-hand sides)
};
inline KBatch kbatch1(float k)
{
    KBatch b{};
    b.k[0] = k;
    b.zstride = 0;
    return b;
}

hipError_t launch_fill(float2* out, int64_t rows, int64_t cols, int64_t ld, float2 val, hipStream_t s);
hipError_t launch_assemble(const MeshDev& m, int op, int basis, int quad, int clamp, int parts, int64_t row0,
                           int64_t row1, float k, float2* out, int64_t ld, hipStream_t s);
// nk operators of one kind on one mesh (operator z at out + z * stride,
// [n, ld] each), one launch per kernel for up to kMaxBatchK wavenumbers.
hipError_t launch_assemble_batch_k(const MeshDev& m, int op, int basis, int quad, int clamp, const float* ks, int nk,
                                   float2* out, int64_t ld, int64_t stride, hipStream_t s);
// One (mesh, k) item of a batched fused-CBIE assembly (device-resident table).
struct BatchItemDev {
    const float* rec;   // 6-point records of the mesh
    const int4* tris;
    int64_t nt;
    float k;
    int pad;
    const float2* psi;  // [nt]
    float2* lhs;        // [nt, ld]
    int64_t ld;
    float2* prow;       // row partials of the fused CBIE pass
    float2* pcol;       // column partials
    int64_t task0;      // first task of the item in the queue
};
hipError_t launch_batch_cbie_persistent(const BatchItemDev* items, int nitems, int64_t total, int rt,
                                        unsigned long long* counter, int clamp, hipStream_t s);
int64_t batch_cbie_tasks(int64_t nt, int rt);
// Scratch of the fused DP0 CBIE pass for row tile rt (16, 32 or 64 rows per
// task): row partials, column partials (64 / rt slots per 64-row block) and
// the singular V psi terms.
struct CbieWork {
    float2* prow;
    float2* pcol;
    float2* vpsi;
    int nsub;
};
int64_t cbie_work_bytes(const MeshDev& m, int rt);
CbieWork cbie_work_split(const MeshDev& m, int rt, void* work);
hipError_t launch_cbie_dp0(const MeshDev& m, float k, int quad, int clamp, int parts, const float2* psi, float2* lhs,
                           int64_t ld, float2* rhs, const CbieWork& w, hipStream_t s);
int64_t adjoint_work_bytes(const MeshDev& m);
hipError_t launch_adjoint_pairs(const MeshDev& m, float k, int clamp, const float2* lam, const float2* psi,
                                const float2* phi, double* gvert, double* gk, double* vhl, void* work, hipStream_t s);
hipError_t launch_adjoint_potential(const MeshDev& m, const float* pts, int64_t np, float k, int clamp,
                                    const float2* pbar, const float2* psi, const float2* phi, double* gvert, double* gk,
                                    hipStream_t s);
// P1 Burton-Miller adjoint: vertex gradients (FP64, accumulated into gvert
// [nv, 3]) of 2 Re lambda^H (dR g - dL phi) (regular and neighbour pairs, the
// mass terms) and of 2 Re pbar^H (dD_pot phi - dS_pot g) (plus the listener
// gradients into gpts [np, 3] when it is not null).
hipError_t launch_adjoint_bm_p1(const MeshDev& m, float k, int clamp, const float2* lam, const float2* g,
                                const float2* phi, double* gvert, hipStream_t s);
hipError_t launch_adjoint_potential_p1(const MeshDev& m, const float* pts, int64_t np, float k, int clamp,
                                       const float2* pbar, const float2* g, const float2* phi, double* gvert,
                                       double* gpts, hipStream_t s);
hipError_t launch_add_identity(const MeshDev& m, int basis, float2 alpha, float2* out, int64_t ld, hipStream_t s);
// neumann2dirichlet's P1 system: lhs = bm_lhs + M/2 [nv, ld], rhs = (bm_rhs -
// (i/2k) M) g [nv] without storing bm_rhs (contracted inside the LHS kernels)
hipError_t launch_bm_system_p1(const MeshDev& m, float k, int quad, int clamp, const float2* g, float2* lhs,
                               int64_t ld, float2* rhs, hipStream_t s);
// the same for nk wavenumbers on one mesh (gstride 0) or for the ns samples of
// a geometry set (mg its region view): system z at lhs + z * stride, g and rhs
// [nk, nv] contiguous
hipError_t launch_bm_systems_p1(const MeshDev& mg, int64_t gstride, const float* ks, int nk, int quad, int clamp,
                                const float2* g, float2* lhs, int64_t ld, int64_t stride, float2* rhs,
                                hipStream_t s);
hipError_t launch_identity(const MeshDev& m, int basis, float2* out, int64_t ld, hipStream_t s);
// records msg as the calling thread's bem_last_error() and returns code
int report_failure(int code, const std::string& msg);
hipError_t launch_tri2vert(const MeshDev& m, const float2* tv, float2* out, hipStream_t s);
// Geometry sets: mg is the view of sample 0 (the mesh's connectivity with
// position-dependent pointers into the set), sample z's arrays lie gstride
// floats further on.  One wavenumber per sample (ks host [ns]); operator z
// at out + z * stride.  Potentials: points device [ns, np, 3].  Mass terms:
// alphas host [ns] complex (re, im).
hipError_t launch_assemble_samples(const MeshDev& mg, int64_t gstride, const float* ks, int ns, int op, int basis,
                                   int quad, int clamp, float2* out, int64_t ld, int64_t stride, hipStream_t s);
hipError_t launch_add_identity_samples(const MeshDev& mg, int64_t gstride, int basis, const float* alphas, int ns,
                                       float2* out, int64_t ld, int64_t stride, hipStream_t s);
hipError_t launch_potential_samples(const MeshDev& mg, int64_t gstride, const float* pts, int64_t np, const float* ks,
                                    int ns, int op, int basis, int clamp, float2* out, int64_t ld, int64_t stride,
                                    hipStream_t s);
// work: int32 [2 nv + 3 nt + 2] (counts, list pointers, lists, error flag)
hipError_t launch_tri2vert_raw(const int32_t* tris, int64_t nt, int64_t nv, const float2* tv, float2* out,
                               int32_t* work, hipStream_t s);
hipError_t launch_potential(const MeshDev& m, const float* pts, int64_t np, float k, int op, int basis, int clamp,
                            float2* out, int64_t ld, hipStream_t s);
hipError_t launch_potential_batch_k(const MeshDev& m, const float* pts, int64_t np, const float* ks, int nk, int op,
                                    int basis, int clamp, float2* out, int64_t ld, int64_t stride, hipStream_t s);

// Builds all device data from host copies of the mesh; returns 0 or a
// bem_status code with a message in `err`.
int mesh_build(const float* h_verts, int64_t nv, const int32_t* h_tris, int64_t nt, MeshDev& m, std::string& err);
// New vertex positions for a built mesh (same triangles): recomputes the
// position-dependent arrays on the host and uploads them in one copy on s,
// then waits for it.  Connectivity data and the tiling are kept.
int mesh_set_vertices(MeshDev& m, const float* h_verts, hipStream_t s, std::string& err);
// The position-dependent region of m (the first m.geom_bytes of its block:
// vertices, quadrature records, curls, singular-pair geometry) for new
// positions h_verts [nv, 3] of the same triangles, written to region (host).
int mesh_geometry_region(const MeshDev& m, const float* h_verts, char* region, std::string& err);
void mesh_free(MeshDev& m);

}  // namespace bem
