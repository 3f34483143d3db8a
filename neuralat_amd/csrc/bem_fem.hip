// Linear-tetrahedron FEM matrices for modal analysis (the step upstream of
// the BEM path that turns a solid into vibration modes and their Neumann
// data), replacing the reference's per-tet COO triplet kernels
// (src/modalobj/cuda/computeMatrix.cu:6-142) that write 144 uncoalesced
// entries per tetrahedron for torch to sum.
//
// Layout: 3 degrees of freedom per vertex (dof 3 v + c); both matrices are
// stored as 3 x 3 blocks over the vertex adjacency graph (block CSR, rows and
// columns sorted), built once per mesh on the host.  Assembly is a gather:
// one thread per block sums the contributions of the tetrahedra sharing its
// two vertices in a fixed order (host-built lists), so the values are
// deterministic and no atomics are used.  A per-tet record (the four shape
// function gradients and the volume) is computed once; stiffness blocks are
// formed from it on the fly:
//   K_(ia),(jb) = V (lambda g_i[a] g_j[b] + mu (delta_ab g_i.g_j + g_j[a] g_i[b]))
// (the isotropic form of B^T D B V of computeMatrix.cu:58-127), and mass
// blocks are rho V / 20 (1 + delta_ij) I_3 (computeMatrix.cu:28-40).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/bem_hip.h"

namespace bem {

constexpr int kFemRec = 16;  // floats per tet record: 4 gradients (12), volume, pad

struct FemDev {
    int64_t nv = 0, nt = 0, nnzb = 0;
    float* rec = nullptr;        // [nt * kFemRec]
    int32_t* row_ptr = nullptr;  // [nv + 1] block rows
    int32_t* col = nullptr;      // [nnzb] block columns (sorted per row)
    int32_t* brow = nullptr;     // [nnzb] block row of each block
    int32_t* cptr = nullptr;     // [nnzb + 1] contribution ranges
    int32_t* contrib = nullptr;  // [16 nt]: tet * 16 + 4 * local row vertex + local column vertex
    void* block = nullptr;
};

// Shape-function gradients and volume of each tetrahedron, in FP64 then
// rounded: x = p0 + J xi, grad N_i = row i of J^-1 (i = 1..3), grad N_0 =
// -(grad N_1 + grad N_2 + grad N_3), V = |det J| / 6.
__global__ __launch_bounds__(256) void fem_records_kernel(const float* __restrict__ verts,
                                                          const int32_t* __restrict__ tets, int64_t nt,
                                                          float* __restrict__ rec)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nt) return;
    double p[4][3];
    for (int i = 0; i < 4; ++i) {
        const int64_t v = tets[4 * t + i];
        for (int c = 0; c < 3; ++c) p[i][c] = verts[3 * v + c];
    }
    double J[3][3];  // columns p_i - p_0
    for (int r = 0; r < 3; ++r)
        for (int i = 0; i < 3; ++i) J[r][i] = p[i + 1][r] - p[0][r];
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const double id = 1.0 / det;
    // inverse = adjugate / det; row i of the inverse is grad N_{i+1}
    double inv[3][3];
    inv[0][0] = c00 * id;
    inv[1][0] = c01 * id;
    inv[2][0] = c02 * id;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    float* R = rec + t * kFemRec;
    for (int c = 0; c < 3; ++c) {
        R[c] = (float)(-(inv[0][c] + inv[1][c] + inv[2][c]));
        for (int i = 0; i < 3; ++i) R[3 * (i + 1) + c] = (float)inv[i][c];
    }
    R[12] = (float)(fabs(det) / 6.0);
    R[13] = R[14] = R[15] = 0.f;
}

// values[e] (3 x 3, row-major) = sum over the block's contributions (in
// FP64, stored as T); which = 0 mass (c0 = density), 1 stiffness (c0 =
// lambda, c1 = mu).  One thread per block.
template <typename T>
__global__ __launch_bounds__(256) void fem_assemble_kernel(const float* __restrict__ rec,
                                                           const int32_t* __restrict__ cptr,
                                                           const int32_t* __restrict__ contrib, int64_t nnzb,
                                                           int which, double c0, double c1, T* __restrict__ values)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= nnzb) return;
    double b[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int32_t q = cptr[e]; q < cptr[e + 1]; ++q) {
        const int32_t code = contrib[q];
        const int64_t t = code >> 4;
        const int li = (code >> 2) & 3, lj = code & 3;
        const float* R = rec + t * kFemRec;
        const double V = R[12];
        if (which == 0) {
            const double m = c0 * V / 20.0 * (li == lj ? 2.0 : 1.0);
            b[0] += m;
            b[4] += m;
            b[8] += m;
        } else {
            const double gi[3] = {R[3 * li], R[3 * li + 1], R[3 * li + 2]};
            const double gj[3] = {R[3 * lj], R[3 * lj + 1], R[3 * lj + 2]};
            const double dot = gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2];
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c)
                    b[3 * a + c] += V * (c0 * gi[a] * gj[c] + c1 * ((a == c ? dot : 0.0) + gj[a] * gi[c]));
        }
    }
    for (int q = 0; q < 9; ++q) values[e * 9 + q] = (T)b[q];
}

// Y = A X for block-CSR A (3 x 3 blocks) and X, Y row-major [3 nv][m] with
// leading dimensions ldx, ldy, all of type T (float or double).  One
// wavefront per block row; lane j owns column j (columns beyond 64 in further
// passes); the block values are wave-uniform loads, the X rows of a block
// column are contiguous across lanes.
template <typename T>
__global__ __launch_bounds__(256) void fem_spmm_kernel(const int32_t* __restrict__ row_ptr,
                                                       const int32_t* __restrict__ col,
                                                       const T* __restrict__ values, int64_t nv,
                                                       const T* __restrict__ X, int64_t ldx, int m,
                                                       T* __restrict__ Y, int64_t ldy)
{
    const int64_t u = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (u >= nv) return;
    const int lane = threadIdx.x & 63;
    const int32_t e0 = row_ptr[u], e1 = row_ptr[u + 1];
    for (int j = lane; j < m; j += 64) {
        T y0 = 0, y1 = 0, y2 = 0;
        for (int32_t e = e0; e < e1; ++e) {
            const T* A = values + (int64_t)e * 9;
            const int64_t c = col[e];
            const T x0 = X[(3 * c) * ldx + j], x1 = X[(3 * c + 1) * ldx + j], x2 = X[(3 * c + 2) * ldx + j];
            y0 = fma(A[0], x0, fma(A[1], x1, fma(A[2], x2, y0)));
            y1 = fma(A[3], x0, fma(A[4], x1, fma(A[5], x2, y1)));
            y2 = fma(A[6], x0, fma(A[7], x1, fma(A[8], x2, y2)));
        }
        Y[(3 * u) * ldy + j] = y0;
        Y[(3 * u + 1) * ldy + j] = y1;
        Y[(3 * u + 2) * ldy + j] = y2;
    }
}

}  // namespace bem

struct bem_fem {
    bem::FemDev dev;
};

using namespace bem;

namespace {
thread_local std::string g_fem_err;

int fem_fail(int rc, const std::string& msg)
{
    g_fem_err = msg;
    return rc;
}

size_t al256(size_t b) { return (b + 255) / 256 * 256; }
}  // namespace

#define FEM_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) return fem_fail(BEM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

extern "C" {

const char* bem_fem_last_error(void) { return g_fem_err.c_str(); }

int bem_fem_create(const float* verts, int64_t nv, const int32_t* tets, int64_t ntet, bem_fem** out)
{
    if (!verts || !tets || !out) return fem_fail(BEM_ERR_BAD_ARG, "null pointer");
    *out = nullptr;
    if (nv < 4 || ntet < 1) return fem_fail(BEM_ERR_BAD_ARG, "need at least 4 vertices and 1 tetrahedron");
    if (ntet > ((int64_t)1 << 27) || nv > ((int64_t)1 << 30)) return fem_fail(BEM_ERR_BAD_ARG, "mesh too large");
    for (int64_t t = 0; t < ntet; ++t) {
        for (int i = 0; i < 4; ++i) {
            const int32_t v = tets[4 * t + i];
            if (v < 0 || v >= nv) return fem_fail(BEM_ERR_BAD_MESH, "tetrahedron " + std::to_string(t) + " has a vertex index out of range");
            for (int j = 0; j < i; ++j)
                if (tets[4 * t + j] == v) return fem_fail(BEM_ERR_BAD_MESH, "tetrahedron " + std::to_string(t) + " repeats a vertex");
        }
        double e[3][3];
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 3; ++c)
                e[i][c] = (double)verts[3 * (int64_t)tets[4 * t + i + 1] + c] - verts[3 * (int64_t)tets[4 * t] + c];
        const double det = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
                           e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
                           e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
        if (!(std::fabs(det) > 0.0) || !std::isfinite(det))
            return fem_fail(BEM_ERR_BAD_MESH, "tetrahedron " + std::to_string(t) + " has zero volume");
    }
    // block pattern and contribution lists: all 16 (row vertex, column vertex)
    // pairs of every tetrahedron, sorted by (row, column, tet)
    std::vector<int64_t> key((size_t)ntet * 16);
    for (int64_t t = 0; t < ntet; ++t)
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) key[(size_t)(16 * t + 4 * i + j)] = 16 * t + 4 * i + j;
    auto rv = [&](int64_t c) { return (int64_t)tets[4 * (c >> 4) + ((c >> 2) & 3)]; };
    auto cv = [&](int64_t c) { return (int64_t)tets[4 * (c >> 4) + (c & 3)]; };
    std::sort(key.begin(), key.end(), [&](int64_t a, int64_t b) {
        const int64_t ra = rv(a), rb = rv(b);
        if (ra != rb) return ra < rb;
        const int64_t ca = cv(a), cb = cv(b);
        if (ca != cb) return ca < cb;
        return a < b;
    });
    std::vector<int32_t> row_ptr((size_t)nv + 1, 0), col, brow, cptr, contrib(key.size());
    for (size_t q = 0; q < key.size(); ++q) {
        contrib[q] = (int32_t)key[q];
        const int64_t r = rv(key[q]), c = cv(key[q]);
        if (q == 0 || r != rv(key[q - 1]) || c != cv(key[q - 1])) {
            col.push_back((int32_t)c);
            brow.push_back((int32_t)r);
            cptr.push_back((int32_t)q);
            ++row_ptr[(size_t)r + 1];
        }
    }
    cptr.push_back((int32_t)key.size());
    for (int64_t v = 0; v < nv; ++v) row_ptr[(size_t)v + 1] += row_ptr[(size_t)v];
    const int64_t nnzb = (int64_t)col.size();

    auto* h = new bem_fem();
    FemDev& d = h->dev;
    d.nv = nv;
    d.nt = ntet;
    d.nnzb = nnzb;
    const size_t o_verts = 0, o_tets = o_verts + al256(12 * (size_t)nv);
    const size_t o_rec = o_tets + al256(16 * (size_t)ntet);
    const size_t o_rp = o_rec + al256(4 * kFemRec * (size_t)ntet);
    const size_t o_col = o_rp + al256(4 * ((size_t)nv + 1));
    const size_t o_brow = o_col + al256(4 * (size_t)nnzb);
    const size_t o_cptr = o_brow + al256(4 * (size_t)nnzb);
    const size_t o_con = o_cptr + al256(4 * ((size_t)nnzb + 1));
    const size_t total = o_con + al256(4 * contrib.size());
    if (hipMalloc(&d.block, total) != hipSuccess) {
        (void)hipGetLastError();
        delete h;
        return fem_fail(BEM_ERR_HIP, "device allocation of " + std::to_string(total) + " bytes failed");
    }
    char* B = (char*)d.block;
    float* dverts = (float*)(B + o_verts);
    int32_t* dtets = (int32_t*)(B + o_tets);
    d.rec = (float*)(B + o_rec);
    d.row_ptr = (int32_t*)(B + o_rp);
    d.col = (int32_t*)(B + o_col);
    d.brow = (int32_t*)(B + o_brow);
    d.cptr = (int32_t*)(B + o_cptr);
    d.contrib = (int32_t*)(B + o_con);
    bool ok = hipMemcpy(dverts, verts, 12 * (size_t)nv, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(dtets, tets, 16 * (size_t)ntet, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d.row_ptr, row_ptr.data(), 4 * row_ptr.size(), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d.col, col.data(), 4 * col.size(), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d.brow, brow.data(), 4 * brow.size(), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d.cptr, cptr.data(), 4 * cptr.size(), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d.contrib, contrib.data(), 4 * contrib.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(fem_records_kernel, dim3((unsigned)((ntet + 255) / 256)), dim3(256), 0, nullptr, dverts,
                           dtets, ntet, d.rec);
        ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess;
    }
    if (!ok) {
        (void)hipFree(d.block);
        delete h;
        return fem_fail(BEM_ERR_HIP, "upload of the FEM mesh failed");
    }
    *out = h;
    return BEM_OK;
}

void bem_fem_free(bem_fem* fem)
{
    if (!fem) return;
    if (fem->dev.block) (void)hipFree(fem->dev.block);
    delete fem;
}

int64_t bem_fem_num_blocks(const bem_fem* fem) { return fem ? fem->dev.nnzb : 0; }

int bem_fem_pattern(const bem_fem* fem, void* row_ptr, void* block_rows, void* cols, void* stream)
{
    if (!fem || !row_ptr || !block_rows || !cols) return fem_fail(BEM_ERR_BAD_ARG, "null pointer");
    const FemDev& d = fem->dev;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    FEM_TRY(hipMemcpyAsync(row_ptr, d.row_ptr, 4 * ((size_t)d.nv + 1), hipMemcpyDeviceToDevice, s));
    FEM_TRY(hipMemcpyAsync(block_rows, d.brow, 4 * (size_t)d.nnzb, hipMemcpyDeviceToDevice, s));
    FEM_TRY(hipMemcpyAsync(cols, d.col, 4 * (size_t)d.nnzb, hipMemcpyDeviceToDevice, s));
    return BEM_OK;
}

int bem_fem_assemble(const bem_fem* fem, int which, float density, float youngs, float poisson, int fp64,
                     void* values, void* stream)
{
    if (!fem || !values) return fem_fail(BEM_ERR_BAD_ARG, "null pointer");
    const FemDev& d = fem->dev;
    double c0 = 0.0, c1 = 0.0;
    if (which == 0) {
        if (!(density > 0.f) || !std::isfinite(density)) return fem_fail(BEM_ERR_BAD_ARG, "density must be > 0");
        c0 = density;
    } else if (which == 1) {
        if (!(youngs > 0.f) || !(poisson > -1.f && poisson < 0.5f))
            return fem_fail(BEM_ERR_BAD_ARG, "need Young's modulus > 0 and -1 < Poisson's ratio < 0.5");
        const double E = youngs, nu = poisson;
        c0 = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));  // lambda
        c1 = E / (2.0 * (1.0 + nu));                     // mu
    } else {
        return fem_fail(BEM_ERR_BAD_ARG, "which must be 0 (mass) or 1 (stiffness)");
    }
    const dim3 g((unsigned)((d.nnzb + 255) / 256));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (fp64)
        hipLaunchKernelGGL(fem_assemble_kernel<double>, g, dim3(256), 0, s, d.rec, d.cptr, d.contrib, d.nnzb, which,
                           c0, c1, (double*)values);
    else
        hipLaunchKernelGGL(fem_assemble_kernel<float>, g, dim3(256), 0, s, d.rec, d.cptr, d.contrib, d.nnzb, which,
                           c0, c1, (float*)values);
    FEM_TRY(hipGetLastError());
    return BEM_OK;
}

int bem_fem_spmm(const bem_fem* fem, int fp64, const void* values, const void* X, int64_t ldx, int m, void* Y,
                 int64_t ldy, void* stream)
{
    if (!fem || !values || !X || !Y) return fem_fail(BEM_ERR_BAD_ARG, "null pointer");
    if (m <= 0 || ldx < m || ldy < m) return fem_fail(BEM_ERR_BAD_ARG, "bad shape");
    const FemDev& d = fem->dev;
    const dim3 g((unsigned)((d.nv + 3) / 4));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (fp64)
        hipLaunchKernelGGL(fem_spmm_kernel<double>, g, dim3(256), 0, s, d.row_ptr, d.col, (const double*)values, d.nv,
                           (const double*)X, ldx, m, (double*)Y, ldy);
    else
        hipLaunchKernelGGL(fem_spmm_kernel<float>, g, dim3(256), 0, s, d.row_ptr, d.col, (const float*)values, d.nv,
                           (const float*)X, ldx, m, (float*)Y, ldy);
    FEM_TRY(hipGetLastError());
    return BEM_OK;
}

}  // extern "C"
