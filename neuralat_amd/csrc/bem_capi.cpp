// extern "C" entry points of libbem_hip.so (declared in include/bem_hip.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bem_hip.h"
#include "bem_device.hpp"
#include "bem_internal.hpp"
#include "bem_linalg.hpp"
#include "bem_lowp.hpp"
#include "bem_msolve.hpp"

struct bem_mesh {
    bem::MeshDev dev;
};

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* where)
{
    return fail(BEM_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                        \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return hip_fail(_e, #expr);    \
    } while (0)

// Runs on the mesh's device, restoring the caller's device afterwards.
struct DeviceGuard {
    int prev = -1;
    bool ok = true;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) {
            ok = false;
            return;
        }
        if (prev != dev && hipSetDevice(dev) != hipSuccess) ok = false;
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }
constexpr int64_t kDp0RowTileHost = bem::kDp0RowTile;

}  // namespace

namespace bem {
int report_failure(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace bem

namespace {

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

struct GmresLayout {
    size_t V, w, r, H, h, h2, cs, sn, g, y, res, sc, tw, hp, total;
};
GmresLayout gmres_layout(int64_t n, int m)
{
    GmresLayout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o += align256(bytes);
        return at;
    };
    L.V = take(sizeof(float2) * (size_t)n * (size_t)(m + 1));
    L.w = take(sizeof(float2) * (size_t)n);
    L.r = take(sizeof(float2) * (size_t)n);
    L.H = take(sizeof(double2) * (size_t)(m + 1) * (size_t)m);
    L.h = take(sizeof(double2) * (size_t)(m + 1));
    L.h2 = take(sizeof(double2) * (size_t)(m + 1));
    L.cs = take(sizeof(double2) * (size_t)m);
    L.sn = take(sizeof(double2) * (size_t)m);
    L.g = take(sizeof(double2) * (size_t)(m + 1));
    L.y = take(sizeof(double2) * (size_t)m);
    L.res = take(sizeof(double) * (size_t)m);
    L.sc = take(sizeof(double) * 8);
    L.tw = take(sizeof(float2) * (size_t)bem::gemv_work_elems(n, n, 1));
    L.hp = take(sizeof(double2) * (size_t)bem::multidot_part_elems(m + 1));
    L.total = o;
    return L;
}

int read_double(const double* d, double* h, hipStream_t s)
{
    HIP_TRY(hipMemcpyAsync(h, d, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BEM_OK;
}

}  // namespace

extern "C" {

const char* bem_version(void) { return "neuralat_amd-bem 0.1 (gfx950)"; }

const char* bem_last_error(void) { return g_last_error.c_str(); }

int bem_mesh_create(const float* verts, int64_t nv, const int32_t* tris, int64_t nt, int on_device, void* stream,
                    bem_mesh** out)
{
    if (!verts || !tris || !out) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (nv <= 0 || nt <= 0) return fail(BEM_ERR_BAD_ARG, "empty mesh");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(BEM_ERR_NO_DEVICE, "no HIP device");
    std::vector<float> hv;
    std::vector<int32_t> ht;
    const float* pv = verts;
    const int32_t* pt = tris;
    if (on_device) {
        hv.resize((size_t)nv * 3);
        ht.resize((size_t)nt * 3);
        HIP_TRY(hipMemcpyAsync(hv.data(), verts, hv.size() * 4, hipMemcpyDeviceToHost, S(stream)));
        HIP_TRY(hipMemcpyAsync(ht.data(), tris, ht.size() * 4, hipMemcpyDeviceToHost, S(stream)));
        HIP_TRY(hipStreamSynchronize(S(stream)));
        pv = hv.data();
        pt = ht.data();
    }
    bem_mesh* m = new (std::nothrow) bem_mesh();
    if (!m) return fail(BEM_ERR_OOM, "host allocation failed");
    std::string err;
    int rc = bem::mesh_build(pv, nv, pt, nt, m->dev, err);
    if (rc != 0) {
        delete m;
        return fail(rc, err);
    }
    *out = m;
    return BEM_OK;
}

int bem_mesh_set_vertices(bem_mesh* mesh, const float* verts, int on_device, void* stream)
{
    if (!mesh || !verts) return fail(BEM_ERR_BAD_ARG, "null pointer");
    auto& m = mesh->dev;
    DeviceGuard g(m.device);
    if (!g.ok) return fail(BEM_ERR_HIP, "cannot select the mesh's device");
    std::vector<float> hv;
    const float* pv = verts;
    if (on_device) {
        hv.resize((size_t)m.nv * 3);
        HIP_TRY(hipMemcpyAsync(hv.data(), verts, hv.size() * 4, hipMemcpyDeviceToHost, S(stream)));
        HIP_TRY(hipStreamSynchronize(S(stream)));
        pv = hv.data();
    }
    std::string err;
    const int rc = bem::mesh_set_vertices(m, pv, S(stream), err);
    if (rc != 0) return fail(rc == 3 ? BEM_ERR_HIP : (rc == 1 ? BEM_ERR_BAD_ARG : BEM_ERR_BAD_MESH), err);
    return BEM_OK;
}

int bem_mesh_destroy(bem_mesh* mesh)
{
    if (!mesh) return BEM_OK;
    DeviceGuard g(mesh->dev.device);
    bem::mesh_free(mesh->dev);
    delete mesh;
    return BEM_OK;
}

int bem_mesh_info(const bem_mesh* mesh, int64_t* nv, int64_t* nt, int64_t* npairs, int32_t* maxnb)
{
    if (!mesh) return fail(BEM_ERR_BAD_ARG, "null mesh");
    if (nv) *nv = mesh->dev.nv;
    if (nt) *nt = mesh->dev.nt;
    if (npairs) *npairs = mesh->dev.npairs;
    if (maxnb) *maxnb = mesh->dev.max_neighbors;
    return BEM_OK;
}

int bem_mesh_normals_curls(const bem_mesh* mesh, float* normals, float* curls, void* stream)
{
    if (!mesh) return fail(BEM_ERR_BAD_ARG, "null mesh");
    DeviceGuard g(mesh->dev.device);
    const auto& m = mesh->dev;
    if (normals)
        HIP_TRY(hipMemcpy2DAsync(normals, 12, m.geo + 9, 64, 12, (size_t)m.nt, hipMemcpyDeviceToDevice, S(stream)));
    if (curls)
        HIP_TRY(hipMemcpy2DAsync(curls, 36, m.curls, 48, 36, (size_t)m.nt, hipMemcpyDeviceToDevice, S(stream)));
    return BEM_OK;
}

static int check_op(int op, int basis, int quad)
{
    if (op < 0 || op > 5) return fail(BEM_ERR_BAD_ARG, "unknown operator " + std::to_string(op));
    if (basis != BEM_BASIS_P1 && basis != BEM_BASIS_DP0) return fail(BEM_ERR_BAD_ARG, "unknown basis");
    if (quad < 0 || quad > 2) return fail(BEM_ERR_BAD_ARG, "unknown quadrature");
    if (basis == BEM_BASIS_DP0 && (op == BEM_OP_HYPERSINGULAR || op == BEM_OP_BM_LHS))
        return fail(BEM_ERR_BAD_ARG, "the hypersingular and Burton-Miller LHS operators need the P1 basis");
    return BEM_OK;
}

int bem_assemble(const bem_mesh* mesh, float k, int op, int basis, int quad, int clamp, void* out, int64_t ld,
                 void* stream)
{
    return bem_assemble_ex(mesh, k, op, basis, quad, clamp, 3, out, ld, stream);
}

int bem_assemble_ex(const bem_mesh* mesh, float k, int op, int basis, int quad, int clamp, int parts, void* out,
                    int64_t ld, void* stream)
{
    if (parts < 1 || parts > 15 || (parts & 3) == 0)
        return fail(BEM_ERR_BAD_ARG, "parts must be 1, 2 or 3, optionally with bits 4 and 8 set");
    if (!mesh || !out) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (int rc = check_op(op, basis, quad)) return rc;
    const auto& m = mesh->dev;
    const int64_t n = basis == BEM_BASIS_DP0 ? m.nt : m.nv;
    if (ld < n) return fail(BEM_ERR_BAD_ARG, "ld smaller than the matrix width");
    if (!std::isfinite(k)) return fail(BEM_ERR_BAD_ARG, "non-finite wavenumber");
    if ((op == BEM_OP_BM_LHS || op == BEM_OP_BM_RHS) && k == 0.f)
        return fail(BEM_ERR_BAD_ARG, "Burton-Miller coupling i/k needs k != 0");
    DeviceGuard g(m.device);
    if (!g.ok) return fail(BEM_ERR_HIP, "cannot select the mesh's device");
    hipError_t e = bem::launch_assemble(m, op, basis, quad, clamp, parts, 0, m.nt, k, (float2*)out, ld, S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_assemble launch");
    return BEM_OK;
}

int bem_burton_miller_system(const bem_mesh* mesh, float k, int quad, int clamp, const void* g, void* lhs, int64_t ld,
                             void* rhs, void* stream)
{
    if (!mesh || !g || !lhs || !rhs) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (quad < 0 || quad > 2) return fail(BEM_ERR_BAD_ARG, "unknown quadrature");
    const auto& m = mesh->dev;
    if (ld < m.nv) return fail(BEM_ERR_BAD_ARG, "ld smaller than the matrix width");
    if (!std::isfinite(k) || k == 0.f) return fail(BEM_ERR_BAD_ARG, "Burton-Miller coupling i/k needs a finite k != 0");
    DeviceGuard dg(m.device);
    if (!dg.ok) return fail(BEM_ERR_HIP, "cannot select the mesh's device");
    hipError_t e = bem::launch_bm_system_p1(m, k, quad, clamp, (const float2*)g, (float2*)lhs, ld, (float2*)rhs,
                                            S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_burton_miller_system launch");
    return BEM_OK;
}

static int check_bm_ks(const float* ks, int nk)
{
    if (!ks || nk < 1) return fail(BEM_ERR_BAD_ARG, "need at least one wavenumber");
    for (int z = 0; z < nk; ++z)
        if (!std::isfinite(ks[z]) || ks[z] == 0.f)
            return fail(BEM_ERR_BAD_ARG, "Burton-Miller coupling i/k needs finite wavenumbers != 0");
    return BEM_OK;
}

int bem_burton_miller_systems(const bem_mesh* mesh, const float* ks, int nk, int quad, int clamp, const void* g,
                              void* lhs, int64_t ld, int64_t stride, void* rhs, void* stream)
{
    if (!mesh || !g || !lhs || !rhs) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (quad < 0 || quad > 2) return fail(BEM_ERR_BAD_ARG, "unknown quadrature");
    if (int rc = check_bm_ks(ks, nk)) return rc;
    const auto& m = mesh->dev;
    if (ld < m.nv) return fail(BEM_ERR_BAD_ARG, "ld smaller than the matrix width");
    if (nk > 1 && stride < m.nv * ld) return fail(BEM_ERR_BAD_ARG, "stride smaller than one matrix (nv * ld)");
    DeviceGuard dg(m.device);
    if (!dg.ok) return fail(BEM_ERR_HIP, "cannot select the mesh's device");
    hipError_t e = bem::launch_bm_systems_p1(m, 0, ks, nk, quad, clamp, (const float2*)g, (float2*)lhs, ld, stride,
                                             (float2*)rhs, S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_burton_miller_systems launch");
    return BEM_OK;
}

int bem_assemble_dk(const bem_mesh* mesh, float k, int op, int quad, int clamp, void* out, int64_t ld, void* stream)
{
    if (!mesh || !out) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (op == BEM_OP_HYPERSINGULAR) return fail(BEM_ERR_BAD_ARG, "no wavenumber derivative of the hypersingular operator");
    if (int rc = check_op(op, BEM_BASIS_P1, quad)) return rc;
    const auto& m = mesh->dev;
    if (ld < m.nv) return fail(BEM_ERR_BAD_ARG, "ld smaller than the matrix width");
    if (!std::isfinite(k)) return fail(BEM_ERR_BAD_ARG, "non-finite wavenumber");
    if ((op == BEM_OP_BM_LHS || op == BEM_OP_BM_RHS) && k == 0.f)
        return fail(BEM_ERR_BAD_ARG, "Burton-Miller coupling i/k needs k != 0");
    DeviceGuard g(m.device);
    if (!g.ok) return fail(BEM_ERR_HIP, "cannot select the mesh's device");
    hipError_t e = bem::launch_assemble(m, op | bem::kDk, BEM_BASIS_P1, quad, clamp, 3, 0, m.nt, k, (float2*)out, ld,
                                        S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_assemble_dk launch");
    return BEM_OK;
}

static int check_ks(const float* ks, int nk, int op)
{
    if (!ks || nk < 1) return fail(BEM_ERR_BAD_ARG, "need at least one wavenumber");
    for (int z = 0; z < nk; ++z) {
        if (!std::isfinite(ks[z])) return fail(BEM_ERR_BAD_ARG, "non-finite wavenumber");
        if ((op == BEM_OP_BM_LHS || op == BEM_OP_BM_RHS) && ks[z] == 0.f)
            return fail(BEM_ERR_BAD_ARG, "Burton-Miller coupling i/k needs k != 0");
    }
    return BEM_OK;
}

int bem_assemble_batch_k(const bem_mesh* mesh, const float* ks, int nk, int op, int basis, int quad, int clamp,
                         void* out, int64_t ld, int64_t stride, void* stream)
{
    if (!mesh || !out) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (int rc = check_op(op, basis, quad)) return rc;
    if (int rc = check_ks(ks, nk, op)) return rc;
    const auto& m = mesh->dev;
    const int64_t n = basis == BEM_BASIS_DP0 ? m.nt : m.nv;
    if (ld < n) return fail(BEM_ERR_BAD_ARG, "ld smaller than the matrix width");
    if (nk > 1 && stride < n * ld) return fail(BEM_ERR_BAD_ARG, "stride smaller than one matrix (n * ld)");
    DeviceGuard g(m.device);
    if (!g.ok) return fail(BEM_ERR_HIP, "cannot select the mesh's device");
    hipError_t e = bem::launch_assemble_batch_k(m, op, basis, quad, clamp, ks, nk, (float2*)out, ld, stride,
                                                S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_assemble_batch_k launch");
    return BEM_OK;
}

// ---- geometry sets: several placements of one mesh's triangles ----------------

struct bem_geomset {
    const bem_mesh* mesh = nullptr;
    bem::MeshDev view;  // the mesh's connectivity, position-dependent pointers at sample 0 of block
    int64_t gstride = 0;  // floats from one sample's region to the next
    int ns = 0;
    void* block = nullptr;
};

// Pinned host staging for geometry-set uploads (a DMA copy instead of a
// staged pageable one: 9.6 MB of 48 two-sphere samples took 0.5-1 ms
// pageable).  One process-wide buffer, grown on demand up to
// kPinnedStageMax and kept until the process ends; larger sets stage in
// pageable memory.  g_stage_mu is held from the first write to the end of
// the copy.
static constexpr size_t kPinnedStageMax = (size_t)64 << 20;
static std::mutex g_stage_mu;
static char* g_stage = nullptr;
static size_t g_stage_cap = 0;

int bem_geomset_create(const bem_mesh* mesh, const float* verts, int ns, void* stream, bem_geomset** out)
{
    if (!mesh || !verts || !out) return fail(BEM_ERR_BAD_ARG, "null pointer");
    *out = nullptr;
    if (ns < 1) return fail(BEM_ERR_BAD_ARG, "need at least one sample");
    const auto& m = mesh->dev;
    const size_t gb = m.geom_bytes;
    DeviceGuard g(m.device);
    if (!g.ok) return fail(BEM_ERR_HIP, "cannot select the mesh's device");
    const size_t bytes = gb * (size_t)ns;
    std::unique_lock<std::mutex> lk(g_stage_mu, std::defer_lock);
    char* host = nullptr;
    std::vector<char> pageable;
    if (bytes <= kPinnedStageMax) {
        lk.lock();
        if (g_stage_cap < bytes) {
            if (g_stage) (void)hipHostFree(g_stage);
            g_stage = nullptr;
            g_stage_cap = 0;
            const size_t want = std::min(kPinnedStageMax, (bytes + ((size_t)1 << 20) - 1) >> 20 << 20);
            if (hipHostMalloc((void**)&g_stage, want, hipHostMallocDefault) == hipSuccess) {
                g_stage_cap = want;
            } else {
                (void)hipGetLastError();
                g_stage = nullptr;
            }
        }
        if (g_stage) host = g_stage;
        else lk.unlock();
    }
    if (!host) {
        pageable.resize(bytes);
        host = pageable.data();
    }
    // Each sample's geometry is independent host work (quadrature records,
    // curls): up to 16 threads, each on an interleaved share of the samples.
    std::vector<int> rcs((size_t)ns, 0);
    std::vector<std::string> errs((size_t)ns);
    auto work = [&](int t0, int step) {
        for (int z = t0; z < ns; z += step)
            rcs[z] = bem::mesh_geometry_region(m, verts + (size_t)z * 3 * (size_t)m.nv, host + z * gb, errs[z]);
    };
    const int nth = (int)std::min<int64_t>({(int64_t)ns, 16, (int64_t)std::max(1u, std::thread::hardware_concurrency()),
                                            std::max<int64_t>(1, (int64_t)ns * m.nt / 4096)});
    if (nth <= 1) {
        work(0, 1);
    } else {
        // a thread that cannot be started (std::system_error) leaves its
        // share to this one: no exception crosses the C ABI
        std::vector<std::thread> pool;
        int started = 1;
        try {
            pool.reserve(nth - 1);
            for (; started < nth; ++started) pool.emplace_back(work, started, nth);
        } catch (...) {
        }
        work(0, nth);
        for (int t = started; t < nth; ++t) work(t, nth);
        for (auto& th : pool) th.join();
    }
    for (int z = 0; z < ns; ++z)
        if (rcs[z])
            return fail(rcs[z] == 2 ? BEM_ERR_BAD_MESH : BEM_ERR_BAD_ARG, "sample " + std::to_string(z) + ": " + errs[z]);
    auto* gs = new bem_geomset();
    if (hipMalloc(&gs->block, bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete gs;
        return fail(BEM_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    hipError_t e = hipMemcpyAsync(gs->block, host, bytes, hipMemcpyHostToDevice, S(stream));
    if (e == hipSuccess) e = hipStreamSynchronize(S(stream));
    if (e != hipSuccess) {
        (void)hipFree(gs->block);
        delete gs;
        return hip_fail(e, "bem_geomset_create copy");
    }
    gs->mesh = mesh;
    gs->ns = ns;
    gs->gstride = (int64_t)(gb / sizeof(float));
    gs->view = m;
    auto rebase = [&](float*& p) { p = (float*)((char*)gs->block + ((char*)p - (char*)m.block)); };
    rebase(gs->view.verts);
    rebase(gs->view.rec6);
    rebase(gs->view.rec3);
    rebase(gs->view.rec1);
    rebase(gs->view.curls);
    rebase(gs->view.geo);
    *out = gs;
    return BEM_OK;
}

int bem_geomset_destroy(bem_geomset* gs)
{
    if (!gs) return BEM_OK;
    DeviceGuard g(gs->view.device);
    if (gs->block) (void)hipFree(gs->block);
    delete gs;
    return BEM_OK;
}

int bem_geomset_count(const bem_geomset* gs) { return gs ? gs->ns : 0; }

int bem_assemble_samples(const bem_geomset* gs, const float* ks, int op, int basis, int quad, int clamp, void* out,
                         int64_t ld, int64_t stride, void* stream)
{
    if (!gs || !out) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (int rc = check_op(op, basis, quad)) return rc;
    if (int rc = check_ks(ks, gs->ns, op)) return rc;
    const auto& m = gs->view;
    const int64_t n = basis == BEM_BASIS_DP0 ? m.nt : m.nv;
    if (ld < n) return fail(BEM_ERR_BAD_ARG, "ld smaller than the matrix width");
    if (gs->ns > 1 && stride < n * ld) return fail(BEM_ERR_BAD_ARG, "stride smaller than one matrix (n * ld)");
    DeviceGuard g(m.device);
    if (!g.ok) return fail(BEM_ERR_HIP, "cannot select the mesh's device");
    hipError_t e = bem::launch_assemble_samples(m, gs->gstride, ks, gs->ns, op, basis, quad, clamp, (float2*)out, ld,
                                                stride, S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_assemble_samples launch");
    return BEM_OK;
}

int bem_add_identity_samples(const bem_geomset* gs, int basis, const float* alphas, void* out, int64_t ld,
                             int64_t stride, void* stream)
{
    if (!gs || !alphas || !out) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (basis != BEM_BASIS_P1 && basis != BEM_BASIS_DP0) return fail(BEM_ERR_BAD_ARG, "unknown basis");
    const auto& m = gs->view;
    const int64_t n = basis == BEM_BASIS_DP0 ? m.nt : m.nv;
    if (ld < n) return fail(BEM_ERR_BAD_ARG, "ld smaller than the matrix width");
    if (gs->ns > 1 && stride < n * ld) return fail(BEM_ERR_BAD_ARG, "stride smaller than one matrix (n * ld)");
    DeviceGuard g(m.device);
    hipError_t e = bem::launch_add_identity_samples(m, gs->gstride, basis, alphas, gs->ns, (float2*)out, ld, stride,
                                                    S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_add_identity_samples launch");
    return BEM_OK;
}

int bem_burton_miller_samples(const bem_geomset* gs, const float* ks, int quad, int clamp, const void* g, void* lhs,
                              int64_t ld, int64_t stride, void* rhs, void* stream)
{
    if (!gs || !g || !lhs || !rhs) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (quad < 0 || quad > 2) return fail(BEM_ERR_BAD_ARG, "unknown quadrature");
    if (int rc = check_bm_ks(ks, gs->ns)) return rc;
    const auto& m = gs->view;
    if (ld < m.nv) return fail(BEM_ERR_BAD_ARG, "ld smaller than the matrix width");
    if (gs->ns > 1 && stride < m.nv * ld) return fail(BEM_ERR_BAD_ARG, "stride smaller than one matrix (nv * ld)");
    DeviceGuard dg(m.device);
    if (!dg.ok) return fail(BEM_ERR_HIP, "cannot select the mesh's device");
    hipError_t e = bem::launch_bm_systems_p1(m, gs->gstride, ks, gs->ns, quad, clamp, (const float2*)g, (float2*)lhs,
                                             ld, stride, (float2*)rhs, S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_burton_miller_samples launch");
    return BEM_OK;
}

int bem_potential_samples(const bem_geomset* gs, const float* points, int64_t np, const float* ks, int op, int basis,
                          int clamp, void* out, int64_t ld, int64_t stride, void* stream)
{
    if (!gs || !out || (np > 0 && !points)) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (op != BEM_OP_SINGLE && op != BEM_OP_DOUBLE)
        return fail(BEM_ERR_BAD_ARG, "potential operators are SINGLE and DOUBLE");
    if (basis != BEM_BASIS_P1 && basis != BEM_BASIS_DP0) return fail(BEM_ERR_BAD_ARG, "unknown basis");
    if (int rc = check_ks(ks, gs->ns, op)) return rc;
    const auto& m = gs->view;
    if (np < 0) return fail(BEM_ERR_BAD_ARG, "negative point count");
    const int64_t n = basis == BEM_BASIS_DP0 ? m.nt : m.nv;
    if (ld < n) return fail(BEM_ERR_BAD_ARG, "ld too small");
    if (gs->ns > 1 && stride < np * ld) return fail(BEM_ERR_BAD_ARG, "stride smaller than one matrix (np * ld)");
    DeviceGuard g(m.device);
    hipError_t e = bem::launch_potential_samples(m, gs->gstride, points, np, ks, gs->ns, op, basis, clamp,
                                                 (float2*)out, ld, stride, S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_potential_samples launch");
    return BEM_OK;
}

int bem_matvec_strided(const void* A, int64_t rows, int64_t cols, int64_t lda, int64_t stride_a, const void* x,
                       int64_t stride_x, void* y, int64_t stride_y, int nz, float alpha_re, float alpha_im,
                       float beta_re, float beta_im, void* stream)
{
    if (!A || !x || !y) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (rows < 0 || cols < 0 || lda < cols || nz < 0) return fail(BEM_ERR_BAD_ARG, "bad shape");
    if (nz > 1 && (stride_a < rows * lda || stride_x < cols || stride_y < rows))
        return fail(BEM_ERR_BAD_ARG, "strides smaller than one matrix / vector");
    hipError_t e = bem::launch_gemv_strided((const float2*)A, rows, cols, lda, stride_a, (const float2*)x, stride_x,
                                            (float2*)y, stride_y, nz, make_float2(alpha_re, alpha_im),
                                            make_float2(beta_re, beta_im), S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_matvec_strided launch");
    return BEM_OK;
}

int bem_assemble_rows(const bem_mesh* mesh, float k, int op, int quad, int clamp, int64_t row0, int64_t row1,
                      void* out, int64_t ld, void* stream)
{
    if (!mesh || !out) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (int rc = check_op(op, BEM_BASIS_DP0, quad)) return rc;
    const auto& m = mesh->dev;
    if (row0 < 0 || row1 > m.nt || row0 >= row1) return fail(BEM_ERR_BAD_ARG, "bad row range");
    if (ld < m.nt) return fail(BEM_ERR_BAD_ARG, "ld smaller than the matrix width");
    if ((op == BEM_OP_BM_RHS) && k == 0.f) return fail(BEM_ERR_BAD_ARG, "Burton-Miller coupling i/k needs k != 0");
    DeviceGuard g(m.device);
    hipError_t e =
        bem::launch_assemble(m, op, BEM_BASIS_DP0, quad, clamp, 3, row0, row1, k, (float2*)out, ld, S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_assemble_rows launch");
    return BEM_OK;
}

int bem_boundary_matrix(const float* verts, int64_t nv, const int32_t* tris, int64_t nt, float k, int op, int quad,
                        void* out, void* stream)
{
    bem_mesh* m = nullptr;
    int rc = bem_mesh_create(verts, nv, tris, nt, 1, stream, &m);
    if (rc) return rc;
    rc = bem_assemble(m, k, op, BEM_BASIS_P1, quad, 1, out, nv, stream);
    if (rc == BEM_OK) {
        hipError_t e = hipStreamSynchronize(S(stream));
        if (e != hipSuccess) rc = hip_fail(e, "bem_boundary_matrix sync");
    }
    bem_mesh_destroy(m);
    return rc;
}

int bem_identity(const bem_mesh* mesh, int basis, void* out, int64_t ld, void* stream)
{
    if (!mesh || !out) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (basis != BEM_BASIS_P1 && basis != BEM_BASIS_DP0) return fail(BEM_ERR_BAD_ARG, "unknown basis");
    const auto& m = mesh->dev;
    if (ld < (basis == BEM_BASIS_DP0 ? m.nt : m.nv)) return fail(BEM_ERR_BAD_ARG, "ld too small");
    DeviceGuard g(m.device);
    hipError_t e = bem::launch_identity(m, basis, (float2*)out, ld, S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_identity launch");
    return BEM_OK;
}

int bem_add_identity(const bem_mesh* mesh, int basis, float alpha_re, float alpha_im, void* A, int64_t ld,
                     void* stream)
{
    if (!mesh || !A) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (basis != BEM_BASIS_P1 && basis != BEM_BASIS_DP0) return fail(BEM_ERR_BAD_ARG, "unknown basis");
    const auto& m = mesh->dev;
    if (ld < (basis == BEM_BASIS_DP0 ? m.nt : m.nv)) return fail(BEM_ERR_BAD_ARG, "ld too small");
    DeviceGuard g(m.device);
    hipError_t e = bem::launch_add_identity(m, basis, make_float2(alpha_re, alpha_im), (float2*)A, ld, S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_add_identity launch");
    return BEM_OK;
}

int64_t bem_cbie_work_size(const bem_mesh* mesh)
{
    if (!mesh) return 0;
    return bem::cbie_work_bytes(mesh->dev, 64);
}

int bem_cbie_dp0(const bem_mesh* mesh, float k, int quad, int clamp, const void* psi, void* lhs, int64_t ld,
                 void* rhs, void* work, void* stream)
{
    return bem_cbie_dp0_ex(mesh, k, quad, clamp, 3, psi, lhs, ld, rhs, work, stream);
}

int bem_cbie_dp0_ex(const bem_mesh* mesh, float k, int quad, int clamp, int parts, const void* psi, void* lhs,
                    int64_t ld, void* rhs, void* work, void* stream)
{
    if (!mesh || !psi || !lhs || !rhs || !work) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (parts < 1 || parts > 3) return fail(BEM_ERR_BAD_ARG, "parts must be 1, 2 or 3");
    if (quad < 0 || quad > 2) return fail(BEM_ERR_BAD_ARG, "unknown quadrature");
    const auto& m = mesh->dev;
    if (ld < m.nt) return fail(BEM_ERR_BAD_ARG, "ld smaller than the matrix width");
    if (!std::isfinite(k)) return fail(BEM_ERR_BAD_ARG, "non-finite wavenumber");
    if ((m.nt + kDp0RowTileHost - 1) / kDp0RowTileHost > 65535) return fail(BEM_ERR_BAD_ARG, "mesh too large");
    DeviceGuard g(m.device);
    hipError_t e = bem::launch_cbie_dp0(m, k, quad, clamp, parts, (const float2*)psi, (float2*)lhs, ld, (float2*)rhs,
                                        bem::cbie_work_split(m, 64, work), S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_cbie_dp0 launch");
    return BEM_OK;
}

int64_t bem_adjoint_work_size(const bem_mesh* mesh)
{
    if (!mesh) return 0;
    return bem::adjoint_work_bytes(mesh->dev);
}

int bem_adjoint_cbie_dp0(const bem_mesh* mesh, float k, int clamp, const void* lam, const void* psi, const void* phi,
                         double* grad_vertices, double* grad_k, double* vhl, void* work, void* stream)
{
    if (!mesh || !lam || !psi || !phi || !grad_vertices || !grad_k || !vhl || !work)
        return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (!std::isfinite(k)) return fail(BEM_ERR_BAD_ARG, "non-finite wavenumber");
    DeviceGuard g(mesh->dev.device);
    hipError_t e = bem::launch_adjoint_pairs(mesh->dev, k, clamp, (const float2*)lam, (const float2*)psi,
                                             (const float2*)phi, grad_vertices, grad_k, vhl, work, S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_adjoint_cbie_dp0 launch");
    return BEM_OK;
}

int bem_adjoint_potential_dp0(const bem_mesh* mesh, const float* points, int64_t np, float k, int clamp,
                              const void* pbar, const void* psi, const void* phi, double* grad_vertices,
                              double* grad_k, void* stream)
{
    if (!mesh || !points || !pbar || !psi || !phi || !grad_vertices || !grad_k)
        return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (np < 0) return fail(BEM_ERR_BAD_ARG, "negative point count");
    DeviceGuard g(mesh->dev.device);
    hipError_t e = bem::launch_adjoint_potential(mesh->dev, points, np, k, clamp, (const float2*)pbar,
                                                 (const float2*)psi, (const float2*)phi, grad_vertices, grad_k,
                                                 S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_adjoint_potential_dp0 launch");
    return BEM_OK;
}

int bem_adjoint_bm_p1(const bem_mesh* mesh, float k, int clamp, const void* lam, const void* g, const void* phi,
                      double* grad_vertices, void* stream)
{
    if (!mesh || !lam || !g || !phi || !grad_vertices) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (!std::isfinite(k) || k == 0.f) return fail(BEM_ERR_BAD_ARG, "Burton-Miller coupling i/k needs a finite k != 0");
    DeviceGuard dg(mesh->dev.device);
    hipError_t e = bem::launch_adjoint_bm_p1(mesh->dev, k, clamp, (const float2*)lam, (const float2*)g,
                                             (const float2*)phi, grad_vertices, S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_adjoint_bm_p1 launch");
    return BEM_OK;
}

int bem_adjoint_potential_p1(const bem_mesh* mesh, const float* points, int64_t np, float k, int clamp,
                             const void* pbar, const void* g, const void* phi, double* grad_vertices,
                             double* grad_points, void* stream)
{
    if (!mesh || (np > 0 && (!points || !pbar)) || !g || !phi || !grad_vertices)
        return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (np < 0) return fail(BEM_ERR_BAD_ARG, "negative point count");
    if (!std::isfinite(k)) return fail(BEM_ERR_BAD_ARG, "non-finite wavenumber");
    DeviceGuard dg(mesh->dev.device);
    hipError_t e = bem::launch_adjoint_potential_p1(mesh->dev, points, np, k, clamp, (const float2*)pbar,
                                                    (const float2*)g, (const float2*)phi, grad_vertices, grad_points,
                                                    S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_adjoint_potential_p1 launch");
    return BEM_OK;
}

int bem_triangle2vertex(const bem_mesh* mesh, const void* tri_values, void* vert_values, void* stream)
{
    if (!mesh || !tri_values || !vert_values) return fail(BEM_ERR_BAD_ARG, "null pointer");
    DeviceGuard g(mesh->dev.device);
    hipError_t e = bem::launch_tri2vert(mesh->dev, (const float2*)tri_values, (float2*)vert_values, S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_triangle2vertex launch");
    return BEM_OK;
}

int64_t bem_triangle2vertex_work_size(int64_t nv, int64_t nt)
{
    if (nv <= 0 || nt <= 0) return 0;
    return (int64_t)sizeof(int32_t) * (2 * nv + 3 * nt + 2);
}

int bem_triangle2vertex_raw(const int32_t* tris, int64_t nt, int64_t nv, const void* tri_values, void* vert_values,
                            void* work, void* stream)
{
    if (!tris || !tri_values || !vert_values || !work) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (nv <= 0 || nt <= 0) return fail(BEM_ERR_BAD_ARG, "empty mesh");
    if (3 * nt >= INT32_MAX || nv >= INT32_MAX) return fail(BEM_ERR_BAD_ARG, "mesh too large for 32-bit lists");
    int32_t* w = (int32_t*)work;
    hipError_t e = bem::launch_tri2vert_raw(tris, nt, nv, (const float2*)tri_values, (float2*)vert_values, w,
                                           S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_triangle2vertex_raw launch");
    int32_t flag = 0;
    e = hipMemcpyAsync(&flag, w + 2 * nv + 1 + 3 * nt, sizeof(flag), hipMemcpyDeviceToHost, S(stream));
    if (e == hipSuccess) e = hipStreamSynchronize(S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_triangle2vertex_raw sync");
    if (flag) return fail(BEM_ERR_BAD_MESH, "triangle vertex index out of range");
    return BEM_OK;
}

int bem_potential(const bem_mesh* mesh, const float* points, int64_t np, float k, int op, int basis, int clamp,
                  void* out, int64_t ld, void* stream)
{
    if (!mesh || !out || (np > 0 && !points)) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (op != BEM_OP_SINGLE && op != BEM_OP_DOUBLE)
        return fail(BEM_ERR_BAD_ARG, "potential operators are SINGLE and DOUBLE");
    if (basis != BEM_BASIS_P1 && basis != BEM_BASIS_DP0) return fail(BEM_ERR_BAD_ARG, "unknown basis");
    const auto& m = mesh->dev;
    if (np < 0) return fail(BEM_ERR_BAD_ARG, "negative point count");
    if (ld < (basis == BEM_BASIS_DP0 ? m.nt : m.nv)) return fail(BEM_ERR_BAD_ARG, "ld too small");
    DeviceGuard g(m.device);
    hipError_t e = bem::launch_potential(m, points, np, k, op, basis, clamp, (float2*)out, ld, S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_potential launch");
    return BEM_OK;
}

int bem_potential_dk(const bem_mesh* mesh, const float* points, int64_t np, float k, int op, int basis, int clamp,
                     void* out, int64_t ld, void* stream)
{
    if (!mesh || !out || (np > 0 && !points)) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (op != BEM_OP_SINGLE && op != BEM_OP_DOUBLE)
        return fail(BEM_ERR_BAD_ARG, "potential operators are SINGLE and DOUBLE");
    if (basis != BEM_BASIS_P1 && basis != BEM_BASIS_DP0) return fail(BEM_ERR_BAD_ARG, "unknown basis");
    const auto& m = mesh->dev;
    if (np < 0) return fail(BEM_ERR_BAD_ARG, "negative point count");
    if (ld < (basis == BEM_BASIS_DP0 ? m.nt : m.nv)) return fail(BEM_ERR_BAD_ARG, "ld too small");
    DeviceGuard g(m.device);
    hipError_t e = bem::launch_potential(m, points, np, k, op | bem::kDk, basis, clamp, (float2*)out, ld, S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_potential_dk launch");
    return BEM_OK;
}

int64_t bem_matvec_work_size(int64_t rows, int64_t cols, int trans)
{
    return (int64_t)sizeof(float2) * bem::gemv_work_elems(rows, cols, trans);
}

int bem_potential_batch_k(const bem_mesh* mesh, const float* points, int64_t np, const float* ks, int nk, int op,
                          int basis, int clamp, void* out, int64_t ld, int64_t stride, void* stream)
{
    if (!mesh || !out || (np > 0 && !points)) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (op != BEM_OP_SINGLE && op != BEM_OP_DOUBLE)
        return fail(BEM_ERR_BAD_ARG, "potential operators are SINGLE and DOUBLE");
    if (basis != BEM_BASIS_P1 && basis != BEM_BASIS_DP0) return fail(BEM_ERR_BAD_ARG, "unknown basis");
    if (int rc = check_ks(ks, nk, op)) return rc;
    const auto& m = mesh->dev;
    if (np < 0) return fail(BEM_ERR_BAD_ARG, "negative point count");
    if (ld < (basis == BEM_BASIS_DP0 ? m.nt : m.nv)) return fail(BEM_ERR_BAD_ARG, "ld too small");
    if (nk > 1 && stride < np * ld) return fail(BEM_ERR_BAD_ARG, "stride smaller than one matrix (np * ld)");
    DeviceGuard g(m.device);
    hipError_t e = bem::launch_potential_batch_k(m, points, np, ks, nk, op, basis, clamp, (float2*)out, ld, stride,
                                                 S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_potential_batch_k launch");
    return BEM_OK;
}

int bem_matvec(const void* A, int64_t rows, int64_t cols, int64_t lda, int trans, const void* x, void* y,
               float alpha_re, float alpha_im, float beta_re, float beta_im, void* work, void* stream)
{
    if (!A || !x || !y) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (rows < 0 || cols < 0 || lda < cols) return fail(BEM_ERR_BAD_ARG, "bad matrix shape");
    if (trans < 0 || trans > 2) return fail(BEM_ERR_BAD_ARG, "trans must be 0, 1 or 2");
    if (trans != 0 && bem::gemv_work_elems(rows, cols, trans) > 0 && !work)
        return fail(BEM_ERR_BAD_ARG, "transposed matvec needs a work buffer");
    hipError_t e = bem::launch_gemv((const float2*)A, rows, cols, lda, trans, (const float2*)x, (float2*)y,
                                    make_float2(alpha_re, alpha_im), make_float2(beta_re, beta_im), (float2*)work,
                                    S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_matvec launch");
    return BEM_OK;
}

int bem_combine(const void* A, int64_t lda, const void* B, int64_t ldb, void* C, int64_t ldc, int64_t rows,
                int64_t cols, float a_re, float a_im, float b_re, float b_im, void* stream)
{
    if (!C) return fail(BEM_ERR_BAD_ARG, "null output");
    if (rows < 0 || cols < 0 || ldc < cols || (A && lda < cols) || (B && ldb < cols))
        return fail(BEM_ERR_BAD_ARG, "bad shape");
    hipError_t e = bem::launch_combine((const float2*)A, lda, (const float2*)B, ldb, (float2*)C, ldc, rows, cols,
                                       make_float2(a_re, a_im), make_float2(b_re, b_im), S(stream));
    if (e != hipSuccess) return hip_fail(e, "bem_combine launch");
    return BEM_OK;
}

int64_t bem_gmres_work_size(int64_t n, int restart)
{
    if (n <= 0 || restart <= 0) return 0;
    return (int64_t)gmres_layout(n, restart).total;
}

int bem_gmres(const void* A_, int64_t n, int64_t lda, const void* b_, void* x_, float tol, int restart, int maxiter,
              void* work, bem_solve_info* info, void* stream)
{
    return bem_gmres_ex(A_, n, lda, 0, b_, x_, tol, restart, maxiter, work, info, stream);
}

int bem_gmres_ex(const void* A_, int64_t n, int64_t lda, int trans, const void* b_, void* x_, float tol, int restart,
                 int maxiter, void* work, bem_solve_info* info, void* stream)
{
    return bem_gmres_timed(A_, n, lda, trans, b_, x_, tol, restart, maxiter, work, info, nullptr, nullptr, stream);
}

int bem_gmres_timed(const void* A_, int64_t n, int64_t lda, int trans, const void* b_, void* x_, float tol,
                    int restart, int maxiter, void* work, bem_solve_info* info, double* gemv_ms, int64_t* gemv_count,
                    void* stream)
{
    if (!A_ || !b_ || !x_ || !work) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (trans < 0 || trans > 2) return fail(BEM_ERR_BAD_ARG, "trans must be 0, 1 or 2");
    if (n <= 0 || lda < n || restart <= 0 || maxiter < 0 || !(tol > 0.f))
        return fail(BEM_ERR_BAD_ARG, "bad GMRES argument");
    const int m = restart;
    const GmresLayout L = gmres_layout(n, m);
    char* W = (char*)work;
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
    double* sc = (double*)(W + L.sc);
    float2* tw = (float2*)(W + L.tw);
    double2* hp = (double2*)(W + L.hp);
    const float2* A = (const float2*)A_;
    const float2* b = (const float2*)b_;
    float2* x = (float2*)x_;
    hipStream_t s = S(stream);
    const float2 one = make_float2(1.f, 0.f), zero = make_float2(0.f, 0.f), mone = make_float2(-1.f, 0.f);

    HIP_TRY(bem::launch_norm2(b, n, sc, s));
    double bn2 = 0.0;
    if (int rc = read_double(sc, &bn2, s)) return rc;
    const double bnorm = std::sqrt(bn2);
    bem_solve_info inf{0, 0, 0.0, 0.0};
    if (bnorm == 0.0) {
        HIP_TRY(hipMemsetAsync(x, 0, sizeof(float2) * (size_t)n, s));
        inf.converged = 1;
        if (info) *info = inf;
        return BEM_OK;
    }
    const double target = (double)tol * bnorm;
    // optional HIP-event timing of the matrix-vector products; each product
    // is followed by a host read of a reduction, so its events are complete
    // when the time is read
    hipEvent_t te0 = nullptr, te1 = nullptr;
    const bool timed = gemv_ms || gemv_count;
    if (timed) {
        HIP_TRY(hipEventCreate(&te0));
        HIP_TRY(hipEventCreate(&te1));
        if (gemv_ms) *gemv_ms = 0.0;
        if (gemv_count) *gemv_count = 0;
    }
    auto gemv_timed = [&](const float2* xv, float2* yv, float2 al, float2 be) -> hipError_t {
        if (timed) {
            hipError_t e = hipEventRecord(te0, s);
            if (e != hipSuccess) return e;
        }
        hipError_t e = bem::launch_gemv(A, n, n, lda, trans, xv, yv, al, be, tw, s);
        if (e != hipSuccess || !timed) return e;
        return hipEventRecord(te1, s);
    };
    auto gemv_account = [&]() -> int {
        if (!timed) return BEM_OK;
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, te0, te1));
        if (gemv_ms) *gemv_ms += ms;
        if (gemv_count) *gemv_count += 1;
        return BEM_OK;
    };
    struct EventFree {
        hipEvent_t a, b;
        ~EventFree()
        {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } event_free{te0, te1};
    int its = 0;
    double rnorm = 0.0, implicit = 0.0;
    bool conv = false;
    const int chk = (n <= bem::kArnoldiSmallN && !timed) ? 8 : 1;  // steps per read of the residuals
    while (true) {
        // r = b - A x
        HIP_TRY(hipMemcpyAsync(r, b, sizeof(float2) * (size_t)n, hipMemcpyDeviceToDevice, s));
        HIP_TRY(gemv_timed(x, r, mone, one));
        HIP_TRY(bem::launch_norm2(r, n, sc, s));
        double rn2 = 0.0;
        if (int rc = read_double(sc, &rn2, s)) return rc;
        if (int rc = gemv_account()) return rc;
        rnorm = std::sqrt(rn2);
        if (rnorm <= target) {
            conv = true;
            break;
        }
        if (its >= maxiter) break;
        HIP_TRY(bem::launch_normalize(r, n, sc, V, s));
        HIP_TRY(bem::launch_init_g(g, m, sc, s));
        int nk = 0, checked = 0;
        for (int j = 0; j < m && its < maxiter; ++j) {
            float2* vj = V + (size_t)j * n;
            HIP_TRY(gemv_timed(vj, w, one, zero));
            if (n <= bem::kArnoldiSmallN) {  // launch-bound sizes: the step in one workgroup
                HIP_TRY(bem::launch_arnoldi_small(V, n, j, w, h, h2, H, m, cs, sn, g, res, sc + 1, s));
            } else {
                // classical Gram-Schmidt, twice
                HIP_TRY(bem::launch_multidot(V, n, n, j + 1, w, h, hp, s));
                HIP_TRY(bem::launch_gs_update(V, n, n, j + 1, h, w, s));
                HIP_TRY(bem::launch_multidot(V, n, n, j + 1, w, h2, hp, s));
                HIP_TRY(bem::launch_gs_update(V, n, n, j + 1, h2, w, s));
                HIP_TRY(bem::launch_norm2(w, n, sc + 1, s));
                HIP_TRY(bem::launch_normalize(w, n, sc + 1, V + (size_t)(j + 1) * n, s));
                HIP_TRY(bem::launch_givens(H, m, j, h, h2, sc + 1, cs, sn, g, res, s));
            }
            ++its;
            nk = j + 1;
            // the implicit residuals are read every `chk` steps (small systems:
            // the read is most of a step's time); on convergence the steps past
            // the first converged one are dropped from the update (columns and
            // g entries of earlier steps are final when a step ends), so the
            // solution is the one of stopping there
            const bool last = j + 1 == m || its >= maxiter;
            if (nk - checked < chk && !last) continue;
            // at most chk (<= 8) residuals per read: a fixed host array
            double rs[8];
            const int nrs = nk - checked;
            HIP_TRY(hipMemcpyAsync(rs, res + checked, sizeof(double) * (size_t)nrs, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (int rc = gemv_account()) return rc;
            bool hit = false;
            for (int q = 0; q < nrs; ++q) {
                implicit = rs[q];
                if (implicit <= target) {
                    its -= nrs - 1 - q;
                    nk = checked + (int)q + 1;
                    hit = true;
                    break;
                }
            }
            checked = nk;
            if (hit) break;
        }
        HIP_TRY(bem::launch_backsolve(H, m, nk, g, y, s));
        HIP_TRY(bem::launch_update_x(V, n, n, nk, y, x, s));
    }
    inf.iterations = its;
    inf.converged = conv ? 1 : 0;
    inf.rel_residual = rnorm / bnorm;
    inf.implicit_rel_residual = implicit / bnorm;
    if (info) *info = inf;
    if (!conv)
        return fail(BEM_ERR_NOT_CONVERGED, "GMRES reached maxiter with relative residual " +
                                               std::to_string(inf.rel_residual));
    return BEM_OK;
}

int64_t bem_bicgstab_work_size(int64_t n)
{
    if (n <= 0) return 0;
    return (int64_t)(7 * align256(sizeof(float2) * (size_t)n) + align256(sizeof(double2) * 16));
}

int bem_bicgstab(const void* A_, int64_t n, int64_t lda, const void* b_, void* x_, int nsteps, double tol, double atol,
                 void* work, bem_solve_info* info, void* stream)
{
    if (!A_ || !b_ || !x_ || !work) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (n <= 0 || lda < n || nsteps < 0) return fail(BEM_ERR_BAD_ARG, "bad BiCGSTAB argument");
    const size_t vb = align256(sizeof(float2) * (size_t)n);
    char* W = (char*)work;
    float2* r = (float2*)(W);
    float2* rh = (float2*)(W + vb);
    float2* p = (float2*)(W + 2 * vb);
    float2* v = (float2*)(W + 3 * vb);
    float2* sv = (float2*)(W + 4 * vb);
    float2* t = (float2*)(W + 5 * vb);
    float2* q = (float2*)(W + 6 * vb);
    double2* sc = (double2*)(W + 7 * vb);
    double* scd = (double*)(sc + 8);  // real scratch
    const float2* A = (const float2*)A_;
    const float2* b = (const float2*)b_;
    float2* x = (float2*)x_;
    hipStream_t s = S(stream);
    const float2 one = make_float2(1.f, 0.f), zero = make_float2(0.f, 0.f), mone = make_float2(-1.f, 0.f);

    // residual_tol = tol * ||b||^2 (BiCGSTAB.py:40)
    HIP_TRY(bem::launch_norm2(b, n, scd, s));
    double bn2 = 0.0;
    if (int rc = read_double(scd, &bn2, s)) return rc;
    const double residual_tol = tol * bn2;
    // q = b - A x, ||q||^2 (BiCGSTAB.py:51-57)
    auto residual = [&](double* out2) -> int {
        HIP_TRY(hipMemcpyAsync(q, b, sizeof(float2) * (size_t)n, hipMemcpyDeviceToDevice, s));
        HIP_TRY(bem::launch_gemv(A, n, n, lda, 0, x, q, mone, one, nullptr, s));
        HIP_TRY(bem::launch_norm2(q, n, scd, s));
        return read_double(scd, out2, s);
    };
    bem_solve_info inf{0, 0, 0.0, 0.0};
    double rr = 0.0;
    if (int rc = residual(&rr)) return rc;
    bool conv = rr < residual_tol || rr < atol;
    if (!conv) {
        HIP_TRY(hipMemcpyAsync(r, q, sizeof(float2) * (size_t)n, hipMemcpyDeviceToDevice, s));
        // rho = alpha = omega = 1, v = p = 0, r_hat = r (BiCGSTAB.py:44-49)
        double2 init[8] = {};
        init[0] = make_double2(1.0, 0.0);
        init[1] = make_double2(1.0, 0.0);
        init[2] = make_double2(1.0, 0.0);
        HIP_TRY(hipMemcpyAsync(sc, init, sizeof(init), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemsetAsync(v, 0, sizeof(float2) * (size_t)n, s));
        HIP_TRY(hipMemsetAsync(p, 0, sizeof(float2) * (size_t)n, s));
        HIP_TRY(hipMemcpyAsync(rh, r, sizeof(float2) * (size_t)n, hipMemcpyDeviceToDevice, s));
        int left = nsteps;
        while (left) {
            // BiCGSTAB.py:59-83
            HIP_TRY(bem::launch_dot(r, rh, n, 0, sc + 3, s));  // rho = <r, r_hat> (bilinear)
            HIP_TRY(bem::launch_bicg_scalar(0, sc, s));        // beta; rho_prev = rho
            HIP_TRY(bem::launch_bicg_p(r, p, v, sc, n, s));    // p = r + beta (p - omega v)
            HIP_TRY(bem::launch_gemv(A, n, n, lda, 0, p, v, one, zero, nullptr, s));
            HIP_TRY(bem::launch_dot(rh, v, n, 0, sc + 5, s));
            HIP_TRY(bem::launch_bicg_scalar(1, sc, s));                   // alpha = rho / <r_hat, v>
            HIP_TRY(bem::launch_bicg_axpy(r, v, sc, 1, -1.0, sv, n, s));  // s = r - alpha v
            HIP_TRY(bem::launch_gemv(A, n, n, lda, 0, sv, t, one, zero, nullptr, s));
            HIP_TRY(bem::launch_dot(t, sv, n, 0, sc + 6, s));
            HIP_TRY(bem::launch_dot(t, t, n, 0, sc + 7, s));
            HIP_TRY(bem::launch_bicg_scalar(2, sc, s));  // omega = <t, s> / <t, t>
            HIP_TRY(bem::launch_bicg_x(x, p, sv, sc, n, s));  // x += alpha p + omega s
            ++inf.iterations;
            if (int rc = residual(&rr)) return rc;
            if (rr < residual_tol || rr < atol) {
                conv = true;
                break;
            }
            HIP_TRY(bem::launch_bicg_axpy(sv, t, sc, 2, -1.0, r, n, s));  // r = s - omega t
            double2 rho_prev;
            HIP_TRY(hipMemcpyAsync(&rho_prev, sc, sizeof(double2), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (rho_prev.x == 0.0 && rho_prev.y == 0.0) break;  // BiCGSTAB.py:103
            --left;
        }
    }
    inf.converged = conv ? 1 : 0;
    inf.rel_residual = bn2 > 0 ? std::sqrt(rr / bn2) : 0.0;
    if (info) *info = inf;
    return BEM_OK;
}

// ---------------------------------------------------------------------------
// mixed precision: 16-bit storage, matrix-core products, refinement
// ---------------------------------------------------------------------------
int64_t bem_lowp_ld(int64_t cols) { return cols <= 0 ? 0 : (cols + 31) / 32 * 32; }

int64_t bem_lowp_plane_elems(int64_t rows, int64_t cols)
{
    if (rows <= 0 || cols <= 0) return 0;
    return bem::lowp_plane_elems(rows, bem_lowp_ld(cols));
}

int bem_convert_lowp(const void* A, int64_t rows, int64_t cols, int64_t lda, int dtype, float scale, void* Ar,
                     void* Ai, int64_t ldlp, float* scale_used, void* stream)
{
    if (!A || !Ar || !Ai) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (dtype != bem::kDtBF16 && dtype != bem::kDtFP16) return fail(BEM_ERR_BAD_ARG, "dtype must be BF16 or FP16");
    if (rows <= 0 || cols <= 0 || lda < cols || ldlp < bem_lowp_ld(cols) || ldlp % 32)
        return fail(BEM_ERR_BAD_ARG, "bad shape (ldlp must be a multiple of 32 >= cols)");
    hipStream_t st = S(stream);
    if (!(scale > 0.f)) {
        if (dtype == bem::kDtBF16) {
            scale = 1.f;
        } else {
            // largest power of two keeping max |A| * scale <= 2^14 (fp16 max 65504)
            const int nb = 1024;
            float* part = nullptr;
            HIP_TRY(hipMallocAsync((void**)&part, sizeof(float) * nb, st));
            HIP_TRY(bem::launch_absmax((const float2*)A, rows, cols, lda, part, nb, st));
            std::vector<float> hp(nb);
            HIP_TRY(hipMemcpyAsync(hp.data(), part, sizeof(float) * nb, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipFreeAsync(part, st));
            HIP_TRY(hipStreamSynchronize(st));
            float mx = 0.f;
            for (float v : hp) mx = std::max(mx, v);
            if (!std::isfinite(mx)) return fail(BEM_ERR_BAD_ARG, "matrix has non-finite entries");
            scale = mx > 0.f ? std::ldexp(1.f, (int)std::floor(std::log2(16384.0 / mx))) : 1.f;
        }
    }
    if (scale_used) *scale_used = scale;
    HIP_TRY(bem::launch_to_lowp(dtype, (const float2*)A, rows, cols, lda, scale, (uint16_t*)Ar, (uint16_t*)Ai, ldlp,
                                st));
    return BEM_OK;
}

int64_t bem_matmul_work_size(int64_t rows)
{
    return rows > 0 ? bem::splitk_part_elems(rows) * (int64_t)sizeof(float2) : 0;
}

int bem_matmul(const void* A, int64_t rows, int64_t cols, int64_t lda, const void* X, int64_t ldx, int nrhs, void* Y,
               int64_t ldy, float alpha_re, float alpha_im, float beta_re, float beta_im, void* work, void* stream)
{
    if (!A || !X || !Y) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (rows <= 0 || cols <= 0 || lda < cols || nrhs <= 0 || ldx < cols || ldy < rows)
        return fail(BEM_ERR_BAD_ARG, "bad shape");
    for (int s0 = 0; s0 < nrhs; s0 += 16) {
        const int cnt = std::min(16, nrhs - s0);
        HIP_TRY(bem::launch_gemm_c64((const float2*)A, rows, cols, lda, (const float2*)X + (int64_t)s0 * ldx, ldx, cnt,
                                     (float2*)Y + (int64_t)s0 * ldy, ldy, make_float2(alpha_re, alpha_im),
                                     make_float2(beta_re, beta_im), (float2*)work, S(stream)));
    }
    return BEM_OK;
}

int64_t bem_matmul_lowp_work_size(int64_t ldlp) { return (int64_t)bem::lowp_scratch_bytes(ldlp); }

int bem_matmul_lowp(const void* Ar, const void* Ai, int dtype, int64_t rows, int64_t cols, int64_t ldlp,
                    float inv_scale, const void* X, int64_t ldx, int nrhs, void* Y, int64_t ldy, float alpha_re,
                    float alpha_im, float beta_re, float beta_im, void* work, void* stream)
{
    if (!Ar || !Ai || !X || !Y || !work) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (dtype != bem::kDtBF16 && dtype != bem::kDtFP16) return fail(BEM_ERR_BAD_ARG, "dtype must be BF16 or FP16");
    if (rows <= 0 || cols <= 0 || ldlp < bem_lowp_ld(cols) || ldlp % 32 || nrhs <= 0 || ldx < cols || ldy < rows)
        return fail(BEM_ERR_BAD_ARG, "bad shape");
    uint16_t* Xr = (uint16_t*)work;
    uint16_t* Xi = Xr + 32 * ldlp;
    float2* part = rows <= ldlp ? bem::lowp_scratch_part(work, ldlp) : nullptr;
    hipStream_t st = S(stream);
    for (int s0 = 0; s0 < nrhs; s0 += 16) {
        const int cnt = std::min(16, nrhs - s0);
        HIP_TRY(bem::launch_vec_to_lowp(dtype, (const float2*)X + (int64_t)s0 * ldx, cols, ldx, cnt, Xr, Xi, ldlp, st));
        HIP_TRY(bem::launch_gemm_lowp(dtype, (const uint16_t*)Ar, (const uint16_t*)Ai, rows, ldlp, inv_scale, Xr, Xi,
                                      ldlp, cnt, (float2*)Y + (int64_t)s0 * ldy, ldy, make_float2(alpha_re, alpha_im),
                                      make_float2(beta_re, beta_im), part, st));
    }
    return BEM_OK;
}

// scratch of the multi-RHS operators: 16-bit right-hand-side planes and the
// split-K partial sums, laid out for max(ldlp, padded n)
static size_t lowp_work_bytes(int64_t n, int64_t ldlp)
{
    return bem::lowp_scratch_bytes(std::max<int64_t>(ldlp, bem_lowp_ld(n)));
}

static int make_op(bem::MultiOp& op, const void* A, int64_t lda, const void* Ar, const void* Ai, int dtype,
                   int64_t ldlp, float inv_scale, int64_t n, void* scratch)
{
    op.n = n;
    op.dtype = dtype;
    const int64_t ls = std::max<int64_t>(ldlp, bem_lowp_ld(n));  // layout of the scratch (see lowp_work_bytes)
    if (scratch) op.part = bem::lowp_scratch_part(scratch, ls);
    if (dtype == 0) {
        if (!A || lda < n) return fail(BEM_ERR_BAD_ARG, "complex64 matrix missing or lda < n");
        op.A = (const float2*)A;
        op.lda = lda;
        return BEM_OK;
    }
    if (dtype != bem::kDtBF16 && dtype != bem::kDtFP16) return fail(BEM_ERR_BAD_ARG, "unknown dtype");
    if (!Ar || !Ai || ldlp < bem_lowp_ld(n) || ldlp % 32) return fail(BEM_ERR_BAD_ARG, "16-bit planes missing/bad ldlp");
    op.Ar = (const uint16_t*)Ar;
    op.Ai = (const uint16_t*)Ai;
    op.ldlp = ldlp;
    op.inv_scale = inv_scale;
    op.Xr = (uint16_t*)scratch;  // Xr (16 ls) + its low plane (16 ls), then Xi likewise
    op.Xi = op.Xr + 32 * ls;
    return BEM_OK;
}

int64_t bem_gmres_multi_work_size(int64_t n, int nrhs, int restart, int64_t ldlp)
{
    if (n <= 0 || nrhs <= 0 || restart <= 0) return 0;
    return (int64_t)(bem::gmres_multi_bytes(n, nrhs, restart) + lowp_work_bytes(n, ldlp));
}

int bem_gmres_multi(const void* A, int64_t lda, const void* Ar, const void* Ai, int dtype, int64_t ldlp,
                    float inv_scale, int64_t n, int nrhs, const void* B, int64_t ldb, void* X, int64_t ldx, float tol,
                    int restart, int maxiter, void* work, bem_solve_info* info, void* stream)
{
    if (!B || !X || !work) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (n <= 0 || nrhs <= 0 || restart <= 0 || maxiter < 0 || ldb < n || ldx < n || !(tol > 0.f))
        return fail(BEM_ERR_BAD_ARG, "bad argument");
    bem::MultiOp op;
    char* scratch = (char*)work + bem::gmres_multi_bytes(n, nrhs, restart);
    if (int rc = make_op(op, A, lda, Ar, Ai, dtype, ldlp, inv_scale, n, scratch)) return rc;
    bem::MultiInfo mi;
    HIP_TRY(bem::gmres_multi(op, n, nrhs, (const float2*)B, ldb, (float2*)X, ldx, tol, restart, maxiter, work, &mi,
                             S(stream)));
    if (info)
        for (int q = 0; q < nrhs; ++q) {
            info[q].iterations = mi.iterations;
            info[q].converged = mi.rel_residual[q] <= tol ? 1 : 0;
            info[q].rel_residual = mi.rel_residual[q];
            info[q].implicit_rel_residual = 0.0;
        }
    if (!mi.converged) return fail(BEM_ERR_NOT_CONVERGED, "multi-RHS GMRES reached maxiter");
    return BEM_OK;
}

int64_t bem_solve_ir_work_size(int64_t n, int nrhs, int restart, int64_t ldlp)
{
    if (n <= 0 || nrhs <= 0 || restart <= 0) return 0;
    return (int64_t)(bem::solve_ir_bytes(n, nrhs, restart) + lowp_work_bytes(n, ldlp));
}

int bem_solve_ir(const void* A, int64_t lda, const void* Ar, const void* Ai, int dtype, int64_t ldlp, float inv_scale,
                 int64_t n, int nrhs, const void* B, int64_t ldb, void* X, int64_t ldx, float tol, float inner_tol,
                 int restart, int max_inner, int max_outer, void* work, bem_solve_info* info,
                 int32_t* outer_iterations, void* stream)
{
    if (!B || !X || !work || !A) return fail(BEM_ERR_BAD_ARG, "null pointer");
    if (n <= 0 || nrhs <= 0 || restart <= 0 || max_inner < 1 || max_outer < 1 || ldb < n || ldx < n ||
        !(tol > 0.f) || !(inner_tol > 0.f))
        return fail(BEM_ERR_BAD_ARG, "bad argument");
    if (dtype != bem::kDtBF16 && dtype != bem::kDtFP16) return fail(BEM_ERR_BAD_ARG, "inner dtype must be BF16/FP16");
    bem::MultiOp hi, lo;
    char* scratch = (char*)work + bem::solve_ir_bytes(n, nrhs, restart);
    if (int rc = make_op(hi, A, lda, nullptr, nullptr, 0, ldlp, 1.f, n, scratch)) return rc;
    if (int rc = make_op(lo, nullptr, 0, Ar, Ai, dtype, ldlp, inv_scale, n, scratch)) return rc;
    bem::IrInfo ir;
    HIP_TRY(bem::solve_ir(hi, lo, n, nrhs, (const float2*)B, ldb, (float2*)X, ldx, tol, inner_tol, restart, max_inner,
                          max_outer, work, &ir, S(stream)));
    if (outer_iterations) *outer_iterations = ir.outer;
    if (info)
        for (int q = 0; q < nrhs; ++q) {
            info[q].iterations = ir.inner;
            info[q].converged = ir.rel_residual[q] <= tol ? 1 : 0;
            info[q].rel_residual = ir.rel_residual[q];
            info[q].implicit_rel_residual = 0.0;
        }
    if (!ir.converged) return fail(BEM_ERR_NOT_CONVERGED, "iterative refinement reached max_outer");
    return BEM_OK;
}

}  // extern "C"
