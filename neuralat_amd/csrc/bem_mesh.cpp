// Host-side mesh preprocessing: validation, per-triangle quadrature records,
// surface curls, neighbour pairs, vertex->triangle lists and the P1 tiling.
//
// The reference recomputes normals and curls in torch (solver.py:40-58) and
// finds neighbour pairs with an O(N) scan per triangle inside the singular
// kernel (bind.cu:90-97); here they are built once per mesh in O(N log N).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "bem_device.hpp"
#include "bem_internal.hpp"

namespace bem {

namespace {

struct V3 {
    double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 mul(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// quadrature rules in double (same points as the device TriRule<Q>)
struct HostRule {
    std::vector<double> x1, x2, w;
};
HostRule host_rule(int q)
{
    HostRule r;
    if (q == 6) {
        const double a = 0.445948490915965, b = 0.091576213509771;
        const double wa = 0.223381589678011, wb = 0.109951743655322;
        r.x1 = {a, b, 1 - 2 * a, a, 1 - 2 * b, b};
        r.x2 = {a, b, a, 1 - 2 * a, b, 1 - 2 * b};
        r.w = {wa, wb, wa, wa, wb, wb};
    } else if (q == 3) {
        r.x1 = {1.0 / 6, 2.0 / 3, 1.0 / 6};
        r.x2 = {1.0 / 6, 1.0 / 6, 2.0 / 3};
        r.w = {1.0 / 3, 1.0 / 3, 1.0 / 3};
    } else {
        r.x1 = {1.0 / 3};
        r.x2 = {1.0 / 3};
        r.w = {1.0};
    }
    return r;
}

uint32_t spread10(uint32_t v)
{
    v &= 0x3ff;
    v = (v | (v << 16)) & 0x030000FF;
    v = (v | (v << 8)) & 0x0300F00F;
    v = (v | (v << 4)) & 0x030C30C3;
    v = (v | (v << 2)) & 0x09249249;
    return v;
}

struct Tiling {
    std::vector<int32_t> ptr, vptr, verts;
    std::vector<int32_t> slot;  // 4 per position
};

// Greedy tiling of the ordered triangle list: a tile closes when it holds
// max_tris triangles, one more would exceed max_verts distinct vertices, or
// the next position is in stops (sorted; tiles never cross one).
Tiling make_tiling(const std::vector<int32_t>& perm, const int32_t* tris, int max_tris, int max_verts,
                   const std::vector<int32_t>& stops = {})
{
    Tiling T;
    const int64_t nt = (int64_t)perm.size();
    T.slot.assign((size_t)nt * 4, 0);
    T.ptr.push_back(0);
    T.vptr.push_back(0);
    int64_t pos = 0;
    std::vector<int32_t> local;
    size_t next_stop = 0;
    while (pos < nt) {
        local.clear();
        int64_t start = pos;
        while (next_stop < stops.size() && stops[next_stop] <= start) ++next_stop;
        const int64_t limit = next_stop < stops.size() ? std::min<int64_t>(nt, stops[next_stop]) : nt;
        while (pos < limit && pos - start < max_tris) {
            const int32_t* t = tris + 3 * (int64_t)perm[pos];
            int fresh = 0;
            for (int a = 0; a < 3; ++a)
                if (std::find(local.begin(), local.end(), t[a]) == local.end()) ++fresh;
            if ((int)local.size() + fresh > max_verts) break;
            for (int a = 0; a < 3; ++a)
                if (std::find(local.begin(), local.end(), t[a]) == local.end()) local.push_back(t[a]);
            ++pos;
        }
        std::vector<int32_t> sorted = local;
        std::sort(sorted.begin(), sorted.end());
        for (int64_t p = start; p < pos; ++p) {
            const int32_t* t = tris + 3 * (int64_t)perm[p];
            for (int a = 0; a < 3; ++a)
                T.slot[4 * p + a] = (int32_t)(std::lower_bound(sorted.begin(), sorted.end(), t[a]) - sorted.begin());
        }
        T.verts.insert(T.verts.end(), sorted.begin(), sorted.end());
        T.ptr.push_back((int32_t)pos);
        T.vptr.push_back((int32_t)T.verts.size());
    }
    return T;
}

}  // namespace

// Position-dependent host data of a mesh: quadrature records, P1 curls,
// singular-kernel geometry, centroids, bounding-box diagonal (FP64, rounded
// once).  Fails on non-finite coordinates and zero-area triangles.
struct HostGeom {
    std::vector<float> rec6, rec3, rec1, curls, geo;
    std::vector<V3> cent;
    V3 lo{0.0, 0.0, 0.0}, hi{0.0, 0.0, 0.0};  // bounding box
    double diameter = 0.0;
};

static int build_geometry(const float* h_verts, int64_t nv, const int32_t* h_tris, int64_t nt, HostGeom& g,
                          std::string& err)
{
    for (int64_t v = 0; v < 3 * nv; ++v)
        if (!std::isfinite(h_verts[v])) {
            err = "non-finite vertex coordinate";
            return 2;
        }
    constexpr int S6 = Rec<6>::kFloats, S3 = Rec<3>::kFloats, S1 = Rec<1>::kFloats;
    g.rec6.assign((size_t)nt * S6, 0.f);
    g.rec3.assign((size_t)nt * S3, 0.f);
    g.rec1.assign((size_t)nt * S1, 0.f);
    g.curls.assign((size_t)nt * 12, 0.f);
    g.geo.assign((size_t)nt * 16, 0.f);
    g.cent.assign((size_t)nt, V3{0.0, 0.0, 0.0});
    auto& rec6 = g.rec6;
    auto& rec3 = g.rec3;
    auto& rec1 = g.rec1;
    auto& curls = g.curls;
    auto& geo = g.geo;
    auto& cent = g.cent;
    const HostRule R6 = host_rule(6), R3 = host_rule(3), R1 = host_rule(1);
    V3 lo{1e300, 1e300, 1e300}, hi{-1e300, -1e300, -1e300};
    for (int64_t v = 0; v < nv; ++v) {
        lo.x = std::min(lo.x, (double)h_verts[3 * v]);
        hi.x = std::max(hi.x, (double)h_verts[3 * v]);
        lo.y = std::min(lo.y, (double)h_verts[3 * v + 1]);
        hi.y = std::max(hi.y, (double)h_verts[3 * v + 1]);
        lo.z = std::min(lo.z, (double)h_verts[3 * v + 2]);
        hi.z = std::max(hi.z, (double)h_verts[3 * v + 2]);
    }
    g.diameter = std::sqrt(dot(sub(hi, lo), sub(hi, lo)));
    g.lo = lo;
    g.hi = hi;
    for (int64_t t = 0; t < nt; ++t) {
        const int32_t* tr = h_tris + 3 * t;
        V3 v[3];
        for (int a = 0; a < 3; ++a) {
            v[a] = {h_verts[3 * (int64_t)tr[a]], h_verts[3 * (int64_t)tr[a] + 1], h_verts[3 * (int64_t)tr[a] + 2]};
        }
        const V3 e1 = sub(v[1], v[0]), e2 = sub(v[2], v[0]);
        const V3 c = cross(e1, e2);
        const double J = std::sqrt(dot(c, c));
        if (!(J > 0.0) || !std::isfinite(J)) {
            err = "triangle " + std::to_string(t) + " has zero area";
            return 2;
        }
        const V3 n = mul(1.0 / J, c);
        cent[t] = mul(1.0 / 3.0, add(add(v[0], v[1]), v[2]));
        auto fill = [&](float* r, const HostRule& R) {
            const int Q = (int)R.w.size();
            for (int q = 0; q < Q; ++q) {
                const V3 p = add(v[0], add(mul(R.x1[q], e1), mul(R.x2[q], e2)));
                r[3 * q] = (float)p.x;
                r[3 * q + 1] = (float)p.y;
                r[3 * q + 2] = (float)p.z;
                r[3 * Q + q] = (float)(R.w[q] * 0.5 * J);
            }
            r[4 * Q] = (float)n.x;
            r[4 * Q + 1] = (float)n.y;
            r[4 * Q + 2] = (float)n.z;
        };
        fill(&rec6[(size_t)t * S6], R6);
        fill(&rec3[(size_t)t * S3], R3);
        fill(&rec1[(size_t)t * S1], R1);
        // P1 surface curls: row a = n x grad(phi_a), grad from the dual basis
        const double g11 = dot(e1, e1), g12 = dot(e1, e2), g22 = dot(e2, e2);
        const double det = g11 * g22 - g12 * g12;
        const V3 b1 = add(mul(g22 / det, e1), mul(-g12 / det, e2));
        const V3 b2 = add(mul(-g12 / det, e1), mul(g11 / det, e2));
        const V3 grad[3] = {mul(-1.0, add(b1, b2)), b1, b2};
        for (int a = 0; a < 3; ++a) {
            const V3 cu = cross(n, grad[a]);
            curls[(size_t)t * 12 + 3 * a] = (float)cu.x;
            curls[(size_t)t * 12 + 3 * a + 1] = (float)cu.y;
            curls[(size_t)t * 12 + 3 * a + 2] = (float)cu.z;
        }
        float* gr = &geo[(size_t)t * 16];
        for (int a = 0; a < 3; ++a) {
            gr[3 * a] = (float)v[a].x;
            gr[3 * a + 1] = (float)v[a].y;
            gr[3 * a + 2] = (float)v[a].z;
        }
        gr[9] = (float)n.x;
        gr[10] = (float)n.y;
        gr[11] = (float)n.z;
        gr[12] = (float)J;
    }

    return 0;
}

int mesh_build(const float* h_verts, int64_t nv, const int32_t* h_tris, int64_t nt, MeshDev& m, std::string& err)
{
    if (nv < 3 || nt < 1) {
        err = "mesh needs at least 3 vertices and 1 triangle";
        return 1;
    }
    if (nt > (int64_t)1 << 30 || nv > (int64_t)1 << 30) {
        err = "mesh too large (more than 2^30 vertices or triangles)";
        return 1;
    }
    for (int64_t t = 0; t < nt; ++t) {
        const int32_t* tr = h_tris + 3 * t;
        for (int a = 0; a < 3; ++a)
            if (tr[a] < 0 || tr[a] >= nv) {
                err = "triangle " + std::to_string(t) + " has a vertex index out of range";
                return 2;
            }
        if (tr[0] == tr[1] || tr[1] == tr[2] || tr[0] == tr[2]) {
            err = "triangle " + std::to_string(t) + " repeats a vertex";
            return 2;
        }
    }
    m.nv = nv;
    m.nt = nt;

    HostGeom hg;
    if (int rc = build_geometry(h_verts, nv, h_tris, nt, hg, err)) return rc;
    m.diameter = hg.diameter;
    const auto& rec6 = hg.rec6;
    const auto& rec3 = hg.rec3;
    const auto& rec1 = hg.rec1;
    const auto& curls = hg.curls;
    const auto& geo = hg.geo;
    const auto& cent = hg.cent;
    const V3 lo = hg.lo, hi = hg.hi;
    std::vector<int32_t> tris4((size_t)nt * 4, 0);
    for (int64_t t = 0; t < nt; ++t)
        for (int a = 0; a < 3; ++a) tris4[4 * t + a] = h_tris[3 * t + a];
    m.h_tris.assign(h_tris, h_tris + 3 * nt);

    // ---- vertex -> triangle CSR ----
    std::vector<int32_t> vt_ptr((size_t)nv + 1, 0), vt_idx((size_t)nt * 3), vt_corner((size_t)nt * 3);
    for (int64_t t = 0; t < 3 * nt; ++t) vt_ptr[h_tris[t] + 1]++;
    for (int64_t v = 0; v < nv; ++v) vt_ptr[v + 1] += vt_ptr[v];
    {
        std::vector<int32_t> fillp(vt_ptr.begin(), vt_ptr.end() - 1);
        for (int64_t t = 0; t < nt; ++t)
            for (int a = 0; a < 3; ++a) {
                const int32_t q = fillp[h_tris[3 * t + a]]++;
                vt_idx[q] = (int32_t)t;
                vt_corner[q] = (int32_t)(3 * t + a);
            }
    }
    for (int64_t v = 0; v < nv; ++v)
        if (vt_ptr[v + 1] == vt_ptr[v]) {
            err = "vertex " + std::to_string(v) + " belongs to no triangle";
            return 2;
        }

    // ---- neighbour pairs (share >= 1 vertex, self included) ----
    std::vector<int32_t> pairs;
    pairs.reserve((size_t)nt * 14 * 4);
    m.h_pair_ptr.assign((size_t)nt + 1, 0);
    std::vector<int32_t> cand;
    int32_t maxnb = 0;
    for (int64_t t = 0; t < nt; ++t) {
        cand.clear();
        for (int a = 0; a < 3; ++a) {
            const int32_t v = h_tris[3 * t + a];
            cand.insert(cand.end(), vt_idx.begin() + vt_ptr[v], vt_idx.begin() + vt_ptr[v + 1]);
        }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        maxnb = std::max(maxnb, (int32_t)cand.size());
        for (int32_t u : cand) {
            int s = 0;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) s += (h_tris[3 * t + a] == h_tris[3 * (int64_t)u + b]);
            if (s == 3 && u != t) {
                err = "triangles " + std::to_string(t) + " and " + std::to_string(u) + " share all three vertices";
                return 2;
            }
            pairs.push_back((int32_t)t);
            pairs.push_back(u);
            pairs.push_back(s);
            pairs.push_back(0);
        }
        m.h_pair_ptr[t + 1] = (int64_t)pairs.size() / 4;
    }
    m.npairs = (int64_t)pairs.size() / 4;
    // the P1 neighbour-pair launches take every pair in any order: grouped by
    // kind, the pairs that share a wavefront take the same quadrature branch
    std::vector<int32_t> pairs_kind(pairs.size());
    {
        std::vector<int64_t> idx((size_t)m.npairs);
        std::iota(idx.begin(), idx.end(), (int64_t)0);
        std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return pairs[4 * a + 2] > pairs[4 * b + 2]; });
        for (int64_t q = 0; q < m.npairs; ++q)
            for (int c = 0; c < 4; ++c) pairs_kind[4 * q + c] = pairs[4 * idx[q] + c];
    }
    m.max_neighbors = maxnb;

    // ---- P1 tiling on a Morton order of the centroids ----
    std::vector<int32_t> perm((size_t)nt);
    {
        std::vector<uint32_t> code((size_t)nt);
        const double sx = hi.x - lo.x > 0 ? 1023.0 / (hi.x - lo.x) : 0.0;
        const double sy = hi.y - lo.y > 0 ? 1023.0 / (hi.y - lo.y) : 0.0;
        const double sz = hi.z - lo.z > 0 ? 1023.0 / (hi.z - lo.z) : 0.0;
        for (int64_t t = 0; t < nt; ++t) {
            const uint32_t ix = (uint32_t)((cent[t].x - lo.x) * sx);
            const uint32_t iy = (uint32_t)((cent[t].y - lo.y) * sy);
            const uint32_t iz = (uint32_t)((cent[t].z - lo.z) * sz);
            code[t] = spread10(ix) | (spread10(iy) << 1) | (spread10(iz) << 2);
        }
        std::iota(perm.begin(), perm.end(), 0);
        std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { return code[a] < code[b]; });
    }
    // P1 numbering: vertices in the order they first appear along the Morton
    // order of the triangles, so the vertices of a tile are (nearly)
    // consecutive and its accumulator flush lands in few cache lines; the
    // P1 kernels accumulate in this numbering (see p1_vnew / p1_vold)
    std::vector<int32_t> vnew((size_t)nv, -1), vold;
    vold.reserve((size_t)nv);
    for (int64_t q = 0; q < nt; ++q)
        for (int a = 0; a < 3; ++a) {
            const int32_t v = h_tris[3 * (int64_t)perm[q] + a];
            if (vnew[v] < 0) {
                vnew[v] = (int32_t)vold.size();
                vold.push_back(v);
            }
        }
    for (int64_t v = 0; v < nv; ++v)  // vertices in no triangle
        if (vnew[v] < 0) {
            vnew[v] = (int32_t)vold.size();
            vold.push_back((int32_t)v);
        }
    std::vector<int32_t> tris_p1((size_t)nt * 3);
    for (int64_t e = 0; e < 3 * nt; ++e) tris_p1[e] = vnew[h_tris[e]];
    // row tiles nest in column tiles, so the symmetric sweep can pair a row
    // tile with the column tiles from its own on
    const Tiling colT = make_tiling(perm, tris_p1.data(), kP1ColTris, kP1ColVerts);
    const Tiling rowT = make_tiling(perm, tris_p1.data(), kP1RowTris, kP1RowVerts, colT.ptr);
    const int nrowt = (int)rowT.ptr.size() - 1, ncolt = (int)colT.ptr.size() - 1;
    std::vector<int32_t> row_ctile((size_t)nrowt);
    std::vector<int64_t> sym_ptr((size_t)nrowt + 1, 0);
    for (int r = 0, c = 0; r < nrowt; ++r) {
        while (colT.ptr[c + 1] <= rowT.ptr[r]) ++c;
        row_ctile[r] = c;
        sym_ptr[r + 1] = sym_ptr[r] + (ncolt - c);
    }
    // incidences of each column-tile vertex slot (CSR over the global slot index)
    std::vector<int32_t> inc_ptr(colT.verts.size() + 1, 0), inc;
    {
        const int ntile = (int)colT.ptr.size() - 1;
        for (int tl = 0; tl < ntile; ++tl)
            for (int32_t p = colT.ptr[tl]; p < colT.ptr[tl + 1]; ++p)
                for (int b = 0; b < 3; ++b) ++inc_ptr[colT.vptr[tl] + colT.slot[4 * (size_t)p + b] + 1];
        for (size_t g = 0; g + 1 < inc_ptr.size(); ++g) inc_ptr[g + 1] += inc_ptr[g];
        inc.assign((size_t)inc_ptr.back(), 0);
        std::vector<int32_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
        for (int tl = 0; tl < ntile; ++tl)
            for (int32_t p = colT.ptr[tl]; p < colT.ptr[tl + 1]; ++p)
                for (int b = 0; b < 3; ++b) {
                    const int32_t g = colT.vptr[tl] + colT.slot[4 * (size_t)p + b];
                    inc[(size_t)fill[g]++] = 4 * (p - colT.ptr[tl]) + b;
                }
    }

    // ---- one device allocation ----
    struct Part {
        const void* src;
        size_t bytes;
        void** dst;
    };
    std::vector<Part> parts;
    float* d_verts = nullptr;
    std::vector<float> vcopy(h_verts, h_verts + 3 * nv);
    // the position-dependent arrays first: mesh_set_vertices rewrites them as
    // one contiguous region
    parts.push_back({vcopy.data(), vcopy.size() * 4, (void**)&d_verts});
    parts.push_back({rec6.data(), rec6.size() * 4, (void**)&m.rec6});
    parts.push_back({rec3.data(), rec3.size() * 4, (void**)&m.rec3});
    parts.push_back({rec1.data(), rec1.size() * 4, (void**)&m.rec1});
    parts.push_back({curls.data(), curls.size() * 4, (void**)&m.curls});
    parts.push_back({geo.data(), geo.size() * 4, (void**)&m.geo});
    constexpr size_t kGeomParts = 6;
    parts.push_back({tris4.data(), tris4.size() * 4, (void**)&m.tris});
    parts.push_back({pairs.data(), pairs.size() * 4, (void**)&m.pairs});
    parts.push_back({pairs_kind.data(), pairs_kind.size() * 4, (void**)&m.pairs_by_kind});
    parts.push_back({m.h_pair_ptr.data(), m.h_pair_ptr.size() * 8, (void**)&m.pair_ptr});
    parts.push_back({vt_ptr.data(), vt_ptr.size() * 4, (void**)&m.vt_ptr});
    parts.push_back({vt_idx.data(), vt_idx.size() * 4, (void**)&m.vt_idx});
    parts.push_back({vt_corner.data(), vt_corner.size() * 4, (void**)&m.vt_corner});
    int32_t *d_perm, *d_rp, *d_rvp, *d_rv, *d_cp, *d_cvp, *d_cv, *d_cip, *d_ci, *d_rct;
    int64_t* d_sym;
    int4 *d_rs, *d_cs;
    parts.push_back({perm.data(), perm.size() * 4, (void**)&d_perm});
    parts.push_back({rowT.ptr.data(), rowT.ptr.size() * 4, (void**)&d_rp});
    parts.push_back({rowT.vptr.data(), rowT.vptr.size() * 4, (void**)&d_rvp});
    parts.push_back({rowT.verts.data(), rowT.verts.size() * 4, (void**)&d_rv});
    parts.push_back({rowT.slot.data(), rowT.slot.size() * 4, (void**)&d_rs});
    parts.push_back({colT.ptr.data(), colT.ptr.size() * 4, (void**)&d_cp});
    parts.push_back({colT.vptr.data(), colT.vptr.size() * 4, (void**)&d_cvp});
    parts.push_back({colT.verts.data(), colT.verts.size() * 4, (void**)&d_cv});
    parts.push_back({colT.slot.data(), colT.slot.size() * 4, (void**)&d_cs});
    parts.push_back({inc_ptr.data(), inc_ptr.size() * 4, (void**)&d_cip});
    parts.push_back({inc.data(), inc.size() * 4, (void**)&d_ci});
    parts.push_back({row_ctile.data(), row_ctile.size() * 4, (void**)&d_rct});
    parts.push_back({sym_ptr.data(), sym_ptr.size() * 8, (void**)&d_sym});
    parts.push_back({vnew.data(), vnew.size() * 4, (void**)&m.p1_vnew});
    parts.push_back({vold.data(), vold.size() * 4, (void**)&m.p1_vold});
    size_t total = 0;
    std::vector<size_t> off;
    for (auto& p : parts) {
        off.push_back(total);
        total += (p.bytes + 255) / 256 * 256;
    }
    m.geom_bytes = off[kGeomParts];
    if (hipGetDevice(&m.device) != hipSuccess) {
        err = "no HIP device";
        return 5;
    }
    if (hipMalloc(&m.block, total) != hipSuccess) {
        err = "device allocation of " + std::to_string(total) + " bytes failed";
        (void)hipGetLastError();
        return 6;
    }
    // every array through one host staging buffer and one copy (a copy per
    // array cost more than the rest of the set-up for small meshes)
    std::vector<char> stage(total);
    for (size_t q = 0; q < parts.size(); ++q) {
        *parts[q].dst = (char*)m.block + off[q];
        if (parts[q].bytes) std::memcpy(stage.data() + off[q], parts[q].src, parts[q].bytes);
    }
    if (hipMemcpy(m.block, stage.data(), total, hipMemcpyHostToDevice) != hipSuccess) {
        err = "host to device copy failed";
        (void)hipFree(m.block);
        m.block = nullptr;
        return 3;
    }
    m.verts = d_verts;
    m.tiles.perm = d_perm;
    m.tiles.row_ptr = d_rp;
    m.tiles.row_vptr = d_rvp;
    m.tiles.row_verts = d_rv;
    m.tiles.row_slot = d_rs;
    m.tiles.col_ptr = d_cp;
    m.tiles.col_vptr = d_cvp;
    m.tiles.col_verts = d_cv;
    m.tiles.col_slot = d_cs;
    m.tiles.col_inc_ptr = d_cip;
    m.tiles.col_inc = d_ci;
    m.tiles.row_ctile = d_rct;
    m.tiles.sym_ptr = d_sym;
    m.tiles.nsym = sym_ptr.back();
    m.tiles.nrow_tiles = nrowt;
    m.tiles.ncol_tiles = ncolt;
    m.h_row_slots = rowT.vptr.empty() ? 0 : rowT.vptr.back();
    return 0;
}

int mesh_geometry_region(const MeshDev& m, const float* h_verts, char* region, std::string& err)
{
    if (!m.block || m.h_tris.size() != (size_t)m.nt * 3) {
        err = "mesh has no device data";
        return 1;
    }
    HostGeom hg;
    if (int rc = build_geometry(h_verts, m.nv, m.h_tris.data(), m.nt, hg, err)) return rc;
    // the same 256-byte aligned layout as mesh_build's first six parts
    const std::pair<const float*, size_t> sec[6] = {{h_verts, (size_t)m.nv * 3},   {hg.rec6.data(), hg.rec6.size()},
                                                    {hg.rec3.data(), hg.rec3.size()}, {hg.rec1.data(), hg.rec1.size()},
                                                    {hg.curls.data(), hg.curls.size()}, {hg.geo.data(), hg.geo.size()}};
    std::memset(region, 0, m.geom_bytes);
    size_t off = 0;
    for (const auto& q : sec) {
        const size_t bytes = q.second * 4;
        if (off + bytes > m.geom_bytes) {
            err = "geometry layout mismatch";
            return 1;
        }
        std::memcpy(region + off, q.first, bytes);
        off += (bytes + 255) / 256 * 256;
    }
    if (off != m.geom_bytes) {
        err = "geometry layout mismatch";
        return 1;
    }
    return 0;
}

int mesh_set_vertices(MeshDev& m, const float* h_verts, hipStream_t s, std::string& err)
{
    std::vector<char> stage(m.geom_bytes, 0);
    if (int rc = mesh_geometry_region(m, h_verts, stage.data(), err)) return rc;
    if (hipMemcpyAsync(m.block, stage.data(), stage.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        err = "host to device copy failed";
        return 3;
    }
    return 0;
}

void mesh_free(MeshDev& m)
{
    if (m.block) (void)hipFree(m.block);
    m.block = nullptr;
}

}  // namespace bem
