// Adjoint (reverse-mode) kernels of the P1 Burton-Miller pipeline for gfx950:
// the vertex (and listener) gradients of
//   (bm_lhs + M/2) phi = (bm_rhs - (i/2k) M) g,   p = D_pot phi - S_pot g
// (neumann2dirichlet + boundary2potential with the approximate rules,
// src/bem/solver.py:118-126,140-147).  For a real loss of p, with
// pbar = dL/dp* and lambda solving (bm_lhs + M/2)^H lambda = D_pot^H pbar,
//   dL = 2 Re[ lambda^H (dR g - dL phi) + pbar^H (dD_pot phi - dS_pot g) ].
// Each derivative block is recomputed on the fly and contracted at once.
//
// Per Gauss-point pair (x on the row triangle, y on the column triangle,
// d = y - x, r = |d|) the contracted integrand is written
//   f = cG G + cT F tau + cS F sigma + cN nu G
// with G = e^{ikr}/r, F = G (1 - ikr), tau = d.n_y / r^2,
// sigma = -d.n_x / r^2, nu = n_x.n_y (4 pi times the kernels; the 1/(4 pi)
// sits in the column weights) and complex coefficients:
//   bm_rhs  -G + (i/k) F sigma     (the x-side normal)      cG -= A, cS += (i/k) A
//   bm_lhs  F tau - i k nu G       (minus: it enters as -dL) cT -= B, cN += i k B
//   bm_lhs  curl term (i/k) (curl_a.curl_b) G                 cG -= (i/k) CL.CF
//   D_pot  -F tau, S_pot G (points have no normal)
// where A = l(x) g(y), B = l(x) phi(y) are the P1 interpolants of conj(lambda)
// on the row triangle and of g, phi on the column triangle, CL = sum_a
// conj(lambda_a) curl_a and CF = sum_b phi_b curl_b.  With q = ik - 1/r,
// F' = k^2 e^{ikr} - F/r and rho = d/r (0 where the r-clamp is active):
//   df/dd   = rho [cG G q + cT (F' tau - 2 F tau / r) + cS (F' sigma - 2 F sigma / r)
//              + cN nu G q] + n_y cT F / r^2 - n_x cS F / r^2
//   df/dn_x = -cS F d / r^2 + cN G n_y,   df/dn_y = cT F d / r^2 + cN G n_x.
// Quadrature points move with the vertices through the P1 shape values,
// the weights with |J|, the normals as c/|c| (c = (v1 - v0) x (v2 - v0)) and
// the curls as curl_0 = (v1 - v2)/|c|, curl_1 = (v2 - v0)/|c|,
// curl_2 = (v0 - v1)/|c|.  Gradients are summed in FP64 with atomics.
#include "bem_device.hpp"
#include "bem_internal.hpp"

namespace bem {

namespace {

__device__ __forceinline__ cf cmul(cf a, cf b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cf cconjmul(float2 a, cf b)  // conj(a) b
{
    return {a.x * b.r + a.y * b.i, a.x * b.i - a.y * b.r};
}
__device__ __forceinline__ cf cf2(float2 a) { return {a.x, a.y}; }
__device__ __forceinline__ cf cconj(float2 a) { return {a.x, -a.y}; }

struct BmAdj {
    float gd[3];   // d(2 Re f)/dd
    float gnx[3];  // d(2 Re f)/dn_x
    float gny[3];  // d(2 Re f)/dn_y
    float e;       // 2 Re f (the Jacobian terms)
    float Gr, Gi;  // G (the curl term's pair integral)
};

template <bool CLAMP>
__device__ __forceinline__ void bm_adj_eval(float dx, float dy, float dz, const float* nx, const float* ny, float k,
                                            float krev, cf cG, cf cT, cf cS, cf cN, BmAdj& o)
{
    float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const bool clamped = CLAMP && r2 < kEps2;
    if (CLAMP) r2 = fmaxf(r2, kEps2);
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;
    float sn, cs;
    sincos_rev(r * krev, sn, cs);
    const cf G = {cs * rinv, sn * rinv};
    const float th = k * r;
    const cf F = {fmaf(th, G.i, G.r), fmaf(-th, G.r, G.i)};  // G (1 - i k r)
    const float kk = k * k;
    const cf Fp = {kk * cs - F.r * rinv, kk * sn - F.i * rinv};
    const cf Gq = cmul(G, cf{-rinv, k});
    const float r2inv = rinv * rinv;
    const float tau = (dx * ny[0] + dy * ny[1] + dz * ny[2]) * r2inv;
    const float sig = -(dx * nx[0] + dy * nx[1] + dz * nx[2]) * r2inv;
    const float nu = nx[0] * ny[0] + nx[1] * ny[1] + nx[2] * ny[2];
    // f
    const cf Ft = {F.r * tau, F.i * tau}, Fs = {F.r * sig, F.i * sig};
    const cf a1 = cmul(cG, G), a2 = cmul(cT, Ft), a3 = cmul(cS, Fs), a4 = cmul(cN, G);
    o.e = 2.f * (a1.r + a2.r + a3.r + a4.r * nu);
    // radial coefficient
    const float two_r = 2.f * rinv;
    const cf Pt = {(Fp.r - F.r * two_r) * tau, (Fp.i - F.i * two_r) * tau};
    const cf Ps = {(Fp.r - F.r * two_r) * sig, (Fp.i - F.i * two_r) * sig};
    const cf p1 = cmul(cG, Gq), p2 = cmul(cT, Pt), p3 = cmul(cS, Ps), p4 = cmul(cN, Gq);
    const float pr = 2.f * (p1.r + p2.r + p3.r + p4.r * nu);
    const float ty = 2.f * cmul(cT, F).r * r2inv;   // n_y term of df/dd, d term of df/dn_y
    const float sx = -2.f * cmul(cS, F).r * r2inv;  // n_x term of df/dd, d term of df/dn_x
    const float ng = 2.f * a4.r;                    // 2 Re(cN G)
    const float rs = clamped ? 0.f : rinv;
    const float d[3] = {dx, dy, dz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        o.gd[c] = pr * d[c] * rs + ty * ny[c] + sx * nx[c];
        o.gnx[c] = sx * d[c] + ng * ny[c];
        o.gny[c] = ty * d[c] + ng * nx[c];
    }
    o.Gr = G.r;
    o.Gi = G.i;
}

__device__ __forceinline__ float wsum(float v) { return wave_sum64(v); }

__device__ __forceinline__ double wsum_d(double v) { return wave_sum64(v); }

// dL/dc (c = (v1 - v0) x (v2 - v0)) to the three vertices, added into gv
__device__ __forceinline__ void chain_c(const float* v, const float* gc, float gv[3][3])
{
    const float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
    const float e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
    const float a[3] = {e2[1] * gc[2] - e2[2] * gc[1], e2[2] * gc[0] - e2[0] * gc[2], e2[0] * gc[1] - e2[1] * gc[0]};
    const float b[3] = {gc[1] * e1[2] - gc[2] * e1[1], gc[2] * e1[0] - gc[0] * e1[2], gc[0] * e1[1] - gc[1] * e1[0]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        gv[1][c] += a[c];
        gv[2][c] += b[c];
        gv[0][c] -= a[c] + b[c];
    }
}

// Triangle-level chain: the normal gradient gn, the Jacobian term eJ (dL/d|c|
// from the weights) and the curl gradients gcu[a][3] (may be null) of the
// triangle whose geo record is g (v0 v1 v2, n, |c|); adds into gv.
__device__ __forceinline__ void triangle_chain(const float* g, const float* gn, float eJ, const float (*gcu)[3],
                                               float gv[3][3])
{
    const float n[3] = {g[9], g[10], g[11]};
    const float J = g[12];
    const float invJ = 1.f / J;
    float gJ = eJ;
    if (gcu) {
        // curl_0 = (v1 - v2)/J, curl_1 = (v2 - v0)/J, curl_2 = (v0 - v1)/J
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gv[0][c] += (gcu[2][c] - gcu[1][c]) * invJ;
            gv[1][c] += (gcu[0][c] - gcu[2][c]) * invJ;
            gv[2][c] += (gcu[1][c] - gcu[0][c]) * invJ;
        }
        // d/dJ of u_a / J: -(sum_a gcu_a . curl_a) / J
        const float u0[3] = {g[3] - g[6], g[4] - g[7], g[5] - g[8]};
        const float u1[3] = {g[6] - g[0], g[7] - g[1], g[8] - g[2]};
        const float u2[3] = {g[0] - g[3], g[1] - g[4], g[2] - g[5]};
        float s = 0.f;
#pragma unroll
        for (int c = 0; c < 3; ++c) s += gcu[0][c] * u0[c] + gcu[1][c] * u1[c] + gcu[2][c] * u2[c];
        gJ -= s * invJ * invJ;
    }
    const float gnn = gn[0] * n[0] + gn[1] * n[1] + gn[2] * n[2];
    float gc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) gc[c] = gJ * n[c] + (gn[c] - gnn * n[c]) * invJ;
    chain_c(g, gc, gv);
}

__device__ __forceinline__ void add_vertex_grads(double* gvert, int4 t, const float gv[3][3])
{
    const int ids[3] = {t.x, t.y, t.z};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) unsafeAtomicAdd(&gvert[3 * (int64_t)ids[a] + c], (double)gv[a][c]);
}

// curl gradients of the contraction 2 Re(-(i/k) w (sum_a s_a curl_a) . V):
// gcu[a][c] = 2 Im(s_a V_c) / k  (s_a, V_c complex)
__device__ __forceinline__ void curl_grads(const cf* s, const cf* V, float invk, float gcu[3][3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) gcu[a][c] = 2.f * cmul(s[a], V[c]).i * invk;
}

}  // namespace

// ---------------------------------------------------------------------------
// regular pairs (3-point rule): lane = column triangle, a row tile of 64 row
// triangles per workgroup row; row-side gradients are reduced over the
// wavefront per row triangle, column-side ones kept per lane.
// ---------------------------------------------------------------------------
constexpr int kAdjP1RowTile = 64;

template <bool CLAMP>
__global__ __launch_bounds__(256) void adj_bm_p1_regular_kernel(const float* __restrict__ rec,
                                                                const float* __restrict__ geo,
                                                                const float* __restrict__ curls,
                                                                const int4* __restrict__ tris, int64_t nt, float k,
                                                                const float2* __restrict__ lam,
                                                                const float2* __restrict__ gn_,
                                                                const float2* __restrict__ phi,
                                                                double* __restrict__ gvert)
{
    constexpr int Q = 3;
    constexpr int S = Rec<Q>::kFloats;
    const int lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool active = j < nt;
    const int64_t jj = active ? j : nt - 1;
    const float* cr = rec + jj * S;
    float py[3 * Q], wy[Q];
#pragma unroll
    for (int b = 0; b < 3 * Q; ++b) py[b] = cr[b];
#pragma unroll
    for (int b = 0; b < Q; ++b) wy[b] = cr[3 * Q + b] * kInv4Pi;
    const float ny[3] = {cr[4 * Q], cr[4 * Q + 1], cr[4 * Q + 2]};
    const int4 tj = tris[jj];
    const int vj[3] = {tj.x, tj.y, tj.z};
    cf gh[Q], fh[Q];  // g and phi at the column points
    cf pv[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) pv[b] = cf2(phi[vj[b]]);
    {
        const cf gvv[3] = {cf2(gn_[vj[0]]), cf2(gn_[vj[1]]), cf2(gn_[vj[2]])};
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const float b1 = TriRule<Q>::x1[q], b2 = TriRule<Q>::x2[q], b0 = 1.f - b1 - b2;
            gh[q] = {b0 * gvv[0].r + b1 * gvv[1].r + b2 * gvv[2].r, b0 * gvv[0].i + b1 * gvv[1].i + b2 * gvv[2].i};
            fh[q] = {b0 * pv[0].r + b1 * pv[1].r + b2 * pv[2].r, b0 * pv[0].i + b1 * pv[1].i + b2 * pv[2].i};
        }
    }
    float cj[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) cj[e] = curls[jj * 12 + e];
    cf CF[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        CF[c] = {pv[0].r * cj[c] + pv[1].r * cj[3 + c] + pv[2].r * cj[6 + c],
                 pv[0].i * cj[c] + pv[1].i * cj[3 + c] + pv[2].i * cj[6 + c]};
    }
    const float krev = k * kInv2Pi;
    const float invk = 1.f / k;

    float gy[Q][3];
#pragma unroll
    for (int q = 0; q < Q; ++q) gy[q][0] = gy[q][1] = gy[q][2] = 0.f;
    float gny[3] = {0.f, 0.f, 0.f}, ecol = 0.f;
    cf colcurl[3] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};  // sum_i Gsum_ij CL_i

    const int64_t i_begin = (int64_t)blockIdx.y * kAdjP1RowTile;
    const int64_t i_end = min(i_begin + (int64_t)kAdjP1RowTile, nt);
    for (int64_t i = i_begin; i < i_end; ++i) {
        const int4 ti = tris[i];
        const bool shared = ti.x == tj.x || ti.x == tj.y || ti.x == tj.z || ti.y == tj.x || ti.y == tj.y ||
                            ti.y == tj.z || ti.z == tj.x || ti.z == tj.y || ti.z == tj.z;
        const bool on = active && !shared;
        const float* rr = rec + i * S;
        const float nx[3] = {rr[4 * Q], rr[4 * Q + 1], rr[4 * Q + 2]};
        const float2 l3[3] = {lam[ti.x], lam[ti.y], lam[ti.z]};
        // CL = sum_a conj(lambda_a) curl_a (row triangle)
        cf CL[3];
        {
            const float* ci = curls + i * 12;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                CL[c] = {l3[0].x * ci[c] + l3[1].x * ci[3 + c] + l3[2].x * ci[6 + c],
                         -(l3[0].y * ci[c] + l3[1].y * ci[3 + c] + l3[2].y * ci[6 + c])};
        }
        const cf dot = {CL[0].r * CF[0].r - CL[0].i * CF[0].i + CL[1].r * CF[1].r - CL[1].i * CF[1].i +
                            CL[2].r * CF[2].r - CL[2].i * CF[2].i,
                        CL[0].r * CF[0].i + CL[0].i * CF[0].r + CL[1].r * CF[1].i + CL[1].i * CF[1].r +
                            CL[2].r * CF[2].i + CL[2].i * CF[2].r};
        const cf C1 = on ? cf{dot.i * invk, -dot.r * invk} : cf{0.f, 0.f};  // -(i/k) CL.CF
        float gxv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        float gnx[3] = {0.f, 0.f, 0.f}, erow = 0.f;
        cf Gs = {0.f, 0.f};
#pragma unroll
        for (int a = 0; a < Q; ++a) {
            const float xa0 = rr[3 * a], xa1 = rr[3 * a + 1], xa2 = rr[3 * a + 2];
            const float ua = rr[3 * Q + a];
            const float p1 = TriRule<Q>::x1[a], p2 = TriRule<Q>::x2[a], p0 = 1.f - p1 - p2;
            // l(x_a) = sum conj(lambda_c) phi_c(x_a)
            const cf la = on ? cf{p0 * l3[0].x + p1 * l3[1].x + p2 * l3[2].x,
                                  -(p0 * l3[0].y + p1 * l3[1].y + p2 * l3[2].y)}
                             : cf{0.f, 0.f};
            float gxa[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int b = 0; b < Q; ++b) {
                const cf A = cmul(la, gh[b]), B = cmul(la, fh[b]);
                const cf cG = {C1.r - A.r, C1.i - A.i};
                const cf cS = {-A.i * invk, A.r * invk};   // (i/k) A
                const cf cT = {-B.r, -B.i};                // -B
                const cf cN = {-k * B.i, k * B.r};         // i k B
                // off-pair lanes (neighbours, the triangle itself: d may vanish)
                // evaluate a harmless unit displacement with zero coefficients
                const float dx = on ? py[3 * b] - xa0 : 1.f;
                const float dy = on ? py[3 * b + 1] - xa1 : 0.f;
                const float dz = on ? py[3 * b + 2] - xa2 : 0.f;
                BmAdj o;
                bm_adj_eval<CLAMP>(dx, dy, dz, nx, ny, k, krev, cG, cT, cS, cN, o);
                const float w = ua * wy[b];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    gxa[c] -= w * o.gd[c];
                    gy[b][c] += w * o.gd[c];
                    gnx[c] += w * o.gnx[c];
                    gny[c] += w * o.gny[c];
                }
                erow += w * o.e;
                Gs.r += w * o.Gr;
                Gs.i += w * o.Gi;
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                gxv[0][c] += p0 * gxa[c];
                gxv[1][c] += p1 * gxa[c];
                gxv[2][c] += p2 * gxa[c];
            }
        }
        ecol += erow;
        if (!on) Gs = {0.f, 0.f};
        // curl term: column side keeps sum_i Gs CL_i, row side reduces sum_j Gs CF_j
        cf RC[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const cf t = cmul(Gs, CL[c]);
            colcurl[c].r += t.r;
            colcurl[c].i += t.i;
            RC[c] = cmul(Gs, CF[c]);
        }
        float red[19];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) red[3 * a + c] = wsum(gxv[a][c]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            red[9 + c] = wsum(gnx[c]);
            red[12 + 2 * c] = wsum(RC[c].r);
            red[13 + 2 * c] = wsum(RC[c].i);
        }
        red[18] = wsum(erow);
        if (lane == 0) {
            const float* g = geo + i * 16;
            float gv[3][3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int c = 0; c < 3; ++c) gv[a][c] = red[3 * a + c];
            const cf s3[3] = {cconj(l3[0]), cconj(l3[1]), cconj(l3[2])};
            const cf V3[3] = {{red[12], red[13]}, {red[14], red[15]}, {red[16], red[17]}};
            float gcu[3][3];
            curl_grads(s3, V3, invk, gcu);
            const float gnv[3] = {red[9], red[10], red[11]};
            triangle_chain(g, gnv, red[18] / g[12], gcu, gv);
            add_vertex_grads(gvert, ti, gv);
        }
    }
    if (active) {
        float gv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            const float p1 = TriRule<Q>::x1[b], p2 = TriRule<Q>::x2[b], p0 = 1.f - p1 - p2;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                gv[0][c] += p0 * gy[b][c];
                gv[1][c] += p1 * gy[b][c];
                gv[2][c] += p2 * gy[b][c];
            }
        }
        float gcu[3][3];
        curl_grads(pv, colcurl, invk, gcu);
        const float* g = geo + jj * 16;
        triangle_chain(g, gny, ecol / g[12], gcu, gv);
        add_vertex_grads(gvert, tj, gv);
    }
}

// ---------------------------------------------------------------------------
// neighbour pairs: Sauter-Schwab with the 2-point line rule, one wavefront
// per pair (the same regions and point maps as singular_kernel)
// ---------------------------------------------------------------------------
template <bool CLAMP>
__global__ __launch_bounds__(256) void adj_bm_p1_singular_kernel(const float* __restrict__ geo,
                                                                 const float* __restrict__ curls,
                                                                 const int4* __restrict__ tris,
                                                                 const int4* __restrict__ pairs, int64_t npairs,
                                                                 float k, const float2* __restrict__ lam,
                                                                 const float2* __restrict__ gn_,
                                                                 const float2* __restrict__ phi,
                                                                 double* __restrict__ gvert)
{
    constexpr int NL = 2;
    const int lane = threadIdx.x & 63;
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= npairs) return;  // uniform per wavefront
    const int4 pr = pairs[p];
    const int i = __builtin_amdgcn_readfirstlane(pr.x);
    const int j = __builtin_amdgcn_readfirstlane(pr.y);
    const int shared = __builtin_amdgcn_readfirstlane(pr.z);
    const int4 ti = tris[i], tj = tris[j];
    const int ia[3] = {ti.x, ti.y, ti.z}, ja[3] = {tj.x, tj.y, tj.z};
    int pi[3] = {0, 1, 2}, pj[3] = {0, 1, 2};
    if (shared == 2) {
        int c = 0;
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a)
                if (ia[a] == ja[b]) { pi[c] = a; pj[c] = b; ++c; }
        pi[2] = 3 - pi[0] - pi[1];
        pj[2] = 3 - pj[0] - pj[1];
    } else if (shared == 1) {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                if (ia[a] == ja[b]) {
                    if (a != 0) { pi[0] = a; pi[a] = 0; }
                    if (b != 0) { pj[0] = b; pj[b] = 0; }
                }
    }
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
    float e1[3], e2[3], f1[3], f2[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        e1[d] = v[1][d] - v[0][d];
        e2[d] = v[2][d] - v[1][d];
        f1[d] = w[1][d] - w[0][d];
        f2[d] = w[2][d] - w[1][d];
    }
    const float nx[3] = {gi[9], gi[10], gi[11]};
    const float ny[3] = {gj[9], gj[10], gj[11]};
    const float jij = gi[12] * gj[12] * kInv4Pi;
    const float krev = k * kInv2Pi;
    const float invk = 1.f / k;
    // reordered local values: conj(lambda) on the row triangle, g and phi on the column one
    cf lr[3], gr3[3], fr3[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lr[a] = cconj(lam[ia[pi[a]]]);
        gr3[a] = cf2(gn_[ja[pj[a]]]);
        fr3[a] = cf2(phi[ja[pj[a]]]);
    }
    // curl combinations in the original local order
    const float2 lo3[3] = {lam[ti.x], lam[ti.y], lam[ti.z]};
    const cf po3[3] = {cf2(phi[tj.x]), cf2(phi[tj.y]), cf2(phi[tj.z])};
    float ci[9], cj[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        ci[e] = curls[(int64_t)i * 12 + e];
        cj[e] = curls[(int64_t)j * 12 + e];
    }
    cf CL[3], CF[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        CL[c] = {lo3[0].x * ci[c] + lo3[1].x * ci[3 + c] + lo3[2].x * ci[6 + c],
                 -(lo3[0].y * ci[c] + lo3[1].y * ci[3 + c] + lo3[2].y * ci[6 + c])};
        CF[c] = {po3[0].r * cj[c] + po3[1].r * cj[3 + c] + po3[2].r * cj[6 + c],
                 po3[0].i * cj[c] + po3[1].i * cj[3 + c] + po3[2].i * cj[6 + c]};
    }
    cf dot = {0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const cf t = cmul(CL[c], CF[c]);
        dot.r += t.r;
        dot.i += t.i;
    }
    const cf C1 = {dot.i * invk, -dot.r * invk};  // -(i/k) CL.CF

    // accumulators (reordered local vertices): row 9, col 9, gnx 3, gny 3, e 1, G 2
    double acc[27];
#pragma unroll
    for (int e = 0; e < 27; ++e) acc[e] = 0.0;
    constexpr int NP = NL * NL * NL * NL;
    ss_dispatch(shared, [&](auto kind) {
        constexpr int S = decltype(kind)::value;
        constexpr int NR = kSSRegions<S>;
        for (int q = lane; q < NP; q += 64) {
            const int q0 = q / (NL * NL * NL), q1 = (q / (NL * NL)) % NL, q2 = (q / NL) % NL, q3 = q % NL;
            const float wq =
                LineRule01<NL>::w[q0] * LineRule01<NL>::w[q1] * LineRule01<NL>::w[q2] * LineRule01<NL>::w[q3];
            float reg[NR][7];
            const float jac = ss_regions<S>(LineRule01<NL>::t[q0], LineRule01<NL>::t[q1], LineRule01<NL>::t[q2],
                                            LineRule01<NL>::t[q3], reg);
            const float W = jac * wq * jij;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const float x1 = reg[r][0], x2 = reg[r][1], y1 = reg[r][2], y2 = reg[r][3];
                float dx, dy, dz;
                ss_disp<S>(reg[r], e1, e2, f1, f2, dx, dy, dz);
                const float px[3] = {1.f - x1, x1 - x2, x2};
                const float pyv[3] = {1.f - y1, y1 - y2, y2};
                const cf la = {px[0] * lr[0].r + px[1] * lr[1].r + px[2] * lr[2].r,
                               px[0] * lr[0].i + px[1] * lr[1].i + px[2] * lr[2].i};
                const cf gy = {pyv[0] * gr3[0].r + pyv[1] * gr3[1].r + pyv[2] * gr3[2].r,
                               pyv[0] * gr3[0].i + pyv[1] * gr3[1].i + pyv[2] * gr3[2].i};
                const cf fy = {pyv[0] * fr3[0].r + pyv[1] * fr3[1].r + pyv[2] * fr3[2].r,
                               pyv[0] * fr3[0].i + pyv[1] * fr3[1].i + pyv[2] * fr3[2].i};
                const cf A = cmul(la, gy), B = cmul(la, fy);
                const cf cG = {C1.r - A.r, C1.i - A.i};
                const cf cS = {-A.i * invk, A.r * invk};
                const cf cT = {-B.r, -B.i};
                const cf cN = {-k * B.i, k * B.r};
                BmAdj o;
                bm_adj_eval<CLAMP>(dx, dy, dz, nx, ny, k, krev, cG, cT, cS, cN, o);
                const float ww = W * reg[r][4];
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        acc[3 * a + c] -= (double)(ww * px[a] * o.gd[c]);
                        acc[9 + 3 * a + c] += (double)(ww * pyv[a] * o.gd[c]);
                    }
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    acc[18 + c] += (double)(ww * o.gnx[c]);
                    acc[21 + c] += (double)(ww * o.gny[c]);
                }
                acc[24] += (double)(ww * o.e);
                acc[25] += (double)(ww * o.Gr);
                acc[26] += (double)(ww * o.Gi);
            }
        }
    });
#pragma unroll
    for (int e = 0; e < 27; ++e) acc[e] = wsum_d(acc[e]);
    if (lane != 0) return;
    float gvi[3][3], gvj[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gvi[pi[a]][c] = (float)acc[3 * a + c];
            gvj[pj[a]][c] = (float)acc[9 + 3 * a + c];
        }
    const float E = (float)acc[24];
    const cf Gs = {(float)acc[25], (float)acc[26]};
    cf RC[3], CC[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        RC[c] = cmul(Gs, CF[c]);
        CC[c] = cmul(Gs, CL[c]);
    }
    {
        const cf s3[3] = {cconj(lo3[0]), cconj(lo3[1]), cconj(lo3[2])};
        float gcu[3][3];
        curl_grads(s3, RC, invk, gcu);
        const float gnv[3] = {(float)acc[18], (float)acc[19], (float)acc[20]};
        triangle_chain(gi, gnv, E / gi[12], gcu, gvi);
    }
    {
        float gcu[3][3];
        curl_grads(po3, CC, invk, gcu);
        const float gnv[3] = {(float)acc[21], (float)acc[22], (float)acc[23]};
        triangle_chain(gj, gnv, E / gj[12], gcu, gvj);
    }
    add_vertex_grads(gvert, ti, gvi);
    add_vertex_grads(gvert, tj, gvj);
}

// P1 mass terms: 2 Re[ sum_ab conj(lambda_a) dM_ab (-(i/2k) g_b - phi_b / 2) ],
// M_ab = |c| (1 + delta_ab) / 24 per triangle
__global__ __launch_bounds__(256) void adj_bm_p1_mass_kernel(const float* __restrict__ geo,
                                                             const int4* __restrict__ tris, int64_t nt, float k,
                                                             const float2* __restrict__ lam,
                                                             const float2* __restrict__ gn_,
                                                             const float2* __restrict__ phi,
                                                             double* __restrict__ gvert)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nt) return;
    const int4 tr = tris[t];
    const int v3[3] = {tr.x, tr.y, tr.z};
    const float h = 0.5f / k;
    float s = 0.f;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float2 la = lam[v3[a]], gb = gn_[v3[b]], pb = phi[v3[b]];
            // u = -(i/2k) g_b - phi_b / 2
            const cf u = {h * gb.y - 0.5f * pb.x, -h * gb.x - 0.5f * pb.y};
            s += (a == b ? 2.f : 1.f) * cconjmul(la, u).r;
        }
    const float* g = geo + t * 16;
    const float gJ = 2.f * s / 24.f;
    float gv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const float gc[3] = {gJ * g[9], gJ * g[10], gJ * g[11]};
    chain_c(g, gc, gv);
    add_vertex_grads(gvert, tr, gv);
}

// ---------------------------------------------------------------------------
// potentials (6-point rule): lane = triangle, loop over the points;
// optional point gradients reduced over the wavefront per point
// ---------------------------------------------------------------------------
template <bool CLAMP>
__global__ __launch_bounds__(256) void adj_potential_p1_kernel(const float* __restrict__ rec,
                                                               const float* __restrict__ geo,
                                                               const int4* __restrict__ tris, int64_t nt,
                                                               const float* __restrict__ pts, int64_t np, float k,
                                                               const float2* __restrict__ pbar,
                                                               const float2* __restrict__ gn_,
                                                               const float2* __restrict__ phi,
                                                               double* __restrict__ gvert, double* __restrict__ gpts)
{
    constexpr int Q = 6;
    constexpr int S = Rec<Q>::kFloats;
    const int lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool active = j < nt;
    const int64_t jj = active ? j : nt - 1;
    const float* cr = rec + jj * S;
    float py[3 * Q], wy[Q];
#pragma unroll
    for (int b = 0; b < 3 * Q; ++b) py[b] = cr[b];
#pragma unroll
    for (int b = 0; b < Q; ++b) wy[b] = active ? cr[3 * Q + b] * kInv4Pi : 0.f;
    const float ny[3] = {cr[4 * Q], cr[4 * Q + 1], cr[4 * Q + 2]};
    const float nx0[3] = {0.f, 0.f, 0.f};
    const int4 tj = tris[jj];
    const cf gvv[3] = {cf2(gn_[tj.x]), cf2(gn_[tj.y]), cf2(gn_[tj.z])};
    const cf pvv[3] = {cf2(phi[tj.x]), cf2(phi[tj.y]), cf2(phi[tj.z])};
    cf gh[Q], fh[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const float b1 = TriRule<Q>::x1[q], b2 = TriRule<Q>::x2[q], b0 = 1.f - b1 - b2;
        gh[q] = {b0 * gvv[0].r + b1 * gvv[1].r + b2 * gvv[2].r, b0 * gvv[0].i + b1 * gvv[1].i + b2 * gvv[2].i};
        fh[q] = {b0 * pvv[0].r + b1 * pvv[1].r + b2 * pvv[2].r, b0 * pvv[0].i + b1 * pvv[1].i + b2 * pvv[2].i};
    }
    const float krev = k * kInv2Pi;
    float gy[Q][3];
#pragma unroll
    for (int q = 0; q < Q; ++q) gy[q][0] = gy[q][1] = gy[q][2] = 0.f;
    float gny[3] = {0.f, 0.f, 0.f}, e = 0.f;
    for (int64_t m = 0; m < np; ++m) {
        const float2 pm = pbar[m];
        const float x0 = pts[3 * m], x1 = pts[3 * m + 1], x2 = pts[3 * m + 2];
        float gx[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            // p = D phi - S g: f = conj(pbar) (-F tau phi(y) - G g(y))
            const cf cT = {-(pm.x * fh[b].r + pm.y * fh[b].i), -(pm.x * fh[b].i - pm.y * fh[b].r)};
            const cf cG = {-(pm.x * gh[b].r + pm.y * gh[b].i), -(pm.x * gh[b].i - pm.y * gh[b].r)};
            BmAdj o;
            bm_adj_eval<CLAMP>(py[3 * b] - x0, py[3 * b + 1] - x1, py[3 * b + 2] - x2, nx0, ny, k, krev, cG, cT,
                               cf{0.f, 0.f}, cf{0.f, 0.f}, o);
            const float w = wy[b];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                gy[b][c] += w * o.gd[c];
                gny[c] += w * o.gny[c];
                gx[c] -= w * o.gd[c];
            }
            e += w * o.e;
        }
        if (gpts) {
            const float s0 = wsum(gx[0]), s1 = wsum(gx[1]), s2 = wsum(gx[2]);
            if (lane == 0) {
                unsafeAtomicAdd(&gpts[3 * m], (double)s0);
                unsafeAtomicAdd(&gpts[3 * m + 1], (double)s1);
                unsafeAtomicAdd(&gpts[3 * m + 2], (double)s2);
            }
        }
    }
    if (active) {
        float gv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            const float p1 = TriRule<Q>::x1[b], p2 = TriRule<Q>::x2[b], p0 = 1.f - p1 - p2;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                gv[0][c] += p0 * gy[b][c];
                gv[1][c] += p1 * gy[b][c];
                gv[2][c] += p2 * gy[b][c];
            }
        }
        const float* g = geo + jj * 16;
        triangle_chain(g, gny, e / g[12], nullptr, gv);
        add_vertex_grads(gvert, tj, gv);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_adjoint_bm_p1(const MeshDev& m, float k, int clamp, const float2* lam, const float2* g,
                                const float2* phi, double* gvert, hipStream_t s)
{
    const int64_t nby = (m.nt + kAdjP1RowTile - 1) / kAdjP1RowTile;
    dim3 grid((unsigned)((m.nt + 255) / 256), (unsigned)nby);
    if (clamp)
        hipLaunchKernelGGL(adj_bm_p1_regular_kernel<true>, grid, dim3(256), 0, s, m.rec3, m.geo, m.curls, m.tris, m.nt,
                           k, lam, g, phi, gvert);
    else
        hipLaunchKernelGGL(adj_bm_p1_regular_kernel<false>, grid, dim3(256), 0, s, m.rec3, m.geo, m.curls, m.tris,
                           m.nt, k, lam, g, phi, gvert);
    if (m.npairs > 0) {
        const unsigned gs = (unsigned)((m.npairs + 3) / 4);
        if (clamp)
            hipLaunchKernelGGL(adj_bm_p1_singular_kernel<true>, dim3(gs), dim3(256), 0, s, m.geo, m.curls, m.tris,
                               m.pairs, m.npairs, k, lam, g, phi, gvert);
        else
            hipLaunchKernelGGL(adj_bm_p1_singular_kernel<false>, dim3(gs), dim3(256), 0, s, m.geo, m.curls, m.tris,
                               m.pairs, m.npairs, k, lam, g, phi, gvert);
    }
    hipLaunchKernelGGL(adj_bm_p1_mass_kernel, dim3((unsigned)((m.nt + 255) / 256)), dim3(256), 0, s, m.geo, m.tris,
                       m.nt, k, lam, g, phi, gvert);
    return hipGetLastError();
}

hipError_t launch_adjoint_potential_p1(const MeshDev& m, const float* pts, int64_t np, float k, int clamp,
                                       const float2* pbar, const float2* g, const float2* phi, double* gvert,
                                       double* gpts, hipStream_t s)
{
    dim3 grid((unsigned)((m.nt + 255) / 256));
    if (clamp)
        hipLaunchKernelGGL(adj_potential_p1_kernel<true>, grid, dim3(256), 0, s, m.rec6, m.geo, m.tris, m.nt, pts, np,
                           k, pbar, g, phi, gvert, gpts);
    else
        hipLaunchKernelGGL(adj_potential_p1_kernel<false>, grid, dim3(256), 0, s, m.rec6, m.geo, m.tris, m.nt, pts,
                           np, k, pbar, g, phi, gvert, gpts);
    return hipGetLastError();
}

}  // namespace bem
