// Adjoint (reverse-mode) kernels of the DP0 CBIE pipeline for gfx950.
//
// For a real loss L of the exterior pressure p = D_pot phi - S_pot psi with
// (M/2 - K) phi = -V psi, the gradient with respect to the vertex positions
// and k is, with pbar = dL/dp* and lambda solving (M/2 - K)^H lambda =
// D_pot^H pbar,
//   dL = 2 Re[ sum_ij alpha_ij dV_ij + beta_ij dK_ij        (alpha = -conj(l_i) psi_j,
//                                                             beta = conj(l_i) phi_j)
//              - 1/4 sum_i conj(l_i) phi_i dJ_i
//              + sum_mi alpha'_mi dS_mi + beta'_mi dD_mi ]   (alpha' = -conj(pbar_m) psi_i,
//                                                             beta' = conj(pbar_m) phi_i)
// Each derivative block is recomputed on the fly, never stored: per
// Gauss-point pair the contracted integrand f = alpha G + beta dG/dn_y is
// differentiated with respect to d = y - x, the normal n_y and k, and chained
// to the vertices through the (linear) quadrature-point maps and the Jacobians
// |J|.  Gradients are summed in FP64 with atomics.
//
// With G = e^{ikr}/(4 pi r), q = ik - 1/r, s = d.n_y, rho = d/r (0 where the
// r-clamp is active), H = G q s / r:
//   grad_d (2 Re f) = 2Re[G (alpha q + beta c1)] rho + 2Re[beta G q] / r * n_y,
//     c1 = q^2 s / r + s / r^3 - q s / r^2
//   grad_n (2 Re f) = 2Re[beta G q] / r * d
//   d(2 Re f)/dk    = -2 Im[ r alpha G + i k s beta G ]
#include "bem_device.hpp"
#include "bem_internal.hpp"

namespace bem {

struct AdjEval {
    float gd[3];  // gradient with respect to d = y - x
    float gn[3];  // with respect to the column normal
    float e;      // 2 Re f (for the Jacobian terms)
    float gk;     // d(2 Re f)/dk
    float gr, gi; // G itself (for the V^H lambda product)
};

// alpha, beta are the complex pair coefficients (already including 4 pi
// folding: the caller scales the weights by 1/(4 pi)).
template <bool CLAMP>
__device__ __forceinline__ void adj_eval(float dx, float dy, float dz, float n0, float n1, float n2, float k, float krev,
                                         float ar, float ai, float br, float bi, AdjEval& o)
{
    float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const bool clamped = CLAMP && r2 < kEps2;
    if (CLAMP) r2 = fmaxf(r2, kEps2);
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;
    float sn, cs;
    sincos_rev(r * krev, sn, cs);
    const float gr = cs * rinv, gi = sn * rinv;  // 4 pi G
    const float s = dx * n0 + dy * n1 + dz * n2;
    // q = -1/r + i k ; q^2 = (1/r^2 - k^2) - 2 i k / r
    const float qr = -rinv, qi = k;
    const float q2r = rinv * rinv - k * k, q2i = -2.f * k * rinv;
    // c1 = s (q^2 / r + 1/r^3 - q / r^2)
    const float r3 = rinv * rinv * rinv;
    const float c1r = s * (q2r * rinv + r3 - qr * rinv * rinv);
    const float c1i = s * (q2i * rinv - qi * rinv * rinv);
    // alpha G, beta G
    const float agr = ar * gr - ai * gi, agi = ar * gi + ai * gr;
    const float bgr = br * gr - bi * gi, bgi = br * gi + bi * gr;
    // A1 = 2 Re(alpha G q + beta G c1), A2 = 2 Re(beta G q) / r
    const float A1 = 2.f * ((agr * qr - agi * qi) + (bgr * c1r - bgi * c1i));
    const float bq = 2.f * (bgr * qr - bgi * qi);
    const float A2 = bq * rinv;
    const float rs = clamped ? 0.f : rinv;
    o.gd[0] = A1 * dx * rs + A2 * n0;
    o.gd[1] = A1 * dy * rs + A2 * n1;
    o.gd[2] = A1 * dz * rs + A2 * n2;
    o.gn[0] = A2 * dx;
    o.gn[1] = A2 * dy;
    o.gn[2] = A2 * dz;
    // f = alpha G + beta G q s / r
    const float hqr = (qr * s) * rinv, hqi = (qi * s) * rinv;
    o.e = 2.f * (agr + (bgr * hqr - bgi * hqi));
    // -2 Im(r alpha G + i k s beta G) = -2 (r agi + k s bgr)
    o.gk = -2.f * (r * agi + k * s * bgr);
    o.gr = gr;
    o.gi = gi;
}

// shape values of the 6-point rule's points (x1, x2), for rolled loops
__constant__ float c_rule6_x1[6] = {0.445948490915965f, 0.091576213509771f, 1.0f - 2.0f * 0.445948490915965f,
                                    0.445948490915965f, 1.0f - 2.0f * 0.091576213509771f, 0.091576213509771f};
__constant__ float c_rule6_x2[6] = {0.445948490915965f, 0.091576213509771f, 0.445948490915965f,
                                    1.0f - 2.0f * 0.445948490915965f, 0.091576213509771f,
                                    1.0f - 2.0f * 0.091576213509771f};

__device__ __forceinline__ float wsum(float v) { return wave_sum64(v); }

// d|J|/dv for the vertex triangle (v0, v1, v2) contracted with g_c (dL/dc,
// c = (v1 - v0) x (v2 - v0)); adds into gv[3][3].
__device__ __forceinline__ void chain_c(const float* v, const float* gc, float gv[3][3])
{
    const float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
    const float e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
    // dL/dv1 = e2 x gc, dL/dv2 = gc x e1, dL/dv0 = -(both)
    const float a[3] = {e2[1] * gc[2] - e2[2] * gc[1], e2[2] * gc[0] - e2[0] * gc[2], e2[0] * gc[1] - e2[1] * gc[0]};
    const float b[3] = {gc[1] * e1[2] - gc[2] * e1[1], gc[2] * e1[0] - gc[0] * e1[2], gc[0] * e1[1] - gc[1] * e1[0]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        gv[1][c] += a[c];
        gv[2][c] += b[c];
        gv[0][c] -= a[c] + b[c];
    }
}

__device__ __forceinline__ void add_vertex_grads(double* gvert, int4 t, const float gv[3][3])
{
    const int ids[3] = {t.x, t.y, t.z};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) unsafeAtomicAdd(&gvert[3 * (int64_t)ids[a] + c], (double)gv[a][c]);
}

// Column-side chain: point gradients gy[b] (d(2Re f)/dy_b summed), normal
// gradient gn, Jacobian term e (sum of 2 Re f), through y_b = sum phi_beta
// v_beta and n = c/|c|, |J| = |c|.
__device__ __forceinline__ void column_chain(const float* geo_j, const float (*gy)[3], const float* x1, const float* x2,
                                             int Q, const float* gn, float e, float gv[3][3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) gv[a][c] = 0.f;
    for (int b = 0; b < Q; ++b) {
        const float p0 = 1.f - x1[b] - x2[b], p1 = x1[b], p2 = x2[b];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gv[0][c] += p0 * gy[b][c];
            gv[1][c] += p1 * gy[b][c];
            gv[2][c] += p2 * gy[b][c];
        }
    }
    const float n[3] = {geo_j[9], geo_j[10], geo_j[11]};
    const float J = geo_j[12];
    const float gnn = gn[0] * n[0] + gn[1] * n[1] + gn[2] * n[2];
    // g_c = (e / J) n + (I - n n^T) gn / J
    float gc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) gc[c] = (e * n[c] + gn[c] - gnn * n[c]) / J;
    chain_c(geo_j, gc, gv);
}

// ---------------------------------------------------------------------------
// regular pairs (DP0, 6-point rule): lane = column triangle, row tile uniform
// ---------------------------------------------------------------------------
constexpr int kAdjRowTile = 64;

template <bool CLAMP>
__global__ __launch_bounds__(256) void adj_regular_kernel(const float* __restrict__ rec, const float* __restrict__ geo,
                                                          const int4* __restrict__ tris, int64_t nt, float k,
                                                          const float2* __restrict__ lam,
                                                          const float2* __restrict__ psi,
                                                          const float2* __restrict__ phi, double* __restrict__ gvert,
                                                          double* __restrict__ gk_out, float2* __restrict__ vhl_part)
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
    for (int b = 0; b < Q; ++b) wy[b] = cr[3 * Q + b] * kInv4Pi;
    const float ny0 = cr[4 * Q], ny1 = cr[4 * Q + 1], ny2 = cr[4 * Q + 2];
    const int4 tj = tris[jj];
    const float2 psj = psi[jj], phj = phi[jj];
    const float krev = k * kInv2Pi;

    float gy[Q][3];
#pragma unroll
    for (int b = 0; b < Q; ++b) gy[b][0] = gy[b][1] = gy[b][2] = 0.f;
    float gn[3] = {0.f, 0.f, 0.f}, ecol = 0.f, gk = 0.f, vr = 0.f, vi = 0.f;

    const int64_t i_begin = (int64_t)blockIdx.y * kAdjRowTile;
    const int64_t i_end = min(i_begin + (int64_t)kAdjRowTile, nt);
    for (int64_t i = i_begin; i < i_end; ++i) {
        const int4 ti = tris[i];
        const bool shared = ti.x == tj.x || ti.x == tj.y || ti.x == tj.z || ti.y == tj.x || ti.y == tj.y ||
                            ti.y == tj.z || ti.z == tj.x || ti.z == tj.y || ti.z == tj.z;
        const bool on = active && !shared;
        const float2 li = lam[i];
        // alpha = -conj(l_i) psi_j, beta = conj(l_i) phi_j (zero off-pair)
        const float ar = on ? -(li.x * psj.x + li.y * psj.y) : 0.f;
        const float ai = on ? -(li.x * psj.y - li.y * psj.x) : 0.f;
        const float br = on ? (li.x * phj.x + li.y * phj.y) : 0.f;
        const float bi = on ? (li.x * phj.y - li.y * phj.x) : 0.f;
        const float* rr = rec + i * S;
        float gxv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        float erow = 0.f, pv_r = 0.f, pv_i = 0.f;
#pragma unroll 1
        for (int a = 0; a < Q; ++a) {
            const float xa0 = rr[3 * a], xa1 = rr[3 * a + 1], xa2 = rr[3 * a + 2];
            const float ua = rr[3 * Q + a];
            float gxa[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int b = 0; b < Q; ++b) {
                AdjEval o;
                adj_eval<CLAMP>(py[3 * b] - xa0, py[3 * b + 1] - xa1, py[3 * b + 2] - xa2, ny0, ny1, ny2, k, krev, ar,
                                ai, br, bi, o);
                const float w = ua * wy[b];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    gxa[c] -= w * o.gd[c];
                    gy[b][c] += w * o.gd[c];
                    gn[c] += w * o.gn[c];
                }
                erow += w * o.e;
                gk += w * o.gk;
                pv_r += w * o.gr;
                pv_i += w * o.gi;
            }
            const float p1 = c_rule6_x1[a], p2 = c_rule6_x2[a], p0 = 1.f - p1 - p2;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                gxv[0][c] += p0 * gxa[c];
                gxv[1][c] += p1 * gxa[c];
                gxv[2][c] += p2 * gxa[c];
            }
        }
        ecol += erow;
        // V^H lambda: conj(V_ij) l_i
        if (on) {
            vr += pv_r * li.x + pv_i * li.y;
            vi += pv_r * li.y - pv_i * li.x;
        }
        // row triangle: reduce over the 64 columns of this wavefront
        float red[10];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) red[3 * a + c] = wsum(gxv[a][c]);
        red[9] = wsum(erow);
        if (lane == 0) {
            const float* gi = geo + i * 16;
            float gv[3][3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int c = 0; c < 3; ++c) gv[a][c] = red[3 * a + c];
            const float eJ = red[9] / gi[12];
            const float gc[3] = {eJ * gi[9], eJ * gi[10], eJ * gi[11]};
            chain_c(gi, gc, gv);
            add_vertex_grads(gvert, ti, gv);
        }
    }
    if (active) {
        float gv[3][3];
        column_chain(geo + jj * 16, gy, TriRule<Q>::x1, TriRule<Q>::x2, Q, gn, ecol, gv);
        add_vertex_grads(gvert, tj, gv);
        vhl_part[(int64_t)blockIdx.y * nt + j] = make_float2(vr, vi);
    }
    const float gks = wsum(gk);
    if (lane == 0) unsafeAtomicAdd(gk_out, (double)gks);
}

// ---------------------------------------------------------------------------
// neighbour pairs: Sauter-Schwab, one wavefront per pair
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wsum_d(double v) { return wave_sum64(v); }

template <bool CLAMP>
__global__ __launch_bounds__(256) void adj_singular_kernel(const float* __restrict__ geo, const int4* __restrict__ tris,
                                                           const int4* __restrict__ pairs, int64_t npairs, float k,
                                                           const float2* __restrict__ lam,
                                                           const float2* __restrict__ psi,
                                                           const float2* __restrict__ phi,
                                                           double* __restrict__ gvert, double* __restrict__ gk_out,
                                                           double* __restrict__ vhl)
{
    constexpr int NL = 4;
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
    const float ny0 = gj[9], ny1 = gj[10], ny2 = gj[11];
    const float jij = gi[12] * gj[12];
    const float krev = k * kInv2Pi;
    const float2 li = lam[i], psj = psi[j], phj = phi[j];
    const float ar = -(li.x * psj.x + li.y * psj.y) * kInv4Pi, ai = -(li.x * psj.y - li.y * psj.x) * kInv4Pi;
    const float br = (li.x * phj.x + li.y * phj.y) * kInv4Pi, bi = (li.x * phj.y - li.y * phj.x) * kInv4Pi;

    // accumulators (reordered local vertices): row 9, col 9, gn 3, e 1, gk 1, V 2
    double acc[25];
#pragma unroll
    for (int e = 0; e < 25; ++e) acc[e] = 0.0;
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
                AdjEval o;
                adj_eval<CLAMP>(dx, dy, dz, ny0, ny1, ny2, k, krev, ar, ai, br, bi, o);
                const float ww = W * reg[r][4];
                const float px[3] = {1.f - x1, x1 - x2, x2};
                const float py[3] = {1.f - y1, y1 - y2, y2};
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        acc[3 * a + c] -= (double)(ww * px[a] * o.gd[c]);
                        acc[9 + 3 * a + c] += (double)(ww * py[a] * o.gd[c]);
                    }
#pragma unroll
                for (int c = 0; c < 3; ++c) acc[18 + c] += (double)(ww * o.gn[c]);
                acc[21] += (double)(ww * o.e);
                acc[22] += (double)(ww * o.gk);
                acc[23] += (double)(ww * o.gr);
                acc[24] += (double)(ww * o.gi);
            }
        }
    });
#pragma unroll
    for (int e = 0; e < 25; ++e) acc[e] = wsum_d(acc[e]);
    if (lane != 0) return;
    // un-permute to the original local vertex order
    float gvi[3][3], gvj[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gvi[pi[a]][c] = (float)acc[3 * a + c];
            gvj[pj[a]][c] = (float)acc[9 + 3 * a + c];
        }
    const float E = (float)acc[21];
    {
        const float eJ = E / gi[12];
        const float gc[3] = {eJ * gi[9], eJ * gi[10], eJ * gi[11]};
        chain_c(gi, gc, gvi);
    }
    {
        const float n[3] = {ny0, ny1, ny2};
        const float gn[3] = {(float)acc[18], (float)acc[19], (float)acc[20]};
        const float gnn = gn[0] * n[0] + gn[1] * n[1] + gn[2] * n[2];
        const float J = gj[12];
        float gc[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) gc[c] = (E * n[c] + gn[c] - gnn * n[c]) / J;
        chain_c(gj, gc, gvj);
    }
    add_vertex_grads(gvert, ti, gvi);
    add_vertex_grads(gvert, tj, gvj);
    unsafeAtomicAdd(gk_out, acc[22]);
    // conj(V_ij) l_i into (V^H lambda)_j ; V_ij = acc[23..24] / (4 pi)
    const double Vr = acc[23] * (double)kInv4Pi, Vi = acc[24] * (double)kInv4Pi;
    unsafeAtomicAdd(&vhl[2 * (int64_t)j], Vr * li.x + Vi * li.y);
    unsafeAtomicAdd(&vhl[2 * (int64_t)j + 1], Vr * li.y - Vi * li.x);
}

// ---------------------------------------------------------------------------
// potentials: lane = triangle, loop over all evaluation points
// ---------------------------------------------------------------------------
template <bool CLAMP>
__global__ __launch_bounds__(256) void adj_potential_kernel(const float* __restrict__ rec, const float* __restrict__ geo,
                                                            const int4* __restrict__ tris, int64_t nt,
                                                            const float* __restrict__ pts, int64_t np, float k,
                                                            const float2* __restrict__ pbar,
                                                            const float2* __restrict__ psi,
                                                            const float2* __restrict__ phi,
                                                            double* __restrict__ gvert, double* __restrict__ gk_out)
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
    for (int b = 0; b < Q; ++b) wy[b] = cr[3 * Q + b] * kInv4Pi;
    const float ny0 = cr[4 * Q], ny1 = cr[4 * Q + 1], ny2 = cr[4 * Q + 2];
    const float2 psj = psi[jj], phj = phi[jj];
    const float krev = k * kInv2Pi;
    float gy[Q][3];
#pragma unroll
    for (int b = 0; b < Q; ++b) gy[b][0] = gy[b][1] = gy[b][2] = 0.f;
    float gn[3] = {0.f, 0.f, 0.f}, e = 0.f, gk = 0.f;
    for (int64_t m = 0; m < np; ++m) {
        const float2 pm = pbar[m];
        // p = D phi - S psi: alpha' = -conj(pbar_m) psi_j, beta' = conj(pbar_m) phi_j
        const float ar = active ? -(pm.x * psj.x + pm.y * psj.y) : 0.f;
        const float ai = active ? -(pm.x * psj.y - pm.y * psj.x) : 0.f;
        const float br = active ? (pm.x * phj.x + pm.y * phj.y) : 0.f;
        const float bi = active ? (pm.x * phj.y - pm.y * phj.x) : 0.f;
        const float x0 = pts[3 * m], x1 = pts[3 * m + 1], x2 = pts[3 * m + 2];
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            AdjEval o;
            adj_eval<CLAMP>(py[3 * b] - x0, py[3 * b + 1] - x1, py[3 * b + 2] - x2, ny0, ny1, ny2, k, krev, ar, ai,
                            br, bi, o);
            const float w = wy[b];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                gy[b][c] += w * o.gd[c];
                gn[c] += w * o.gn[c];
            }
            e += w * o.e;
            gk += w * o.gk;
        }
    }
    if (active) {
        float gv[3][3];
        column_chain(geo + jj * 16, gy, TriRule<Q>::x1, TriRule<Q>::x2, Q, gn, e, gv);
        add_vertex_grads(gvert, tris[jj], gv);
    }
    const float gks = wsum(gk);
    if (lane == 0) unsafeAtomicAdd(gk_out, (double)gks);
}

// mass term: dL += 2 Re(-1/4 conj(l_i) phi_i) d|J_i|; and the final
// V^H lambda reduction over row tiles (fixed order) into vhl (double)
__global__ __launch_bounds__(256) void adj_mass_vhl_kernel(const float* __restrict__ geo, const int4* __restrict__ tris,
                                                           int64_t nt, const float2* __restrict__ lam,
                                                           const float2* __restrict__ phi,
                                                           const float2* __restrict__ vhl_part, int64_t nby,
                                                           double* __restrict__ gvert, double* __restrict__ vhl)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nt) return;
    const float2 li = lam[i], pi = phi[i];
    const float c = -0.5f * (li.x * pi.x + li.y * pi.y);  // 2 Re(-1/4 conj(l) phi)
    const float* g = geo + i * 16;
    float gv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const float gc[3] = {c * g[9], c * g[10], c * g[11]};
    chain_c(g, gc, gv);
    add_vertex_grads(gvert, tris[i], gv);
    double sr = 0.0, si = 0.0;
    for (int64_t b = 0; b < nby; ++b) {
        const float2 v = vhl_part[b * nt + i];
        sr += v.x;
        si += v.y;
    }
    unsafeAtomicAdd(&vhl[2 * i], sr);
    unsafeAtomicAdd(&vhl[2 * i + 1], si);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int64_t adjoint_work_bytes(const MeshDev& m)
{
    const int64_t nby = (m.nt + kAdjRowTile - 1) / kAdjRowTile;
    return (int64_t)sizeof(float2) * nby * m.nt;
}

hipError_t launch_adjoint_pairs(const MeshDev& m, float k, int clamp, const float2* lam, const float2* psi,
                                const float2* phi, double* gvert, double* gk, double* vhl, void* work, hipStream_t s)
{
    const int64_t nby = (m.nt + kAdjRowTile - 1) / kAdjRowTile;
    float2* part = (float2*)work;
    dim3 grid((unsigned)((m.nt + 255) / 256), (unsigned)nby);
    if (clamp)
        hipLaunchKernelGGL(adj_regular_kernel<true>, grid, dim3(256), 0, s, m.rec6, m.geo, m.tris, m.nt, k, lam, psi,
                           phi, gvert, gk, part);
    else
        hipLaunchKernelGGL(adj_regular_kernel<false>, grid, dim3(256), 0, s, m.rec6, m.geo, m.tris, m.nt, k, lam, psi,
                           phi, gvert, gk, part);
    const unsigned g = (unsigned)((m.npairs + 3) / 4);
    if (clamp)
        hipLaunchKernelGGL(adj_singular_kernel<true>, dim3(g), dim3(256), 0, s, m.geo, m.tris, m.pairs, m.npairs, k,
                           lam, psi, phi, gvert, gk, vhl);
    else
        hipLaunchKernelGGL(adj_singular_kernel<false>, dim3(g), dim3(256), 0, s, m.geo, m.tris, m.pairs, m.npairs, k,
                           lam, psi, phi, gvert, gk, vhl);
    hipLaunchKernelGGL(adj_mass_vhl_kernel, dim3((unsigned)((m.nt + 255) / 256)), dim3(256), 0, s, m.geo, m.tris, m.nt,
                       lam, phi, part, nby, gvert, vhl);
    return hipGetLastError();
}

hipError_t launch_adjoint_potential(const MeshDev& m, const float* pts, int64_t np, float k, int clamp,
                                    const float2* pbar, const float2* psi, const float2* phi, double* gvert, double* gk,
                                    hipStream_t s)
{
    dim3 grid((unsigned)((m.nt + 255) / 256));
    if (clamp)
        hipLaunchKernelGGL(adj_potential_kernel<true>, grid, dim3(256), 0, s, m.rec6, m.geo, m.tris, m.nt, pts, np, k,
                           pbar, psi, phi, gvert, gk);
    else
        hipLaunchKernelGGL(adj_potential_kernel<false>, grid, dim3(256), 0, s, m.rec6, m.geo, m.tris, m.nt, pts, np, k,
                           pbar, psi, phi, gvert, gk);
    return hipGetLastError();
}

}  // namespace bem
