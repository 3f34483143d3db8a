// Device-side building blocks of the BEM Helmholtz kernels for gfx950.
//
// Conventions (reference: src/bem/cuda/include/potential.h:18-65,
// integrand.h:187-235):
//   x lies on the ROW (test) triangle i, y on the COLUMN (trial) triangle j;
//   d = y - x; G = e^{ik r} / (4 pi r) with r = |d|, optionally clamped
//   r >= 1e-3 (potential.h:16,21-23).  The 1/(4 pi) factor is folded into the
//   column quadrature weights, so the kernels below return 4 pi * K.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace bem {

enum Op : int { SLP = 0, DLP = 1, ADLP = 2, HYP = 3, BM_LHS = 4, BM_RHS = 5 };
// The derivative with respect to k of operator OP is evaluated as OP | kDk
// (single, double, adjoint double, Burton-Miller LHS and RHS; P1 assembly
// and the potentials).
constexpr int kDk = 8;
constexpr int base_op(int op) { return op & 7; }
constexpr bool is_dk(int op) { return (op & kDk) != 0; }
enum Basis : int { P1 = 0, DP0 = 1 };

constexpr float kEps = 1e-3f;        // potential.h:16
constexpr float kInvEps = 1e3f;
constexpr float kEps2 = 1e-6f;
constexpr float kInv4Pi = 0.0795774715459476678844f;
constexpr float kInv2Pi = 0.159154943091895335769f;

// ---- quadrature rules ------------------------------------------------------
// Triangle rules on (0,0),(1,0),(0,1), sum of weights 1 (gauss.h:56-126):
// 6-point degree-4 Strang-Fix/Cowper, 3-point degree-2, 1-point centroid.
template <int Q> struct TriRule;
template <> struct TriRule<6> {
    static constexpr float a = 0.445948490915965f, b = 0.091576213509771f;
    static constexpr float wa = 0.223381589678011f, wb = 0.109951743655322f;
    static constexpr float x1[6] = {a, b, 1.0f - 2.0f * a, a, 1.0f - 2.0f * b, b};
    static constexpr float x2[6] = {a, b, a, 1.0f - 2.0f * a, b, 1.0f - 2.0f * b};
    static constexpr float w[6] = {wa, wb, wa, wa, wb, wb};
};
template <> struct TriRule<3> {
    static constexpr float x1[3] = {1.0f / 6, 2.0f / 3, 1.0f / 6};
    static constexpr float x2[3] = {1.0f / 6, 1.0f / 6, 2.0f / 3};
    static constexpr float w[3] = {1.0f / 3, 1.0f / 3, 1.0f / 3};
};
template <> struct TriRule<1> {
    static constexpr float x1[1] = {1.0f / 3};
    static constexpr float x2[1] = {1.0f / 3};
    static constexpr float w[1] = {1.0f};
};

// Gauss-Legendre rules mapped to [0,1]: nodes t = (s+1)/2, weights w/2
// (the reference maps inside singular_potential, integrand.h:78-81, and
// divides the product weight by 16, integrand.h:183).
template <int N> struct LineRule01;
template <> struct LineRule01<4> {
    static constexpr float t[4] = {0.5f * (1.0f + 0.861136311594053f), 0.5f * (1.0f + 0.339981043584856f),
                                   0.5f * (1.0f - 0.339981043584856f), 0.5f * (1.0f - 0.861136311594053f)};
    static constexpr float w[4] = {0.5f * 0.347854845137454f, 0.5f * 0.652145154862546f,
                                   0.5f * 0.652145154862546f, 0.5f * 0.347854845137454f};
};
template <> struct LineRule01<2> {
    static constexpr float t[2] = {0.5f * (1.0f + 0.577350269189626f), 0.5f * (1.0f - 0.577350269189626f)};
    static constexpr float w[2] = {0.5f, 0.5f};
};

// ---- per-triangle records ----------------------------------------------------
// Quadrature record of one triangle for rule Q, float4-aligned:
//   pts[3Q]   quadrature points (x,y,z interleaved)
//   wts[Q]    w_q * |J| / 2   (the area-scaled weights)
//   n[3]      unit normal of the vertex order (v0, v1, v2)
//   pad to a multiple of 4 floats
template <int Q> struct Rec {
    static constexpr int kFloats = ((4 * Q + 3) + 3) / 4 * 4;
};

// ---- complex helpers ---------------------------------------------------------
struct cf {
    float r, i;
};

// sin(2 pi t), cos(2 pi t) for t in revolutions, on the hardware
// transcendental unit after exact range reduction to [0, 1).
// Cross-lane sums in every lane of a wavefront.  Within each 16-lane row the
// values move by DPP (quad exchanges [1 0 3 2] and [2 3 0 1], the half-row
// mirror i <-> 7 - i, the row mirror i <-> 15 - i), across rows by two
// permutes (xor 16, xor 32).  Each step adds the same two values in every
// lane, so every lane ends with the same sum.
template <int CTRL>
__device__ __forceinline__ float dpp_row_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_row_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ float wave_sum64(float v)
{
    v += dpp_row_f<0xB1>(v);
    v += dpp_row_f<0x4E>(v);
    v += dpp_row_f<0x141>(v);
    v += dpp_row_f<0x140>(v);
    v += __shfl_xor(v, 16, 64);
    return v + __shfl_xor(v, 32, 64);
}
__device__ __forceinline__ double wave_sum64(double v)
{
    v += dpp_row_d<0xB1>(v);
    v += dpp_row_d<0x4E>(v);
    v += dpp_row_d<0x141>(v);
    v += dpp_row_d<0x140>(v);
    v += __shfl_xor(v, 16, 64);
    return v + __shfl_xor(v, 32, 64);
}
// the same over each 16-lane row alone
__device__ __forceinline__ double row_sum16(double v)
{
    v += dpp_row_d<0xB1>(v);
    v += dpp_row_d<0x4E>(v);
    v += dpp_row_d<0x141>(v);
    return v + dpp_row_d<0x140>(v);
}

__device__ __forceinline__ void sincos_rev(float t, float& s, float& c)
{
    float f = __builtin_amdgcn_fractf(t);
    s = __builtin_amdgcn_sinf(f);
    c = __builtin_amdgcn_cosf(f);
}

// Kernel values for a displacement d = y - x.  Outputs are 4 pi times the
// reference's potential<OP> (potential.h:18-52) and, for BM_LHS, 4 pi times
// the second output of bm_lhs_potential (potential.h:55-65).  HYP evaluates
// the single-layer kernel (its curl form is applied per pair).  OP | kDk:
// the derivatives of the same values with respect to k (with G = e^{ikr}/r,
// dG/dk = i e^{ikr} and d/dk [G (1 - ikr)] = k r e^{ikr}); for BM_LHS | kDk
// the second output is H with d/dk [(i/k) G] = (i/k) H, H = i e^{ikr} - G/k,
// so that the caller's (i/k) scaling of the second output stays as it is.
//   nx: row normal (x side), ny: column normal (y side)
//   k: wavenumber, krev = k / (2 pi)
template <int OP, bool CLAMP>
__device__ __forceinline__ void kernel4pi(float dx, float dy, float dz, float nx0, float nx1, float nx2, float ny0,
                                          float ny1, float ny2, float k, float krev, float& vr, float& vi, float& sr,
                                          float& si)
{
    float r2 = dx * dx + dy * dy + dz * dz;
    if (CLAMP) r2 = fmaxf(r2, kEps2);  // max(r, eps) = sqrt(max(r^2, eps^2))
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;
    float s, c;
    sincos_rev(r * krev, s, c);
    if constexpr (OP == SLP || OP == HYP) {
        vr = c * rinv;
        vi = s * rinv;
    } else if constexpr (OP == DLP || OP == ADLP) {
        // DLP: -G (1 - i k r) (y-x).n_y / r^2 ; ADLP: -G (1 - i k r) (x-y).n_x / r^2
        float dn = (OP == DLP) ? -(dx * ny0 + dy * ny1 + dz * ny2) : (dx * nx0 + dy * nx1 + dz * nx2);
        float th = k * r;
        float f = dn * rinv * rinv * rinv;
        vr = f * fmaf(th, s, c);
        vi = f * fmaf(-th, c, s);
    } else if constexpr (OP == BM_LHS) {
        // (-i k nx.ny + (1 - i k r)(y-x).n_y / r^2) G ; second: (i/k) G
        float a = (dx * ny0 + dy * ny1 + dz * ny2) * rinv * rinv;
        float th = k * r;
        float ai = -k * (nx0 * ny0 + nx1 * ny1 + nx2 * ny2) - th * a;
        float gr = c * rinv, gi = s * rinv;
        vr = a * gr - ai * gi;
        vi = a * gi + ai * gr;
        sr = gr;  // plain G; scaled by i/k after the sum
        si = gi;
    } else if constexpr (OP == BM_RHS) {
        // -G + (i/k) G (1 - i k r) (x-y).n_x / r^2
        float q = -(dx * nx0 + dy * nx1 + dz * nx2) * rinv * rinv;  // (x-y).n_x / r^2
        float th = k * r;
        float gr = c * rinv, gi = s * rinv;
        float invk = 1.0f / k;
        // (i/k)(1 - i th) = th/k + i/k
        float tr = th * invk * q, ti = invk * q;
        vr = -gr + (gr * tr - gi * ti);
        vi = -gi + (gr * ti + gi * tr);
    } else if constexpr (OP == (SLP | kDk)) {
        vr = -s;  // i e^{ikr}
        vi = c;
    } else if constexpr (OP == (DLP | kDk) || OP == (ADLP | kDk)) {
        // the double-layer kernel's e^{ikr} (1 - ikr) / r^3 becomes k e^{ikr} / r
        float dn = (OP == (DLP | kDk)) ? -(dx * ny0 + dy * ny1 + dz * ny2) : (dx * nx0 + dy * nx1 + dz * nx2);
        float f = dn * k * rinv;
        vr = f * c;
        vi = f * s;
    } else if constexpr (OP == (BM_LHS | kDk)) {
        // d/dk of G [(1 - ikr) a - ik nx.ny], a = (y-x).n_y / r^2:
        // a k r e^{ikr} + k nx.ny e^{ikr} - i nx.ny G
        float a = (dx * ny0 + dy * ny1 + dz * ny2) * rinv * rinv;
        float nn = nx0 * ny0 + nx1 * ny1 + nx2 * ny2;
        float t = k * (a * r + nn);
        vr = fmaf(t, c, nn * s * rinv);
        vi = fmaf(t, s, -nn * c * rinv);
        float invk = 1.0f / k;
        sr = -fmaf(c * rinv, invk, s);  // H = i e^{ikr} - G / k
        si = fmaf(-s * rinv, invk, c);
    } else if constexpr (OP == (BM_RHS | kDk)) {
        // d/dk of -G + (i/k) G (1 - ikr) q, q = (x-y).n_x / r^2:
        // -i e^{ikr} + q [-(i/k^2) G (1 - ikr) + i r e^{ikr}]
        float q = -(dx * nx0 + dy * nx1 + dz * nx2) * rinv * rinv;
        float th = k * r;
        float invk = 1.0f / k;
        float u = rinv * invk * invk;  // 1 / (k^2 r)
        vr = s + q * (fmaf(-th, c, s) * u - r * s);
        vi = -c + q * (r * c - fmaf(th, s, c) * u);
    }
}

// Surface-curl product of two triangles' P1 hat functions, A B^T of the
// rows n x grad(phi_a) (common.h:150-164 applied to solver.py:40-58 output).
__device__ __forceinline__ void curl_product(const float* ci, const float* cj, float* cp)
{
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            cp[3 * a + b] = ci[3 * a] * cj[3 * b] + ci[3 * a + 1] * cj[3 * b + 1] + ci[3 * a + 2] * cj[3 * b + 2];
}

// ---- Sauter-Schwab regions of a neighbour pair ------------------------------
// A pair with S shared vertices (3 coincident, 2 edge, 1 vertex; the shared
// ones at local 0 (and 1), integrand.h:117-153) on the reference triangle
// (0,0),(1,0),(1,1): at the tensor point (xs, e_1, e_2, e_3) of [0,1]^4 each
// of the kSSRegions<S> regions gives a row (x1, x2, y1, y2, extra weight, d1,
// d2) with (d1, d2) = (y1 - x1, y2 - x2) formed from the eta products directly
// (coincident; edge: d1 only), which keeps y - x accurate where the points
// nearly coincide; the return value is the Duffy Jacobian.  S is a template
// argument so that callers unroll the region loop and keep the table in
// registers (indexed by a run-time region count it lands in scratch memory).
template <int S> constexpr int kSSRegions = S == 3 ? 6 : (S == 2 ? 5 : 2);

template <int S>
__device__ __forceinline__ float ss_regions(float xs, float e_1, float e_2, float e_3,
                                            float (&reg)[kSSRegions<S>][7])
{
    const float e12 = e_1 * e_2, e123 = e12 * e_3;
    if constexpr (S == 3) {
        const float e12c = e12 * (1.f - e_3);  // eta1 eta2 - eta1 eta2 eta3
        const float t[6][7] = {
            {xs, xs * (1.f - e_1 + e12), xs * (1.f - e123), xs * (1.f - e_1), 1.f, -xs * e123, -xs * e12},
            {xs * (1.f - e123), xs * (1.f - e_1), xs, xs * (1.f - e_1 + e12), 1.f, xs * e123, xs * e12},
            {xs, xs * (e_1 - e12 + e123), xs * (1.f - e12), xs * (e_1 - e12), 1.f, -xs * e12, -xs * e123},
            {xs * (1.f - e12), xs * (e_1 - e12), xs, xs * (e_1 - e12 + e123), 1.f, xs * e12, xs * e123},
            {xs * (1.f - e123), xs * (e_1 - e123), xs, xs * (e_1 - e12), 1.f, xs * e123, -xs * e12c},
            {xs, xs * (e_1 - e12), xs * (1.f - e123), xs * (e_1 - e123), 1.f, -xs * e123, xs * e12c},
        };
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = 0; c < 7; ++c) reg[r][c] = t[r][c];
        return xs * xs * xs * e_1 * e_1 * e_2;
    } else if constexpr (S == 2) {
        const float t[5][7] = {
            {xs, xs * e_1 * e_3, xs * (1.f - e12), xs * e_1 * (1.f - e_2), 1.f, -xs * e12, 0.f},
            {xs, xs * e_1, xs * (1.f - e123), xs * e12 * (1.f - e_3), e_2, -xs * e123, 0.f},
            {xs * (1.f - e12), xs * e_1 * (1.f - e_2), xs, xs * e123, e_2, xs * e12, 0.f},
            {xs * (1.f - e123), xs * e12 * (1.f - e_3), xs, xs * e_1, e_2, xs * e123, 0.f},
            {xs * (1.f - e123), xs * e_1 * (1.f - e_2 * e_3), xs, xs * e12, e_2, xs * e123, 0.f},
        };
#pragma unroll
        for (int r = 0; r < 5; ++r)
#pragma unroll
            for (int c = 0; c < 7; ++c) reg[r][c] = t[r][c];
        return xs * xs * xs * e_1 * e_1;
    } else {
        const float t[2][7] = {
            {xs, xs * e_1, xs * e_2, xs * e_2 * e_3, 1.f, 0.f, 0.f},
            {xs * e_2, xs * e_2 * e_3, xs, xs * e_1, 1.f, 0.f, 0.f},
        };
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 7; ++c) reg[r][c] = t[r][c];
        return xs * xs * xs * e_2;
    }
}

// y - x for a region row, with x = v0 + x1 e1 + x2 e2 and y = w0 + y1 f1 + y2 f2
// (e1 = v1 - v0, e2 = v2 - v1, f likewise on the column triangle); the shared
// local vertex 0 (v0 = w0) cancels, and for an edge pair f1 = e1.
template <int S>
__device__ __forceinline__ void ss_disp(const float (&row)[7], const float* e1, const float* e2, const float* f1,
                                        const float* f2, float& dx, float& dy, float& dz)
{
    const float x1 = row[0], x2 = row[1], y1 = row[2], y2 = row[3], d1 = row[5], d2 = row[6];
    if constexpr (S == 3) {  // same triangle: y - x = d1 e1 + d2 e2
        dx = d1 * e1[0] + d2 * e2[0];
        dy = d1 * e1[1] + d2 * e2[1];
        dz = d1 * e1[2] + d2 * e2[2];
    } else if constexpr (S == 2) {  // common edge: y - x = d1 e1 + y2 f2 - x2 e2
        dx = d1 * e1[0] + y2 * f2[0] - x2 * e2[0];
        dy = d1 * e1[1] + y2 * f2[1] - x2 * e2[1];
        dz = d1 * e1[2] + y2 * f2[2] - x2 * e2[2];
    } else {  // common vertex
        dx = y1 * f1[0] + y2 * f2[0] - x1 * e1[0] - x2 * e2[0];
        dy = y1 * f1[1] + y2 * f2[1] - x1 * e1[1] - x2 * e2[1];
        dz = y1 * f1[2] + y2 * f2[2] - x1 * e1[2] - x2 * e2[2];
    }
}

// f(std::integral_constant<int, S>) for the pair's shared-vertex count
template <class F>
__device__ __forceinline__ void ss_dispatch(int shared, F&& f)
{
    if (shared == 3) f(std::integral_constant<int, 3>{});
    else if (shared == 2) f(std::integral_constant<int, 2>{});
    else f(std::integral_constant<int, 1>{});
}

}  // namespace bem
