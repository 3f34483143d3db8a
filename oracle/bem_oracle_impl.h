/*
 * Precision-generic body of the CPU oracle.  Included twice by bem_oracle.c:
 * once with R = double (the FP64 oracle) and once with R = float (the FP32
 * mirror of the reference's single-precision arithmetic).  TEST
 * INFRASTRUCTURE ONLY: see the header of bem_oracle.c.
 *
 * Required macros: R (scalar type), FN(name) (symbol suffixing), KERNEL_EVAL
 * (pointwise kernel for this precision), IS_F32 (0/1).
 */

/* ---- geometry helpers (restate integrand.h:14-61) ---------------------- */

static inline void FN(sub3)(const R *a, const R *b, R *o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static inline R FN(dot3)(const R *a, const R *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void FN(cross3)(const R *a, const R *b, R *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* Twice the triangle area (integrand.h:30-43) and the unit normal of the
 * vertex order v0,v1,v2 (integrand.h:14-28). */
static void FN(tri_geom)(const R *v0, const R *v1, const R *v2, R *jac, R *nrm)
{
    R e1[3], e2[3], c[3];
    FN(sub3)(v1, v0, e1);
    FN(sub3)(v2, v0, e2);
    FN(cross3)(e1, e2, c);
    R l = (R)sqrt((double)FN(dot3)(c, c));
    if (IS_F32) l = sqrtf((float)FN(dot3)(c, c));
    *jac = l;
    nrm[0] = c[0] / l; nrm[1] = c[1] / l; nrm[2] = c[2] / l;
}

/* Point on triangle (0,0),(1,0),(0,1) with shape values (integrand.h:45-52). */
static inline void FN(map_std)(R x1, R x2, const R *v0, const R *v1, const R *v2, R *p, R *phi)
{
    phi[0] = 1 - x1 - x2; phi[1] = x1; phi[2] = x2;
    for (int d = 0; d < 3; ++d) p[d] = phi[0] * v0[d] + x1 * v1[d] + x2 * v2[d];
}

/* Point on triangle (0,0),(1,0),(1,1) with shape values (integrand.h:54-61). */
static inline void FN(map_ss)(R x1, R x2, const R *v0, const R *v1, const R *v2, R *p, R *phi)
{
    phi[0] = 1 - x1; phi[1] = x1 - x2; phi[2] = x2;
    for (int d = 0; d < 3; ++d) p[d] = phi[0] * v0[d] + phi[1] * v1[d] + x2 * v2[d];
}

/* Surface curls n x grad(phi_a) of the three P1 hat functions of a
 * triangle, row a (restates solver.py:40-58). */
static void FN(tri_curls)(const R *v0, const R *v1, const R *v2, const R *n, R curl[9])
{
    R e1[3], e2[3];
    FN(sub3)(v1, v0, e1);
    FN(sub3)(v2, v0, e2);
    double g11 = FN(dot3)(e1, e1), g12 = FN(dot3)(e1, e2), g22 = FN(dot3)(e2, e2);
    double det = g11 * g22 - g12 * g12;
    double i11 = g22 / det, i12 = -g12 / det, i22 = g11 / det;
    /* columns of J (J^T J)^{-1}: dual basis vectors b1, b2 */
    R b1[3], b2[3];
    for (int d = 0; d < 3; ++d) {
        b1[d] = (R)(e1[d] * i11 + e2[d] * i12);
        b2[d] = (R)(e1[d] * i12 + e2[d] * i22);
    }
    R grad[3][3];
    for (int d = 0; d < 3; ++d) {
        grad[0][d] = -b1[d] - b2[d];
        grad[1][d] = b1[d];
        grad[2][d] = b2[d];
    }
    for (int a = 0; a < 3; ++a) FN(cross3)(n, grad[a], &curl[3 * a]);
}

/* Per-triangle data used by every pair routine. */
typedef struct {
    R v[3][3];
    R jac;
    R n[3];
    R curl[9];
    int32_t idx[3];
} FN(tri_t);

static FN(tri_t) *FN(make_tris)(const R *verts, const int32_t *tris, int64_t nt)
{
    FN(tri_t) *T = (FN(tri_t) *)malloc(sizeof(FN(tri_t)) * (size_t)(nt > 0 ? nt : 1));
    for (int64_t t = 0; t < nt; ++t) {
        for (int a = 0; a < 3; ++a) {
            T[t].idx[a] = tris[3 * t + a];
            for (int d = 0; d < 3; ++d) T[t].v[a][d] = verts[3 * (int64_t)tris[3 * t + a] + d];
        }
        FN(tri_geom)(T[t].v[0], T[t].v[1], T[t].v[2], &T[t].jac, T[t].n);
        FN(tri_curls)(T[t].v[0], T[t].v[1], T[t].v[2], T[t].n, T[t].curl);
    }
    return T;
}

/* Accumulator for one triangle pair: 3x3 basis-weighted block + the
 * un-weighted sum (the DP0 entry; secondary kernel for BM_LHS). */
typedef struct { R re[10], im[10]; } FN(blk_t);

static inline void FN(blk_add)(FN(blk_t) *B, R vr, R vi, R sr, R si, const R *px, const R *py, int op)
{
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            /* (value * phi_x[a]) * phi_y[b], as integrand.h:225-228 */
            R tr = vr * px[a], ti = vi * px[a];
            B->re[3 * a + b] += tr * py[b];
            B->im[3 * a + b] += ti * py[b];
        }
    if (op == ORACLE_BM_LHS) { B->re[9] += sr; B->im[9] += si; }
    else { B->re[9] += vr; B->im[9] += vi; }
}

/* Regular pair (integrand.h:187-235, 280-329): tensor triangle rule. */
static void FN(pair_regular)(const FN(tri_t) *Ti, const FN(tri_t) *Tj, R k, int op, const R *rule, int nq,
                             int clamp, FN(blk_t) *B)
{
    memset(B, 0, sizeof(*B));
    int kop = (op == ORACLE_HYP) ? ORACLE_SLP : op;
    R w1 = (R)(0.25 * (double)Ti->jac * (double)Tj->jac);
    for (int a = 0; a < nq; ++a) {
        R x[3], px[3];
        FN(map_std)(rule[3 * a], rule[3 * a + 1], Ti->v[0], Ti->v[1], Ti->v[2], x, px);
        for (int b = 0; b < nq; ++b) {
            R y[3], py[3];
            FN(map_std)(rule[3 * b], rule[3 * b + 1], Tj->v[0], Tj->v[1], Tj->v[2], y, py);
            R w2 = w1 * rule[3 * a + 2] * rule[3 * b + 2];
            R vr, vi, sr = 0, si = 0;
            KERNEL_EVAL(kop, x, y, Ti->n, Tj->n, k, clamp, &vr, &vi, &sr, &si);
            vr *= w2; vi *= w2; sr *= w2; si *= w2;
            FN(blk_add)(B, vr, vi, sr, si, px, py, op);
        }
    }
}

/* Sauter-Schwab singular pair (integrand.h:63-185, 331-394).  `shared` is
 * the number of common vertices (3 coincident, 2 edge, 1 vertex). */
static void FN(pair_singular)(const FN(tri_t) *Ti, const FN(tri_t) *Tj, R k, int op, const R *line, int nl,
                              int clamp, FN(blk_t) *B)
{
    memset(B, 0, sizeof(*B));
    int kop = (op == ORACLE_HYP) ? ORACLE_SLP : op;
    int shared = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) shared += (Ti->idx[a] == Tj->idx[b]);
    int pi[3] = {0, 1, 2}, pj[3] = {0, 1, 2};
    if (shared == 2) {
        /* shared edge first, in the same order on both triangles, ordered by
         * position in the column triangle (integrand.h:351-364) */
        int c = 0;
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a)
                if (Ti->idx[a] == Tj->idx[b]) { pi[c] = a; pj[c] = b; ++c; }
        pi[2] = 3 - pi[0] - pi[1];
        pj[2] = 3 - pj[0] - pj[1];
    } else if (shared == 1) {
        /* shared vertex moved to local 0 by a swap (integrand.h:365-382) */
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                if (Ti->idx[a] == Tj->idx[b]) {
                    if (a != 0) { pi[0] = a; pi[a] = 0; }
                    if (b != 0) { pj[0] = b; pj[b] = 0; }
                }
    }
    const R *xi_v[3] = {Ti->v[pi[0]], Ti->v[pi[1]], Ti->v[pi[2]]};
    const R *yj_v[3] = {Tj->v[pj[0]], Tj->v[pj[1]], Tj->v[pj[2]]};
    FN(blk_t) L;
    memset(&L, 0, sizeof(L));
    for (int i0 = 0; i0 < nl; ++i0)
        for (int i1 = 0; i1 < nl; ++i1)
            for (int i2 = 0; i2 < nl; ++i2)
                for (int i3 = 0; i3 < nl; ++i3) {
                    /* one tensor Gauss point; its own block is scaled then summed,
                     * as the reference sums per-point results */
                    FN(blk_t) P;
                    memset(&P, 0, sizeof(P));
                    R weight = (R)((double)(line[2 * i0 + 1] * line[2 * i1 + 1] * line[2 * i2 + 1] * line[2 * i3 + 1]) *
                                   (double)Ti->jac * (double)Tj->jac / 16.0);
                    if (IS_F32)
                        weight = line[2 * i0 + 1] * line[2 * i1 + 1] * line[2 * i2 + 1] * line[2 * i3 + 1] * Ti->jac *
                                 Tj->jac / 16;
                    R xs = (R)(0.5 * ((double)line[2 * i0] + 1.0));
                    R e1 = (R)(0.5 * ((double)line[2 * i1] + 1.0));
                    R e2 = (R)(0.5 * ((double)line[2 * i2] + 1.0));
                    R e3 = (R)(0.5 * ((double)line[2 * i3] + 1.0));
                    R e12 = e1 * e2, e123 = e1 * e2 * e3;
                    /* region table: (x1, x2, y1, y2, extra factor) */
                    R reg[6][5];
                    int nreg = 0;
                    R jacw = 0;
                    if (shared == 3) {
                        jacw = xs * xs * xs * e1 * e1 * e2;
                        R t[6][4] = {
                            {xs, xs * (1 - e1 + e12), xs * (1 - e123), xs * (1 - e1)},
                            {xs * (1 - e123), xs * (1 - e1), xs, xs * (1 - e1 + e12)},
                            {xs, xs * (e1 - e12 + e123), xs * (1 - e12), xs * (e1 - e12)},
                            {xs * (1 - e12), xs * (e1 - e12), xs, xs * (e1 - e12 + e123)},
                            {xs * (1 - e123), xs * (e1 - e123), xs, xs * (e1 - e12)},
                            {xs, xs * (e1 - e12), xs * (1 - e123), xs * (e1 - e123)},
                        };
                        nreg = 6;
                        for (int r = 0; r < 6; ++r) { for (int c = 0; c < 4; ++c) reg[r][c] = t[r][c]; reg[r][4] = 1; }
                    } else if (shared == 2) {
                        jacw = xs * xs * xs * e1 * e1;
                        R t[5][5] = {
                            {xs, xs * e1 * e3, xs * (1 - e12), xs * e1 * (1 - e2), 1},
                            {xs, xs * e1, xs * (1 - e123), xs * e1 * e2 * (1 - e3), e2},
                            {xs * (1 - e12), xs * e1 * (1 - e2), xs, xs * e123, e2},
                            {xs * (1 - e123), xs * e12 * (1 - e3), xs, xs * e1, e2},
                            {xs * (1 - e123), xs * e1 * (1 - e2 * e3), xs, xs * e12, e2},
                        };
                        nreg = 5;
                        for (int r = 0; r < 5; ++r) for (int c = 0; c < 5; ++c) reg[r][c] = t[r][c];
                    } else if (shared == 1) {
                        jacw = xs * xs * xs * e2;
                        R t[2][4] = {
                            {xs, xs * e1, xs * e2, xs * e2 * e3},
                            {xs * e2, xs * e2 * e3, xs, xs * e1},
                        };
                        nreg = 2;
                        for (int r = 0; r < 2; ++r) { for (int c = 0; c < 4; ++c) reg[r][c] = t[r][c]; reg[r][4] = 1; }
                    }
                    for (int r = 0; r < nreg; ++r) {
                        R x[3], px[3], y[3], py[3];
                        FN(map_ss)(reg[r][0], reg[r][1], xi_v[0], xi_v[1], xi_v[2], x, px);
                        FN(map_ss)(reg[r][2], reg[r][3], yj_v[0], yj_v[1], yj_v[2], y, py);
                        R vr, vi, sr = 0, si = 0;
                        KERNEL_EVAL(kop, x, y, Ti->n, Tj->n, k, clamp, &vr, &vi, &sr, &si);
                        R ww = reg[r][4];
                        vr *= ww; vi *= ww; sr *= ww; si *= ww;
                        FN(blk_add)(&P, vr, vi, sr, si, px, py, op);
                    }
                    R s = jacw * weight;
                    for (int e = 0; e < 10; ++e) { L.re[e] += P.re[e] * s; L.im[e] += P.im[e] * s; }
                }
    /* undo the vertex reordering (integrand.h:390-393) */
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            B->re[3 * pi[a] + pj[b]] = L.re[3 * a + b];
            B->im[3 * pi[a] + pj[b]] = L.im[3 * a + b];
        }
    B->re[9] = L.re[9];
    B->im[9] = L.im[9];
}

/* HYP / BM_LHS post-combination with surface curls (integrand.h:304-323,
 * bind.cu:127-143). */
static void FN(post_combine)(const FN(tri_t) *Ti, const FN(tri_t) *Tj, R k, int op, FN(blk_t) *B)
{
    if (op != ORACLE_HYP && op != ORACLE_BM_LHS) return;
    R np = FN(dot3)(Ti->n, Tj->n);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            R cp = 0;
            for (int c = 0; c < 3; ++c) cp += Ti->curl[3 * a + c] * Tj->curl[3 * b + c];
            R r = B->re[3 * a + b], i = B->im[3 * a + b];
            if (op == ORACLE_HYP) {
                B->re[3 * a + b] = B->re[9] * cp - r * np * k * k;
                B->im[3 * a + b] = B->im[9] * cp - i * np * k * k;
            } else {
                B->re[3 * a + b] = B->re[9] * cp + r;
                B->im[3 * a + b] = B->im[9] * cp + i;
            }
        }
}

/* Assemble an operator (see bem_oracle.c for the argument meaning). */
int FN(oracle_assemble)(const R *verts, int64_t nv, const int32_t *tris, int64_t nt, R k, int op, int basis,
                        const R *tri_rule, int nq, const R *line_rule, int nl, int clamp, int64_t row_begin,
                        int64_t row_end, int parts, R *out, int nthreads)
{
    if (nv <= 0 || nt <= 0 || row_begin < 0 || row_end > nt || row_begin > row_end) return ORACLE_BAD_ARG;
    if (basis == ORACLE_DP0 && (op == ORACLE_HYP || op == ORACLE_BM_LHS)) return ORACLE_BAD_ARG;
    for (int64_t t = 0; t < 3 * nt; ++t)
        if (tris[t] < 0 || tris[t] >= nv) return ORACLE_BAD_MESH;
    FN(tri_t) *T = FN(make_tris)(verts, tris, nt);
    adj_t A;
    if (build_adjacency(tris, nt, nv, &A) != 0) { free(T); return ORACLE_BAD_MESH; }
    int64_t nrows = row_end - row_begin;
    if (basis == ORACLE_DP0) memset(out, 0, sizeof(R) * 2 * (size_t)nrows * (size_t)nt);
    else memset(out, 0, sizeof(R) * 2 * (size_t)nv * (size_t)nv);
    if (nthreads > 0) omp_set_num_threads(nthreads);

    /* P1: row triangles of one colour share no vertex, so their scatters hit
     * disjoint rows and the result does not depend on the thread count. */
    for (int col = 0; col < (basis == ORACLE_P1 ? A.ncolors : 1); ++col) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t i = row_begin; i < row_end; ++i) {
            if (basis == ORACLE_P1 && A.color[i] != col) continue;
            FN(blk_t) B;
            const FN(tri_t) *Ti = &T[i];
            if (parts & 1) {
                for (int64_t j = 0; j < nt; ++j) {
                    const FN(tri_t) *Tj = &T[j];
                    int sh = 0;
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) sh += (Ti->idx[a] == Tj->idx[b]);
                    if (sh) continue;
                    FN(pair_regular)(Ti, Tj, k, op, tri_rule, nq, clamp, &B);
                    FN(post_combine)(Ti, Tj, k, op, &B);
                    if (basis == ORACLE_DP0) {
                        out[2 * ((i - row_begin) * nt + j)] = B.re[9];
                        out[2 * ((i - row_begin) * nt + j) + 1] = B.im[9];
                    } else {
                        for (int a = 0; a < 3; ++a)
                            for (int b = 0; b < 3; ++b) {
                                int64_t o = 2 * ((int64_t)Ti->idx[a] * nv + Tj->idx[b]);
                                out[o] += B.re[3 * a + b];
                                out[o + 1] += B.im[3 * a + b];
                            }
                    }
                }
            }
            if (parts & 2) {
                for (int64_t e = A.nb_ptr[i]; e < A.nb_ptr[i + 1]; ++e) {
                    int64_t j = A.nb_idx[e];
                    const FN(tri_t) *Tj = &T[j];
                    FN(pair_singular)(Ti, Tj, k, op, line_rule, nl, clamp, &B);
                    FN(post_combine)(Ti, Tj, k, op, &B);
                    if (basis == ORACLE_DP0) {
                        out[2 * ((i - row_begin) * nt + j)] = B.re[9];
                        out[2 * ((i - row_begin) * nt + j) + 1] = B.im[9];
                    } else {
                        for (int a = 0; a < 3; ++a)
                            for (int b = 0; b < 3; ++b) {
                                int64_t o = 2 * ((int64_t)Ti->idx[a] * nv + Tj->idx[b]);
                                out[o] += B.re[3 * a + b];
                                out[o + 1] += B.im[3 * a + b];
                            }
                    }
                }
            }
        }
    }
    free_adjacency(&A);
    free(T);
    return ORACLE_OK;
}

/* DP0 entries at listed (row, column) triangle pairs: out[2e], out[2e+1] =
 * entry (rows[e], cols[e]) of the same matrix oracle_assemble builds for the
 * DP0 basis (the pair is regular when the triangles share no vertex,
 * bind.cu:67-79, else Sauter-Schwab, bind.cu:81-152; the entry is the
 * unweighted sum result[9] of integrand.h:234,392).  For checks of a few
 * rows or columns of a large mesh without assembling it whole. */
int FN(oracle_entries)(const R *verts, int64_t nv, const int32_t *tris, int64_t nt, R k, int op,
                       const R *tri_rule, int nq, const R *line_rule, int nl, int clamp, const int32_t *rows,
                       const int32_t *cols, int64_t n, R *out, int nthreads)
{
    if (nv <= 0 || nt <= 0 || n < 0) return ORACLE_BAD_ARG;
    if (op == ORACLE_HYP || op == ORACLE_BM_LHS) return ORACLE_BAD_ARG;
    for (int64_t t = 0; t < 3 * nt; ++t)
        if (tris[t] < 0 || tris[t] >= nv) return ORACLE_BAD_MESH;
    for (int64_t e = 0; e < n; ++e)
        if (rows[e] < 0 || rows[e] >= nt || cols[e] < 0 || cols[e] >= nt) return ORACLE_BAD_ARG;
    FN(tri_t) *T = FN(make_tris)(verts, tris, nt);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t e = 0; e < n; ++e) {
        const FN(tri_t) *Ti = &T[rows[e]], *Tj = &T[cols[e]];
        int sh = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) sh += (Ti->idx[a] == Tj->idx[b]);
        FN(blk_t) B;
        if (sh) FN(pair_singular)(Ti, Tj, k, op, line_rule, nl, clamp, &B);
        else FN(pair_regular)(Ti, Tj, k, op, tri_rule, nq, clamp, &B);
        FN(post_combine)(Ti, Tj, k, op, &B);
        out[2 * e] = B.re[9];
        out[2 * e + 1] = B.im[9];
    }
    free(T);
    return ORACLE_OK;
}

/* P1 rows of listed vertices: out [nrows, nv] complex = rows row_vertices[e]
 * of the matrix oracle_assemble builds for the P1 basis.  Row a of the P1
 * matrix only receives the blocks of pairs whose row triangle holds vertex a
 * (bind.cu:67-79,144-151 scatter B[alpha][beta] to (T_i.alpha, T_j.beta)), so
 * each row is the sum over the triangles around a of their pairs with every
 * column triangle.  For checks of a few rows of a large mesh. */
int FN(oracle_p1_rows)(const R *verts, int64_t nv, const int32_t *tris, int64_t nt, R k, int op,
                       const R *tri_rule, int nq, const R *line_rule, int nl, int clamp, const int32_t *row_vertices,
                       int64_t nrows, R *out, int nthreads)
{
    if (nv <= 0 || nt <= 0 || nrows < 0) return ORACLE_BAD_ARG;
    for (int64_t t = 0; t < 3 * nt; ++t)
        if (tris[t] < 0 || tris[t] >= nv) return ORACLE_BAD_MESH;
    for (int64_t e = 0; e < nrows; ++e)
        if (row_vertices[e] < 0 || row_vertices[e] >= nv) return ORACLE_BAD_ARG;
    FN(tri_t) *T = FN(make_tris)(verts, tris, nt);
    memset(out, 0, sizeof(R) * 2 * (size_t)nrows * (size_t)nv);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    /* one thread per listed row: its sum runs over its own triangles in
     * ascending order, so the result does not depend on the thread count */
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t e = 0; e < nrows; ++e) {
        int32_t a = row_vertices[e];
        R *row = out + 2 * e * nv;
        for (int64_t i = 0; i < nt; ++i) {
            const FN(tri_t) *Ti = &T[i];
            int alpha = -1;
            for (int c = 0; c < 3; ++c)
                if (Ti->idx[c] == a) alpha = c;
            if (alpha < 0) continue;
            for (int64_t j = 0; j < nt; ++j) {
                const FN(tri_t) *Tj = &T[j];
                int sh = 0;
                for (int c = 0; c < 3; ++c)
                    for (int d = 0; d < 3; ++d) sh += (Ti->idx[c] == Tj->idx[d]);
                FN(blk_t) B;
                if (sh) FN(pair_singular)(Ti, Tj, k, op, line_rule, nl, clamp, &B);
                else FN(pair_regular)(Ti, Tj, k, op, tri_rule, nq, clamp, &B);
                FN(post_combine)(Ti, Tj, k, op, &B);
                for (int beta = 0; beta < 3; ++beta) {
                    row[2 * Tj->idx[beta]] += B.re[3 * alpha + beta];
                    row[2 * Tj->idx[beta] + 1] += B.im[3 * alpha + beta];
                }
            }
        }
    }
    free(T);
    return ORACLE_OK;
}

/* Potential operator from the surface to points (integrand.h:237-255,
 * 396-406; bind.cu:156-176).  out: [np, nv] (P1) or [np, nt] (DP0). */
int FN(oracle_potential)(const R *verts, int64_t nv, const int32_t *tris, int64_t nt, const R *pts, int64_t np, R k,
                         int op, int basis, const R *tri_rule, int nq, int clamp, R *out, int nthreads)
{
    if (nv <= 0 || nt <= 0 || np < 0) return ORACLE_BAD_ARG;
    if (op != ORACLE_SLP && op != ORACLE_DLP) return ORACLE_BAD_ARG;
    for (int64_t t = 0; t < 3 * nt; ++t)
        if (tris[t] < 0 || tris[t] >= nv) return ORACLE_BAD_MESH;
    FN(tri_t) *T = FN(make_tris)(verts, tris, nt);
    int64_t ncol = (basis == ORACLE_DP0) ? nt : nv;
    memset(out, 0, sizeof(R) * 2 * (size_t)np * (size_t)ncol);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    R zero[3] = {0, 0, 0};
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < np; ++p) {
        for (int64_t t = 0; t < nt; ++t) {
            R acc_r[3] = {0, 0, 0}, acc_i[3] = {0, 0, 0};
            R s_r = 0, s_i = 0;
            R w1 = (R)(0.5 * (double)T[t].jac);
            for (int a = 0; a < nq; ++a) {
                R y[3], py[3];
                FN(map_std)(tri_rule[3 * a], tri_rule[3 * a + 1], T[t].v[0], T[t].v[1], T[t].v[2], y, py);
                R vr, vi, sr = 0, si = 0;
                KERNEL_EVAL(op, &pts[3 * p], y, zero, T[t].n, k, clamp, &vr, &vi, &sr, &si);
                R w = w1 * tri_rule[3 * a + 2];
                vr *= w; vi *= w;
                for (int c = 0; c < 3; ++c) { acc_r[c] += vr * py[c]; acc_i[c] += vi * py[c]; }
                s_r += vr; s_i += vi;
            }
            if (basis == ORACLE_DP0) {
                out[2 * (p * ncol + t)] = s_r;
                out[2 * (p * ncol + t) + 1] = s_i;
            } else {
                for (int c = 0; c < 3; ++c) {
                    out[2 * (p * ncol + T[t].idx[c])] += acc_r[c];
                    out[2 * (p * ncol + T[t].idx[c]) + 1] += acc_i[c];
                }
            }
        }
    }
    free(T);
    return ORACLE_OK;
}

/* Mass ("identity") matrix (integrand.h:257-278; bind.cu:10-21).  P1: the
 * triangle rule of sum(w * J/2 * phi_a * phi_b); DP0: diag(J/2).
 * out is dense [n, n] complex. */
int FN(oracle_identity)(const R *verts, int64_t nv, const int32_t *tris, int64_t nt, int basis, const R *tri_rule,
                        int nq, R *out)
{
    if (nv <= 0 || nt <= 0) return ORACLE_BAD_ARG;
    for (int64_t t = 0; t < 3 * nt; ++t)
        if (tris[t] < 0 || tris[t] >= nv) return ORACLE_BAD_MESH;
    FN(tri_t) *T = FN(make_tris)(verts, tris, nt);
    int64_t n = (basis == ORACLE_DP0) ? nt : nv;
    memset(out, 0, sizeof(R) * 2 * (size_t)n * (size_t)n);
    for (int64_t t = 0; t < nt; ++t) {
        R w = (R)(0.5 * (double)T[t].jac);
        if (basis == ORACLE_DP0) { out[2 * (t * n + t)] = w; continue; }
        R blk[9] = {0};
        for (int a = 0; a < nq; ++a) {
            R p[3], ph[3];
            FN(map_std)(tri_rule[3 * a], tri_rule[3 * a + 1], T[t].v[0], T[t].v[1], T[t].v[2], p, ph);
            R lw = tri_rule[3 * a + 2] * w;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) blk[3 * i + j] += lw * ph[i] * ph[j];
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) out[2 * ((int64_t)T[t].idx[i] * n + T[t].idx[j])] += blk[3 * i + j];
    }
    free(T);
    return ORACLE_OK;
}
