/*
 * arap_oracle.c -- FP64 CPU restatement of the As-Rigid-As-Possible deformation of
 * cheind/mesh-deform, inc/deform/arap.h.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP product path: it is
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * nowhere else. The product library (mesh-deform_amd/) never links, loads or calls it.
 *
 * Every function follows the reference loop for loop, in the same iteration orders:
 *   faces ascending (arap.h:189), rows ascending and CSR columns ascending as Eigen's
 *   RowMajor InnerIterator walks them (arap.h:310,319,358,366,397,406), free variables
 *   numbered ascending over vertices (arap.h:267-270).
 * Eigen is not available, so the two numerical library calls are restated from their
 * published algorithms:
 *   - SimplicialLDLT (arap.h:337,420): an up-looking sparse LDL^T (Davis, "Algorithm 849:
 *     A concise sparse Cholesky factorization package", TOMS 2005), with a geometric
 *     nested-dissection fill-reducing ordering in place of Eigen's AMD. Any ordering gives
 *     the same solution up to rounding.
 *   - JacobiSVD (arap.h:376): a one-sided (Hestenes) Jacobi SVD of the 3x3 covariance,
 *     singular values sorted descending so that the reflection fix of arap.h:380-382 flips
 *     the column of the smallest one.
 *
 * Pins (see tests/test_oracle.py): the reference's known-answer cotan test
 * (tests/test_cotan.cpp:20-54), closed-form cotangents, numpy.linalg.svd golden rotations,
 * scipy sparse solves, rigid-motion and translation closed forms, energy monotonicity.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_OK 0
#define OR_ERR_INVALID (-1)
#define OR_ERR_NOMEM (-2)
#define OR_ERR_FACTOR (-3)

/* std::max<Scalar>(a, b) == (a < b) ? b : a.  arap.h:199-218 call it as max(clamp, x),
 * so a NaN x yields the clamp value. */
static double ref_max(double a, double b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------------------------ */
/* Cotangent weights, arap.h:182-239                                                     */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    int r, c;
    int seq; /* insertion order, so duplicates are summed in triplet order like setFromTriplets */
    double v;
} or_trip;

static int trip_cmp(const void *a, const void *b) {
    const or_trip *x = (const or_trip *)a, *y = (const or_trip *)b;
    if (x->r != y->r) return x->r < y->r ? -1 : 1;
    if (x->c != y->c) return x->c < y->c ? -1 : 1;
    return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

/* Builds the symmetric cotan weight matrix as CSR with ascending columns and no diagonal.
 * rowptr has n+1 entries; col/val have room for cap entries (6F always suffices).
 * Returns nnz, or a negative error. */
int or_cotan_csr(int n, int F, const int *tri, const double *p, int *rowptr, int *col,
                 double *val, int cap) {
    if (n < 0 || F < 0) return OR_ERR_INVALID;
    or_trip *t = (or_trip *)malloc(sizeof(or_trip) * (size_t)(6 * (size_t)F + 1));
    if (!t) return OR_ERR_NOMEM;
    int nt = 0;
    for (int f = 0; f < F; ++f) { /* arap.h:189 */
        int i0 = tri[3 * f], i1 = tri[3 * f + 1], i2 = tri[3 * f + 2];
        if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= n || i1 >= n || i2 >= n) {
            free(t);
            return OR_ERR_INVALID;
        }
        const double *v0 = p + 3 * i0, *v1 = p + 3 * i1, *v2 = p + 3 * i2;
        double d[3];
        /* arap.h:199-201: squared lengths clamped at 1e-8 */
        d[0] = v1[0] - v0[0]; d[1] = v1[1] - v0[1]; d[2] = v1[2] - v0[2];
        double l0 = ref_max(1e-8, d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        d[0] = v2[0] - v1[0]; d[1] = v2[1] - v1[1]; d[2] = v2[2] - v1[2];
        double l1 = ref_max(1e-8, d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        d[0] = v0[0] - v2[0]; d[1] = v0[1] - v2[1]; d[2] = v0[2] - v2[2];
        double l2 = ref_max(1e-8, d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        /* arap.h:203-205 */
        l0 = sqrt(l0); l1 = sqrt(l1); l2 = sqrt(l2);
        /* arap.h:207-210: Heron's formula, area clamped at 1e-8 */
        const double semip = 0.5 * (l0 + l1 + l2);
        const double area = ref_max(1e-8, sqrt(semip * (semip - l0) * (semip - l1) * (semip - l2)));
        const double denom = 1.0 / (4.0 * area);
        /* arap.h:212-218: cot of the angle opposite each edge, clamped at 1e-10 */
        double cot0 = (-l0 * l0 + l1 * l1 + l2 * l2) * denom;
        double cot1 = (l0 * l0 - l1 * l1 + l2 * l2) * denom;
        double cot2 = (l0 * l0 + l1 * l1 - l2 * l2) * denom;
        cot0 = ref_max(1e-10, cot0);
        cot1 = ref_max(1e-10, cot1);
        cot2 = ref_max(1e-10, cot2);
        /* arap.h:220-232: edges (0,1),(1,2),(2,0), upper then lower triplets */
        int ea[3] = {i0, i1, i2}, eb[3] = {i1, i2, i0};
        double cw[3] = {cot0 * 0.5, cot1 * 0.5, cot2 * 0.5};
        for (int k = 0; k < 3; ++k) {
            int lo = ea[k] < eb[k] ? ea[k] : eb[k], hi = ea[k] < eb[k] ? eb[k] : ea[k];
            t[nt].r = lo; t[nt].c = hi; t[nt].v = cw[k]; t[nt].seq = nt; ++nt;
        }
        for (int k = 0; k < 3; ++k) {
            int lo = ea[k] < eb[k] ? ea[k] : eb[k], hi = ea[k] < eb[k] ? eb[k] : ea[k];
            t[nt].r = hi; t[nt].c = lo; t[nt].v = cw[k]; t[nt].seq = nt; ++nt;
        }
    }
    qsort(t, (size_t)nt, sizeof(or_trip), trip_cmp);
    /* arap.h:235-238: setFromTriplets sums duplicates */
    int nnz = 0;
    for (int i = 0; i <= n; ++i) rowptr[i] = 0;
    for (int k = 0; k < nt;) {
        int r = t[k].r, c = t[k].c;
        double s = t[k].v;
        int k2 = k + 1;
        while (k2 < nt && t[k2].r == r && t[k2].c == c) { s += t[k2].v; ++k2; }
        if (nnz >= cap) { free(t); return OR_ERR_NOMEM; }
        col[nnz] = c; val[nnz] = s; rowptr[r + 1]++; ++nnz;
        k = k2;
    }
    for (int i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
    free(t);
    return nnz;
}

/* ------------------------------------------------------------------------------------ */
/* 3x3 SVD rotation, arap.h:376-382                                                      */
/* ------------------------------------------------------------------------------------ */

/* One-sided Jacobi SVD of a 3x3 row-major matrix A = U S V^T, S descending.
 * U and V are row-major 3x3 with singular vectors as columns. */
static void svd3_jacobi(const double *A, double *U, double *S, double *V) {
    double B[9];
    memcpy(B, A, sizeof(B));
    for (int i = 0; i < 9; ++i) V[i] = 0.0;
    V[0] = V[4] = V[8] = 1.0;
    double fro = 0.0;
    for (int i = 0; i < 9; ++i) fro += A[i] * A[i];
    /* columns below this squared norm are rounding noise and are left alone */
    const double tiny = 1e-30 * fro;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int pcol = 0; pcol < 2; ++pcol) {
            for (int qcol = pcol + 1; qcol < 3; ++qcol) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int r = 0; r < 3; ++r) {
                    alpha += B[3 * r + pcol] * B[3 * r + pcol];
                    beta += B[3 * r + qcol] * B[3 * r + qcol];
                    gamma += B[3 * r + pcol] * B[3 * r + qcol];
                }
                if (gamma == 0.0 || alpha <= tiny || beta <= tiny) continue;
                double rel = fabs(gamma) / sqrt(alpha * beta);
                if (!(rel > 1e-15)) continue;
                if (rel > off) off = rel;
                double zeta = (beta - alpha) / (2.0 * gamma);
                double tt = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + tt * tt), s = c * tt;
                for (int r = 0; r < 3; ++r) {
                    double bp = B[3 * r + pcol], bq = B[3 * r + qcol];
                    B[3 * r + pcol] = c * bp - s * bq;
                    B[3 * r + qcol] = s * bp + c * bq;
                    double vp = V[3 * r + pcol], vq = V[3 * r + qcol];
                    V[3 * r + pcol] = c * vp - s * vq;
                    V[3 * r + qcol] = s * vp + c * vq;
                }
            }
        }
        if (off <= 1e-15) break;
    }
    double nrm[3];
    for (int k = 0; k < 3; ++k)
        nrm[k] = sqrt(B[k] * B[k] + B[3 + k] * B[3 + k] + B[6 + k] * B[6 + k]);
    /* sort columns by singular value, descending */
    int ord[3] = {0, 1, 2};
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b)
            if (nrm[ord[b]] > nrm[ord[a]]) { int tmp = ord[a]; ord[a] = ord[b]; ord[b] = tmp; }
    double Bs[9], Vs[9];
    for (int k = 0; k < 3; ++k) {
        S[k] = nrm[ord[k]];
        for (int r = 0; r < 3; ++r) {
            Bs[3 * r + k] = B[3 * r + ord[k]];
            Vs[3 * r + k] = V[3 * r + ord[k]];
        }
    }
    memcpy(V, Vs, sizeof(Vs));
    /* U columns: B v / sigma; complete the basis where sigma vanishes */
    double tol = 1e-13 * (S[0] > 0 ? S[0] : 1.0);
    for (int k = 0; k < 3; ++k) {
        if (S[k] > tol) {
            for (int r = 0; r < 3; ++r) U[3 * r + k] = Bs[3 * r + k] / S[k];
        } else if (k == 0) {
            U[0] = 1; U[3] = 0; U[6] = 0;
        } else if (k == 1) {
            /* any unit vector orthogonal to u0 */
            double u0[3] = {U[0], U[3], U[6]};
            double a[3] = {fabs(u0[0]) < 0.9 ? 1.0 : 0.0, fabs(u0[0]) < 0.9 ? 0.0 : 1.0, 0.0};
            double d = a[0] * u0[0] + a[1] * u0[1] + a[2] * u0[2];
            double w[3] = {a[0] - d * u0[0], a[1] - d * u0[1], a[2] - d * u0[2]};
            double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
            U[1] = w[0] / wn; U[4] = w[1] / wn; U[7] = w[2] / wn;
        } else {
            /* u2 = u0 x u1 */
            U[2] = U[3] * U[7] - U[6] * U[4];
            U[5] = U[6] * U[1] - U[0] * U[7];
            U[8] = U[0] * U[4] - U[3] * U[1];
        }
    }
}

static double det3(const double *M) {
    return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
           M[2] * (M[3] * M[7] - M[4] * M[6]);
}

/* R = V * diag(1, 1, det(V U^T)) * U^T  (arap.h:377-382). cov and R are row-major. */
void or_svd_rotation(const double *cov, double *R) {
    double U[9], S[3], V[9];
    svd3_jacobi(cov, U, S, V);
    double VUt[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            VUt[3 * r + c] = V[3 * r] * U[3 * c] + V[3 * r + 1] * U[3 * c + 1] + V[3 * r + 2] * U[3 * c + 2];
    double dd = det3(VUt);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            R[3 * r + c] = V[3 * r] * U[3 * c] + V[3 * r + 1] * U[3 * c + 1] + dd * V[3 * r + 2] * U[3 * c + 2];
}

/* Exposed for golden-vector tests: singular values of a 3x3 row-major matrix. */
void or_svd3(const double *A, double *U, double *S, double *V) { svd3_jacobi(A, U, S, V); }

/* ------------------------------------------------------------------------------------ */
/* Nested-dissection ordering (fill reduction for the sparse LDL^T)                       */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    const int *rp, *ci;      /* adjacency (CSR, may contain the diagonal) */
    const double *xyz;       /* coordinates, 3 per vertex */
    int *stamp, *side;       /* per-vertex work arrays */
    int stamp_ctr;
    int *out;                /* ordering being produced (new -> old) */
    int nout;
    int axis;                /* sort axis for the comparator */
} nd_ctx;

static nd_ctx *g_nd; /* comparator context; the oracle is single-threaded */

static int nd_cmp(const void *a, const void *b) {
    int i = *(const int *)a, j = *(const int *)b;
    double x = g_nd->xyz[3 * i + g_nd->axis], y = g_nd->xyz[3 * j + g_nd->axis];
    if (x < y) return -1;
    if (x > y) return 1;
    return (i > j) - (i < j);
}

#define ND_LEAF 16

static void nd_rec(nd_ctx *c, int *v, int cnt) {
    if (cnt <= ND_LEAF) {
        for (int k = 0; k < cnt; ++k) c->out[c->nout++] = v[k];
        return;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int k = 0; k < cnt; ++k)
        for (int a = 0; a < 3; ++a) {
            double x = c->xyz[3 * v[k] + a];
            if (x < lo[a]) lo[a] = x;
            if (x > hi[a]) hi[a] = x;
        }
    int axis = 0;
    for (int a = 1; a < 3; ++a)
        if (hi[a] - lo[a] > hi[axis] - lo[axis]) axis = a;
    c->axis = axis;
    g_nd = c;
    qsort(v, (size_t)cnt, sizeof(int), nd_cmp);
    int half = cnt / 2;
    int st = ++c->stamp_ctr;
    for (int k = 0; k < cnt; ++k) { c->stamp[v[k]] = st; c->side[v[k]] = k < half ? 0 : 1; }
    /* separator: vertices of the lower half with a neighbour in the upper half */
    int *tmp = (int *)malloc(sizeof(int) * (size_t)cnt);
    int nl = 0, ns = 0;
    int *sep = tmp + half; /* room: at most half entries */
    for (int k = 0; k < half; ++k) {
        int x = v[k], is_sep = 0;
        for (int e = c->rp[x]; e < c->rp[x + 1]; ++e) {
            int y = c->ci[e];
            if (y != x && c->stamp[y] == st && c->side[y] == 1) { is_sep = 1; break; }
        }
        if (is_sep) sep[ns++] = x; else tmp[nl++] = x;
    }
    int *right = (int *)malloc(sizeof(int) * (size_t)(cnt - half));
    memcpy(right, v + half, sizeof(int) * (size_t)(cnt - half));
    int *seps = (int *)malloc(sizeof(int) * (size_t)(ns > 0 ? ns : 1));
    memcpy(seps, sep, sizeof(int) * (size_t)ns);
    int *left = tmp;
    nd_rec(c, left, nl);
    nd_rec(c, right, cnt - half);
    for (int k = 0; k < ns; ++k) c->out[c->nout++] = seps[k];
    free(right);
    free(seps);
    free(tmp);
}

/* perm[new] = old. */
int or_nd_order(int m, const int *rp, const int *ci, const double *xyz, int *perm) {
    nd_ctx c;
    c.rp = rp; c.ci = ci; c.xyz = xyz;
    c.stamp = (int *)calloc((size_t)(m > 0 ? m : 1), sizeof(int));
    c.side = (int *)calloc((size_t)(m > 0 ? m : 1), sizeof(int));
    c.stamp_ctr = 0; c.out = perm; c.nout = 0;
    int *v = (int *)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    if (!c.stamp || !c.side || !v) { free(c.stamp); free(c.side); free(v); return OR_ERR_NOMEM; }
    for (int k = 0; k < m; ++k) v[k] = k;
    nd_rec(&c, v, m);
    free(v); free(c.stamp); free(c.side);
    return c.nout == m ? OR_OK : OR_ERR_INVALID;
}

/* ------------------------------------------------------------------------------------ */
/* Sparse LDL^T (stands in for Eigen::SimplicialLDLT, arap.h:337,420,462)                */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    int n;
    int *P, *Pinv;
    int *Lp, *Li;
    double *Lx, *D;
    long long nnzL;
} or_ldl;

static void ldl_free(or_ldl *f) {
    if (!f) return;
    free(f->P); free(f->Pinv); free(f->Lp); free(f->Li); free(f->Lx); free(f->D);
    memset(f, 0, sizeof(*f));
}

/* Factor the symmetric matrix A (full CSR pattern) permuted by P. Up-looking LDL^T.
 * Returns OR_OK, or OR_ERR_FACTOR when a pivot vanishes (|d_k| <= 1e-14 * max|A_kk|). */
static int ldl_factor(or_ldl *f, int n, const int *Ap, const int *Ai, const double *Ax, const int *P) {
    memset(f, 0, sizeof(*f));
    f->n = n;
    size_t nn = (size_t)(n > 0 ? n : 1);
    f->P = (int *)malloc(sizeof(int) * nn);
    f->Pinv = (int *)malloc(sizeof(int) * nn);
    f->Lp = (int *)malloc(sizeof(int) * (nn + 1));
    f->D = (double *)malloc(sizeof(double) * nn);
    int *Parent = (int *)malloc(sizeof(int) * nn), *Lnz = (int *)malloc(sizeof(int) * nn);
    int *Flag = (int *)malloc(sizeof(int) * nn), *Pattern = (int *)malloc(sizeof(int) * nn);
    double *Y = (double *)malloc(sizeof(double) * nn);
    if (!f->P || !f->Pinv || !f->Lp || !f->D || !Parent || !Lnz || !Flag || !Pattern || !Y) {
        free(Parent); free(Lnz); free(Flag); free(Pattern); free(Y); ldl_free(f);
        return OR_ERR_NOMEM;
    }
    for (int k = 0; k < n; ++k) { f->P[k] = P[k]; f->Pinv[P[k]] = k; }
    /* symbolic: elimination tree and column counts */
    for (int k = 0; k < n; ++k) {
        Parent[k] = -1; Flag[k] = k; Lnz[k] = 0;
        int kk = P[k];
        for (int p = Ap[kk]; p < Ap[kk + 1]; ++p) {
            int i = f->Pinv[Ai[p]];
            if (i < k) {
                for (; Flag[i] != k; i = Parent[i]) {
                    if (Parent[i] == -1) Parent[i] = k;
                    Lnz[i]++;
                    Flag[i] = k;
                }
            }
        }
    }
    f->Lp[0] = 0;
    for (int k = 0; k < n; ++k) f->Lp[k + 1] = f->Lp[k] + Lnz[k];
    f->nnzL = f->Lp[n];
    f->Li = (int *)malloc(sizeof(int) * (size_t)(f->nnzL > 0 ? f->nnzL : 1));
    f->Lx = (double *)malloc(sizeof(double) * (size_t)(f->nnzL > 0 ? f->nnzL : 1));
    if (!f->Li || !f->Lx) {
        free(Parent); free(Lnz); free(Flag); free(Pattern); free(Y); ldl_free(f);
        return OR_ERR_NOMEM;
    }
    double amax = 0.0;
    for (int k = 0; k < n; ++k) {
        int kk = P[k];
        for (int p = Ap[kk]; p < Ap[kk + 1]; ++p)
            if (Ai[p] == kk && fabs(Ax[p]) > amax) amax = fabs(Ax[p]);
    }
    /* numeric */
    int rc = OR_OK;
    for (int k = 0; k < n; ++k) {
        Y[k] = 0.0;
        int top = n;
        Flag[k] = k;
        Lnz[k] = 0;
        int kk = P[k];
        for (int p = Ap[kk]; p < Ap[kk + 1]; ++p) {
            int i = f->Pinv[Ai[p]];
            if (i <= k) {
                Y[i] += Ax[p];
                int len = 0;
                for (; Flag[i] != k; i = Parent[i]) { Pattern[len++] = i; Flag[i] = k; }
                while (len > 0) Pattern[--top] = Pattern[--len];
            }
        }
        f->D[k] = Y[k];
        Y[k] = 0.0;
        for (; top < n; ++top) {
            int i = Pattern[top];
            double yi = Y[i];
            Y[i] = 0.0;
            int p2 = f->Lp[i] + Lnz[i];
            for (int p = f->Lp[i]; p < p2; ++p) Y[f->Li[p]] -= f->Lx[p] * yi;
            double l_ki = yi / f->D[i];
            f->D[k] -= l_ki * yi;
            f->Li[p2] = k;
            f->Lx[p2] = l_ki;
            Lnz[i]++;
        }
        if (!(fabs(f->D[k]) > 1e-14 * amax)) { rc = OR_ERR_FACTOR; break; }
    }
    free(Parent); free(Lnz); free(Flag); free(Pattern); free(Y);
    if (rc != OR_OK) ldl_free(f);
    return rc;
}

/* x = A^{-1} b using the factor; b and x are length n and may alias. */
static void ldl_solve(const or_ldl *f, const double *b, double *x, double *work) {
    int n = f->n;
    for (int k = 0; k < n; ++k) work[k] = b[f->P[k]];
    for (int j = 0; j < n; ++j)
        for (int p = f->Lp[j]; p < f->Lp[j + 1]; ++p) work[f->Li[p]] -= f->Lx[p] * work[j];
    for (int j = 0; j < n; ++j) work[j] /= f->D[j];
    for (int j = n - 1; j >= 0; --j)
        for (int p = f->Lp[j]; p < f->Lp[j + 1]; ++p) work[j] -= f->Lx[p] * work[f->Li[p]];
    for (int k = 0; k < n; ++k) x[f->P[k]] = work[k];
}

/* Stand-alone SPD solve for tests: A (full CSR) x = b, nrhs right-hand sides stored
 * column after column (b[r*n + i]); xyz gives coordinates for the ordering (may be NULL:
 * natural order). Returns OR_OK / OR_ERR_FACTOR; *nnzL receives the strict factor nnz. */
int or_spd_solve(int n, const int *Ap, const int *Ai, const double *Ax, const double *xyz,
                 int nrhs, const double *b, double *x, long long *nnzL) {
    int *perm = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    double *work = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!perm || !work) { free(perm); free(work); return OR_ERR_NOMEM; }
    if (xyz) {
        int rc = or_nd_order(n, Ap, Ai, xyz, perm);
        if (rc) { free(perm); free(work); return rc; }
    } else {
        for (int k = 0; k < n; ++k) perm[k] = k;
    }
    or_ldl f;
    int rc = ldl_factor(&f, n, Ap, Ai, Ax, perm);
    if (rc == OR_OK) {
        for (int r = 0; r < nrhs; ++r) ldl_solve(&f, b + (size_t)r * n, x + (size_t)r * n, work);
        if (nnzL) *nnzL = f.nnzL;
        ldl_free(&f);
    }
    free(perm); free(work);
    return rc;
}

/* ------------------------------------------------------------------------------------ */
/* The ARAP object, arap.h:49-466                                                        */
/* ------------------------------------------------------------------------------------ */

typedef struct or_arap {
    int n, F;
    int *tri;              /* _faces (arap.h:452), 3 per face */
    double *p, *pp;        /* _p, _pprime (arap.h:451), 3 per vertex */
    int *w_rp, *w_ci;      /* _edgeWeights (arap.h:453), CSR with ascending columns */
    double *w_v;
    int nnzW;
    double *R;             /* _rotations (arap.h:454), row-major 3x3 per vertex */
    int nf;                /* _numberOfFreeVariables */
    int *freeIdx;          /* _freeIdxMap (arap.h:457) */
    int *is_con;           /* key set of _constrainedLocations (arap.h:458) */
    double *con;           /* its values */
    int *L_rp, *L_ci;      /* _L (arap.h:460), CSR over free variables */
    double *L_v;
    int nnzL_full;
    double *bFixed, *b;    /* arap.h:461, 3 x nf stored as b[3*i+d] */
    or_ldl fac;            /* _solver (arap.h:462) */
    int have_factor;
    int dirty;             /* arap.h:464 */
    int initialized;       /* rest pose captured at least once (fixed-rest mode) */
    int con_set_changed;   /* constrained index set changed since the last factorization */
    double *tmp;           /* solver workspace */
} or_arap;

void or_arap_destroy(or_arap *a) {
    if (!a) return;
    free(a->tri); free(a->p); free(a->pp); free(a->w_rp); free(a->w_ci); free(a->w_v);
    free(a->R); free(a->freeIdx); free(a->is_con); free(a->con);
    free(a->L_rp); free(a->L_ci); free(a->L_v); free(a->bFixed); free(a->b); free(a->tmp);
    ldl_free(&a->fac);
    free(a);
}

/* ctor + initializeMeshTopology, arap.h:66-70,149-155 */
or_arap *or_arap_create(int n, int F, const int *tri) {
    if (n < 0 || F < 0) return NULL;
    or_arap *a = (or_arap *)calloc(1, sizeof(or_arap));
    if (!a) return NULL;
    a->n = n; a->F = F;
    size_t nn = (size_t)(n > 0 ? n : 1);
    a->tri = (int *)malloc(sizeof(int) * (size_t)(3 * (size_t)F + 1));
    a->p = (double *)calloc(3 * nn, sizeof(double));
    a->pp = (double *)calloc(3 * nn, sizeof(double));
    a->R = (double *)calloc(9 * nn, sizeof(double));
    a->freeIdx = (int *)malloc(sizeof(int) * nn);
    a->is_con = (int *)calloc(nn, sizeof(int));
    a->con = (double *)calloc(3 * nn, sizeof(double));
    a->w_rp = (int *)calloc(nn + 1, sizeof(int));
    a->w_ci = (int *)malloc(sizeof(int) * (size_t)(6 * (size_t)F + 1));
    a->w_v = (double *)malloc(sizeof(double) * (size_t)(6 * (size_t)F + 1));
    a->tmp = (double *)malloc(sizeof(double) * nn);
    if (!a->tri || !a->p || !a->pp || !a->R || !a->freeIdx || !a->is_con || !a->con ||
        !a->w_rp || !a->w_ci || !a->w_v || !a->tmp) {
        or_arap_destroy(a);
        return NULL;
    }
    for (int k = 0; k < 3 * F; ++k) {
        if (tri[k] < 0 || tri[k] >= n) { or_arap_destroy(a); return NULL; }
        a->tri[k] = tri[k];
    }
    a->dirty = 1;
    a->con_set_changed = 1;
    return a;
}

/* setConstraint, arap.h:81-85. Out-of-range indices are an error here (UB in the reference). */
int or_arap_set_constraint(or_arap *a, int vidx, const double *xyz) {
    if (vidx < 0 || vidx >= a->n) return OR_ERR_INVALID;
    if (!a->is_con[vidx]) a->con_set_changed = 1;
    a->is_con[vidx] = 1;
    a->con[3 * vidx] = xyz[0]; a->con[3 * vidx + 1] = xyz[1]; a->con[3 * vidx + 2] = xyz[2];
    a->dirty = 1;
    return OR_OK;
}

/* initializeFreeVariableMapping, arap.h:261-272 */
static void init_free_map(or_arap *a) {
    int fi = 0;
    for (int i = 0; i < a->n; ++i) a->freeIdx[i] = a->is_con[i] ? -1 : fi++;
    a->nf = fi;
}

/* setupLinearSystem, arap.h:292-340. Builds L (CSR, ascending columns, duplicates summed in
 * insertion order) and bFixed; factors when refactor is set. */
static int setup_linear_system(or_arap *a, int refactor) {
    int nf = a->nf;
    free(a->bFixed); free(a->b);
    a->bFixed = (double *)calloc(3 * (size_t)(nf > 0 ? nf : 1), sizeof(double));
    a->b = (double *)calloc(3 * (size_t)(nf > 0 ? nf : 1), sizeof(double));
    if (!a->bFixed || !a->b) return OR_ERR_NOMEM;
    if (refactor) {
        free(a->L_rp); free(a->L_ci); free(a->L_v);
        a->L_rp = (int *)calloc((size_t)nf + 1, sizeof(int));
        a->L_ci = (int *)malloc(sizeof(int) * (size_t)(a->nnzW + nf + 1));
        a->L_v = (double *)malloc(sizeof(double) * (size_t)(a->nnzW + nf + 1));
        if (!a->L_rp || !a->L_ci || !a->L_v) return OR_ERR_NOMEM;
    }
    int nnz = 0;
    for (int vi = 0; vi < a->n; ++vi) { /* arap.h:310 */
        int iidx = a->freeIdx[vi];
        if (iidx == -1) continue; /* arap.h:314-317 */
        /* Row iidx of L: off-diagonals in ascending jidx order (the free map is monotone, so
         * ascending vertex columns give ascending free columns); the diagonal sums every
         * neighbour weight in ascending column order (arap.h:332). */
        double diag = 0.0;
        int row_start = nnz, diag_pos = -1;
        for (int e = a->w_rp[vi]; e < a->w_rp[vi + 1]; ++e) { /* arap.h:319 */
            int vj = a->w_ci[e];
            int jidx = a->freeIdx[vj];
            double wij = a->w_v[e];
            if (jidx == -1) { /* arap.h:325-327 */
                a->bFixed[3 * iidx] += wij * a->con[3 * vj];
                a->bFixed[3 * iidx + 1] += wij * a->con[3 * vj + 1];
                a->bFixed[3 * iidx + 2] += wij * a->con[3 * vj + 2];
            } else if (refactor) { /* arap.h:329 */
                if (diag_pos < 0 && jidx > iidx) { diag_pos = nnz; a->L_ci[nnz] = iidx; a->L_v[nnz] = 0; ++nnz; }
                a->L_ci[nnz] = jidx; a->L_v[nnz] = -wij; ++nnz;
            }
            diag += wij; /* arap.h:332 */
        }
        if (refactor) {
            if (diag_pos < 0) { diag_pos = nnz; a->L_ci[nnz] = iidx; ++nnz; }
            a->L_v[diag_pos] = diag;
            a->L_rp[iidx + 1] = nnz;
            (void)row_start;
        }
    }
    if (refactor) {
        a->nnzL_full = nnz;
        ldl_free(&a->fac);
        a->have_factor = 0;
        /* ordering by rest coordinates of the free vertices */
        double *fx = (double *)malloc(sizeof(double) * 3 * (size_t)(nf > 0 ? nf : 1));
        int *perm = (int *)malloc(sizeof(int) * (size_t)(nf > 0 ? nf : 1));
        if (!fx || !perm) { free(fx); free(perm); return OR_ERR_NOMEM; }
        for (int vi = 0; vi < a->n; ++vi)
            if (a->freeIdx[vi] >= 0) memcpy(fx + 3 * a->freeIdx[vi], a->p + 3 * vi, 3 * sizeof(double));
        int rc = or_nd_order(nf, a->L_rp, a->L_ci, fx, perm);
        if (rc == OR_OK) rc = ldl_factor(&a->fac, nf, a->L_rp, a->L_ci, a->L_v, perm); /* arap.h:337 */
        free(fx); free(perm);
        if (rc != OR_OK) return rc; /* arap.h:339 */
        a->have_factor = 1;
    }
    return OR_OK;
}

/* cotan weights of the current rest pose into the object, arap.h:182-239 */
static int compute_cotan(or_arap *a) {
    int nnz = or_cotan_csr(a->n, a->F, a->tri, a->p, a->w_rp, a->w_ci, a->w_v, 6 * a->F + 1);
    if (nnz < 0) return nnz;
    a->nnzW = nnz;
    return OR_OK;
}

/* estimateRotations, arap.h:354-384 */
static void estimate_rotations(or_arap *a) {
    for (int vi = 0; vi < a->n; ++vi) {
        const double *pi = a->p + 3 * vi, *ppi = a->pp + 3 * vi;
        double cov[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int e = a->w_rp[vi]; e < a->w_rp[vi + 1]; ++e) { /* arap.h:366-374 */
            int vj = a->w_ci[e];
            double wij = a->w_v[e];
            const double *pj = a->p + 3 * vj, *ppj = a->pp + 3 * vj;
            double d[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
            double dp[3] = {ppi[0] - ppj[0], ppi[1] - ppj[1], ppi[2] - ppj[2]};
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) cov[3 * r + c] += wij * d[r] * dp[c];
        }
        or_svd_rotation(cov, a->R + 9 * vi); /* arap.h:376-382 */
    }
}

/* estimatePositions, arap.h:393-430 */
static void estimate_positions(or_arap *a) {
    int nf = a->nf;
    memcpy(a->b, a->bFixed, sizeof(double) * 3 * (size_t)nf); /* arap.h:395 */
    for (int vi = 0; vi < a->n; ++vi) {
        int iidx = a->freeIdx[vi];
        if (iidx == -1) continue;
        const double *Ri = a->R + 9 * vi, *pi = a->p + 3 * vi;
        for (int e = a->w_rp[vi]; e < a->w_rp[vi + 1]; ++e) { /* arap.h:406-413 */
            int vj = a->w_ci[e];
            double wij = a->w_v[e];
            const double *Rj = a->R + 9 * vj, *pj = a->p + 3 * vj;
            double d[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
            for (int r = 0; r < 3; ++r) {
                double rd = (Ri[3 * r] + Rj[3 * r]) * d[0] + (Ri[3 * r + 1] + Rj[3 * r + 1]) * d[1] +
                            (Ri[3 * r + 2] + Rj[3 * r + 2]) * d[2];
                a->b[3 * iidx + r] += rd * wij * 0.5;
            }
        }
    }
    /* arap.h:418-429: one solve per coordinate, scattered to the free vertices in order */
    double *rhs = (double *)malloc(sizeof(double) * (size_t)(nf > 0 ? nf : 1));
    double *u = (double *)malloc(sizeof(double) * (size_t)(nf > 0 ? nf : 1));
    for (int d = 0; d < 3; ++d) {
        for (int k = 0; k < nf; ++k) rhs[k] = a->b[3 * k + d];
        ldl_solve(&a->fac, rhs, u, a->tmp);
        int idx = 0;
        for (int vi = 0; vi < a->n; ++vi)
            if (a->freeIdx[vi] != -1) a->pp[3 * vi + d] = u[idx++];
    }
    free(rhs); free(u);
}

/* ARAP energy sum_i sum_j w_ij |(p'_i - p'_j) - R_i (p_i - p_j)|^2 with the current R, p'. */
double or_arap_energy(const or_arap *a) {
    double E = 0.0;
    for (int vi = 0; vi < a->n; ++vi) {
        const double *Ri = a->R + 9 * vi, *pi = a->p + 3 * vi, *ppi = a->pp + 3 * vi;
        for (int e = a->w_rp[vi]; e < a->w_rp[vi + 1]; ++e) {
            int vj = a->w_ci[e];
            const double *pj = a->p + 3 * vj, *ppj = a->pp + 3 * vj;
            double d[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
            double s = 0.0;
            for (int r = 0; r < 3; ++r) {
                double t = (ppi[r] - ppj[r]) - (Ri[3 * r] * d[0] + Ri[3 * r + 1] * d[1] + Ri[3 * r + 2] * d[2]);
                s += t * t;
            }
            E += a->w_v[e] * s;
        }
    }
    return E;
}

/* deform, arap.h:101-138.
 * mesh_xyz is the adapter's vertex storage (3 per vertex): read when dirty
 * (initializeMeshGeometry, arap.h:162-168) and written back at the end (arap.h:133-135).
 * rest_mode 0 = reference: every dirty call takes the current mesh as the new rest pose.
 * rest_mode 1 = fixed rest: the rest pose, weights and factor of the first dirty call are
 *   kept; later dirty calls only take the current mesh as the initial guess, rebuild bFixed,
 *   and refactor only if the constrained index set changed.
 * energy (may be NULL) receives the ARAP energy after each iteration.
 * Returns 1 (true), 0 (false: factorization failed) or a negative error. */
int or_arap_deform(or_arap *a, int iters, double *mesh_xyz, int rest_mode, double *energy) {
    int n = a->n;
    if (iters < 0) return OR_ERR_INVALID;
    if (a->dirty) {
        int full = (rest_mode == 0) || !a->initialized;
        if (full) {
            memcpy(a->p, mesh_xyz, sizeof(double) * 3 * (size_t)n); /* arap.h:162-168 */
            memcpy(a->pp, mesh_xyz, sizeof(double) * 3 * (size_t)n);
            int rc = compute_cotan(a); /* arap.h:108 */
            if (rc) return rc;
        } else {
            memcpy(a->pp, mesh_xyz, sizeof(double) * 3 * (size_t)n);
        }
        for (int k = 0; k < 9 * n; ++k) a->R[k] = (k % 9 == 0 || k % 9 == 4 || k % 9 == 8) ? 1.0 : 0.0; /* arap.h:246-249 */
        init_free_map(a); /* arap.h:110 */
        for (int i = 0; i < n; ++i) /* arap.h:277-281 */
            if (a->is_con[i]) memcpy(a->pp + 3 * i, a->con + 3 * i, 3 * sizeof(double));
        a->initialized = 1;
        if (a->nf == n) return 1; /* arap.h:113-114: stays dirty, no write-back */
        int refactor = full || a->con_set_changed || !a->have_factor;
        int rc = setup_linear_system(a, refactor); /* arap.h:116-117 */
        if (rc == OR_ERR_FACTOR) return 0;
        if (rc) return rc;
        a->con_set_changed = 0;
        a->dirty = 0; /* arap.h:119 */
    }
    for (int it = 0; it < iters; ++it) { /* arap.h:122-129 */
        estimate_rotations(a);
        if (a->nf > 0) estimate_positions(a);
        if (energy) energy[it] = or_arap_energy(a);
    }
    memcpy(mesh_xyz, a->pp, sizeof(double) * 3 * (size_t)n); /* arap.h:133-135 */
    return 1;
}

/* White-box accessors (the reference's PrivateAccessor, tests/accessor.h:16-22). */
int or_arap_nnz_weights(const or_arap *a) { return a->nnzW; }
int or_arap_cotan(const or_arap *a, int *rowptr, int *col, double *val) {
    memcpy(rowptr, a->w_rp, sizeof(int) * (size_t)(a->n + 1));
    memcpy(col, a->w_ci, sizeof(int) * (size_t)a->nnzW);
    memcpy(val, a->w_v, sizeof(double) * (size_t)a->nnzW);
    return a->nnzW;
}
void or_arap_rotations(const or_arap *a, double *R) { memcpy(R, a->R, sizeof(double) * 9 * (size_t)a->n); }
void or_arap_pprime(const or_arap *a, double *pp) { memcpy(pp, a->pp, sizeof(double) * 3 * (size_t)a->n); }
void or_arap_rest(const or_arap *a, double *p) { memcpy(p, a->p, sizeof(double) * 3 * (size_t)a->n); }
int or_arap_nfree(const or_arap *a) { return a->nf; }
void or_arap_bfixed(const or_arap *a, double *b) { memcpy(b, a->bFixed, sizeof(double) * 3 * (size_t)a->nf); }
void or_arap_rhs(const or_arap *a, double *b) { memcpy(b, a->b, sizeof(double) * 3 * (size_t)a->nf); }
long long or_arap_factor_nnz(const or_arap *a) { return a->have_factor ? a->fac.nnzL : -1; }
int or_arap_laplacian_nnz(const or_arap *a) { return a->nnzL_full; }

/* Single building blocks on caller state, for kernel-level parity tests. */
void or_rotations(int n, const int *rp, const int *ci, const double *w, const double *p,
                  const double *pp, double *R) {
    for (int vi = 0; vi < n; ++vi) {
        double cov[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int e = rp[vi]; e < rp[vi + 1]; ++e) {
            int vj = ci[e];
            double d[3], dp[3];
            for (int k = 0; k < 3; ++k) { d[k] = p[3 * vi + k] - p[3 * vj + k]; dp[k] = pp[3 * vi + k] - pp[3 * vj + k]; }
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) cov[3 * r + c] += w[e] * d[r] * dp[c];
        }
        or_svd_rotation(cov, R + 9 * vi);
    }
}
