/*
 * Element integration (COO), HRZ lumped mass and canonical COO->CSR (host side).
 *
 * The reference integrates every element with a fixed quadrature rule and
 * accumulates, per COO slot, the quadrature terms in quadrature order
 * (assemble/mass_sparse_q4.c:199-256, stiffness_sparse_q4.c:199-271 and the
 * T3/T6/Q9 siblings).  Here one pass per element evaluates all seven operators
 * (M, S1..S6); each slot still receives its terms in quadrature order with the
 * same operand order, so the values are bit-identical.  The lumped mass keeps
 * the reference's single global accumulation order (element, point, i, j)
 * because its scale factor depends on it.
 *
 * Shared tables (kept equal to the reference on purpose; the values are the
 * reference's decimal literals and any change alters the CSR bits):
 *   36-point rule on [0,1]^2 for Q4/Q9: mass_sparse_q4.c:40-156
 *   12-point rule on the unit triangle for T3/T6: mass_sparse_t3.c:37-85
 * Shape functions: shape/shape_{q4,q9,t3,t6}.c (same operand order).
 * Build with -ffp-contract=off and without -ffast-math.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "seisfem_host.h"

/* ---- quadrature rules ---------------------------------------------------- */

/* 1-D abscissae of the 6x6 rule; point q has r = G6[q/6], s = G6[q%6] */
static const double G6[6] = {0.033765242898424, 0.169395306766868, 0.380690406958402,
                             0.619309593041599, 0.830604693233132, 0.966234757101576};
/* its 2-D weights, row-major over (q/6, q%6) */
static const double W6x6[36] = {
    0.007338020422245, 0.015451823343096, 0.020041279329452, 0.020041279329452, 0.015451823343096, 0.007338020422245,
    0.015451823343096, 0.032537228147042, 0.042201341771897, 0.042201341771897, 0.032537228147042, 0.015451823343096,
    0.020041279329452, 0.042201341771897, 0.054735862541824, 0.054735862541824, 0.042201341771897, 0.020041279329452,
    0.020041279329452, 0.042201341771897, 0.054735862541824, 0.054735862541824, 0.042201341771897, 0.020041279329452,
    0.015451823343096, 0.032537228147042, 0.042201341771897, 0.042201341771897, 0.032537228147042, 0.015451823343096,
    0.007338020422245, 0.015451823343096, 0.020041279329452, 0.020041279329452, 0.015451823343096, 0.007338020422245};

/* 12-point triangle rule: three orbits of 3 and one of 6 */
#define TA 0.87382197101699600
#define TB 0.06308901449150200
#define TC 0.50142650965817900
#define TD 0.24928674517091000
#define TE 0.63650249912139900
#define TF 0.31035245103378500
#define TG 0.05314504984481600
#define TWU 0.05084490637020700
#define TWV 0.11678627572637900
#define TWW 0.08285107561837400
static const double T12r[12] = {TA, TB, TB, TC, TD, TD, TE, TE, TF, TF, TG, TG};
static const double T12s[12] = {TB, TA, TB, TD, TC, TD, TF, TG, TE, TG, TE, TF};
static const double T12w[12] = {TWU, TWU, TWU, TWV, TWV, TWV, TWW, TWW, TWW, TWW, TWW, TWW};

/* ---- shape functions on the reference element ---------------------------- */

static void shape_eval(int type, double r, double s, double *t, double *dr, double *ds)
{
    switch (type) {
    case SFH_Q4:
        t[0] = (1.0 - r) * (1.0 - s);
        t[1] = r * (1.0 - s);
        t[2] = r * s;
        t[3] = (1.0 - r) * s;
        dr[0] = -1.0 + s; dr[1] = 1.0 - s; dr[2] = s;  dr[3] = -s;
        ds[0] = -1.0 + r; ds[1] = -r;      ds[2] = r;  ds[3] = 1.0 - r;
        break;
    case SFH_T3:
        t[0] = 1.0 - r - s; t[1] = r; t[2] = s;
        dr[0] = -1.0; dr[1] = 1.0; dr[2] = 0.0;
        ds[0] = -1.0; ds[1] = 0.0; ds[2] = 1.0;
        break;
    case SFH_T6: {
        const double l = 1.0 - r - s;
        t[0] = 2.0 * l * (0.5 - r - s);
        t[1] = 2.0 * r * (r - 0.5);
        t[2] = 2.0 * s * (s - 0.5);
        t[3] = 4.0 * r * l;
        t[4] = 4.0 * r * s;
        t[5] = 4.0 * s * l;
        dr[0] = -3.0 + 4.0 * r + 4.0 * s;
        dr[1] = -1.0 + 4.0 * r;
        dr[2] = 0.0;
        dr[3] = 4.0 - 8.0 * r - 4.0 * s;
        dr[4] = 4.0 * s;
        dr[5] = -4.0 * s;
        ds[0] = -3.0 + 4.0 * r + 4.0 * s;
        ds[1] = 0.0;
        ds[2] = -1.0 + 4.0 * s;
        ds[3] = -4.0 * r;
        ds[4] = 4.0 * r;
        ds[5] = 4.0 - 4.0 * r - 8.0 * s;
        break;
    }
    case SFH_Q9: {
        const double r1 = r - 1.0, rh = r - 0.5, s1 = s - 1.0, sh = s - 0.5;
        t[0] = 4.0 * r1 * rh * s1 * sh;
        t[1] = 4.0 * r * rh * s1 * sh;
        t[2] = 4.0 * r * rh * s * sh;
        t[3] = 4.0 * r1 * rh * s * sh;
        t[4] = -8.0 * r * r1 * s1 * sh;
        t[5] = -8.0 * r * rh * s * s1;
        t[6] = -8.0 * r * r1 * s * sh;
        t[7] = -8.0 * r1 * rh * s * s1;
        t[8] = 16.0 * r * r1 * s * s1;
        const double a = 2.0 * r - 1.5, b = 2.0 * r - 0.5, c = 2.0 * r - 1.0;
        dr[0] = 4.0 * a * s1 * sh;
        dr[1] = 4.0 * b * s1 * sh;
        dr[2] = 4.0 * b * s * sh;
        dr[3] = 4.0 * a * s * sh;
        dr[4] = -8.0 * c * s1 * sh;
        dr[5] = -8.0 * b * s * s1;
        dr[6] = -8.0 * c * s * sh;
        dr[7] = -8.0 * a * s * s1;
        dr[8] = 16.0 * c * s * s1;
        const double p = 2.0 * s - 1.5, q = 2.0 * s - 0.5, u = 2.0 * s - 1.0;
        ds[0] = 4.0 * r1 * rh * p;
        ds[1] = 4.0 * r * rh * p;
        ds[2] = 4.0 * r * rh * q;
        ds[3] = 4.0 * r1 * rh * q;
        ds[4] = -8.0 * r * r1 * p;
        ds[5] = -8.0 * r * rh * u;
        ds[6] = -8.0 * r * r1 * q;
        ds[7] = -8.0 * r1 * rh * u;
        ds[8] = 16.0 * r * r1 * u;
        break;
    }
    }
}

/* ---- per-element geometry + basis at every quadrature point --------------- */

typedef struct {
    int order, nq;
    double area;
    double wq[36];                       /* rule weight per point */
    double phi[36][9], dx[36][9], dy[36][9];
} elem_basis;

static int is_quad(int type) { return type == SFH_Q4 || type == SFH_Q9; }

/* Fills b for element e.  Returns SFH_E_AREA for a zero-area element. */
static int element_basis(int type, const int *en, int e, const double *X, const double *Y,
                         elem_basis *b)
{
    const int o = sfh_element_order(type);
    const int *nd = en + (size_t)e * o;
    b->order = o;
    const double x1 = X[nd[0] - 1], y1 = Y[nd[0] - 1];
    const double x2 = X[nd[1] - 1], y2 = Y[nd[1] - 1];
    const double x3 = X[nd[2] - 1], y3 = Y[nd[2] - 1];
    double t[9], tr[9], ts[9];
    if (is_quad(type)) {
        const double x4 = X[nd[3] - 1], y4 = Y[nd[3] - 1];
        /* two corner triangles; the area replaces |J| in the weights */
        b->area = 0.5 * fabs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2)) +
                  0.5 * fabs(x1 * (y4 - y3) + x4 * (y3 - y1) + x3 * (y1 - y4));
        if (b->area == 0.0) return SFH_E_AREA;
        b->nq = 36;
        for (int q = 0; q < 36; ++q) {
            const double r = G6[q / 6], s = G6[q % 6];
            /* bilinear map through the four corner nodes */
            const double xr = -(1.0 - s) * x1 + (1.0 - s) * x2 + s * x3 - s * x4;
            const double xs = -(1.0 - r) * x1 - r * x2 + r * x3 + (1.0 - r) * x4;
            const double yr = -(1.0 - s) * y1 + (1.0 - s) * y2 + s * y3 - s * y4;
            const double ys = -(1.0 - r) * y1 - r * y2 + r * y3 + (1.0 - r) * y4;
            const double det = xr * ys - xs * yr;
            const double rx = ys / det, ry = -xs / det, sx = -yr / det, sy = xr / det;
            shape_eval(type, r, s, t, tr, ts);
            b->wq[q] = W6x6[q];
            for (int k = 0; k < o; ++k) {
                b->phi[q][k] = t[k];
                b->dx[q][k] = tr[k] * rx + ts[k] * sx;
                b->dy[q][k] = tr[k] * ry + ts[k] * sy;
            }
        }
    } else {
        b->area = 0.5 * fabs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
        if (b->area == 0.0) return SFH_E_AREA;
        /* affine map: constant inverse Jacobian */
        const double det = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
        const double rx = (y3 - y1) / det, ry = (x1 - x3) / det;
        const double sx = (y1 - y2) / det, sy = (x2 - x1) / det;
        b->nq = 12;
        for (int q = 0; q < 12; ++q) {
            shape_eval(type, T12r[q], T12s[q], t, tr, ts);
            b->wq[q] = T12w[q];
            for (int k = 0; k < o; ++k) {
                b->phi[q][k] = t[k];
                b->dx[q][k] = tr[k] * rx + ts[k] * sx;
                b->dy[q][k] = tr[k] * ry + ts[k] * sy;
            }
        }
    }
    return SFH_OK;
}

int sfh_assemble_coo(int type, int node_num, int element_num, const int *en, const double *X,
                     const double *Y, const double *rho, int *ci, int *cj, double *const cx[7])
{
    const int o = sfh_element_order(type);
    if (o < 0 || !en || !X || !Y || element_num <= 0 || node_num <= 0) return SFH_E_ARG;
    if (cx && cx[0] && !rho) return SFH_E_ARG;
    int status = SFH_OK;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < element_num; ++e) {
        elem_basis b;
        if (element_basis(type, en, e, X, Y, &b) != SFH_OK) {
#pragma omp atomic write
            status = SFH_E_AREA;
            continue;
        }
        const int *nd = en + (size_t)e * o;
        const size_t base = (size_t)e * o * o;
        for (int iq = 0; iq < o; ++iq) {
            const int ip = nd[iq] - 1;
            for (int jq = 0; jq < o; ++jq) {
                const size_t k = base + (size_t)iq * o + jq;
                if (ci) ci[k] = ip;
                if (cj) cj[k] = nd[jq] - 1;
                /* per-slot sums over the points, in point order */
                double m = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0, a5 = 0.0, a6 = 0.0;
                for (int q = 0; q < b.nq; ++q) {
                    const double aw = b.area * b.wq[q];
                    if (cx && cx[0]) m = m + aw * rho[ip] * b.phi[q][iq] * b.phi[q][jq];
                    a1 = a1 + aw * (b.dx[q][iq] * b.dx[q][jq]);
                    a2 = a2 + aw * (b.dy[q][iq] * b.dy[q][jq]);
                    a3 = a3 + aw * (b.dx[q][iq] * b.dy[q][jq]);
                    a4 = a4 + aw * (b.dy[q][iq] * b.dx[q][jq]);
                    a5 = a5 + aw * (b.phi[q][iq] * b.dx[q][jq]);
                    a6 = a6 + aw * (b.phi[q][iq] * b.dy[q][jq]);
                }
                if (cx) {
                    if (cx[0]) cx[0][k] = m;
                    if (cx[1]) cx[1][k] = a1;
                    if (cx[2]) cx[2][k] = a2;
                    if (cx[3]) cx[3][k] = a3;
                    if (cx[4]) cx[4][k] = a4;
                    if (cx[5]) cx[5][k] = a5;
                    if (cx[6]) cx[6][k] = a6;
                }
            }
        }
    }
    return status;
}

int sfh_lumped_mass(int type, int node_num, int element_num, const int *en, const double *X,
                    const double *Y, const double *rho, double *mass)
{
    const int o = sfh_element_order(type);
    if (o < 0 || !en || !X || !Y || !rho || !mass || element_num <= 0) return SFH_E_ARG;
    /* Terms are produced element-parallel into a chunk buffer, then summed in
     * the reference's global order (element, point, i, j). */
    const int chunk = 4096;
    elem_basis *bb = (elem_basis *)malloc(sizeof(elem_basis) * (size_t)chunk);
    int *bad = (int *)malloc(sizeof(int) * (size_t)chunk);
    if (!bb || !bad) { free(bb); free(bad); return SFH_E_NOMEM; }
    for (int i = 0; i < node_num; ++i) mass[i] = 0.0;
    double mass_sum = 0.0, diag_sum = 0.0;
    int status = SFH_OK;
    for (int e0 = 0; e0 < element_num && status == SFH_OK; e0 += chunk) {
        const int ne = (element_num - e0 < chunk) ? element_num - e0 : chunk;
#pragma omp parallel for schedule(static)
        for (int k = 0; k < ne; ++k) bad[k] = element_basis(type, en, e0 + k, X, Y, &bb[k]);
        for (int k = 0; k < ne; ++k) {
            if (bad[k] != SFH_OK) { status = bad[k]; break; }
            const elem_basis *b = &bb[k];
            const int *nd = en + (size_t)(e0 + k) * o;
            for (int q = 0; q < b->nq; ++q) {
                const double aw = b->area * b->wq[q];
                for (int iq = 0; iq < o; ++iq) {
                    const int ip = nd[iq] - 1;
                    const double f = aw * rho[ip] * b->phi[q][iq];
                    for (int jq = 0; jq < o; ++jq) {
                        const double v = f * b->phi[q][jq];
                        mass_sum = mass_sum + v;
                        if (ip == nd[jq] - 1) {
                            diag_sum = diag_sum + v;
                            mass[ip] = mass[ip] + v;
                        }
                    }
                }
            }
        }
    }
    free(bb);
    free(bad);
    if (status != SFH_OK) return status;
    for (int i = 0; i < node_num; ++i) mass[i] = mass[i] * mass_sum / diag_sum;
    return SFH_OK;
}

/* ---- canonical COO -> CSR ------------------------------------------------ */

/* Selection sort on v[0..n): repeatedly swaps the first minimum of the tail
 * into place (coo2csr_lib.f90:256-267). */
static void selection_sort_map(int *v, int *m, int n)
{
    for (int i = 0; i < n - 1; ++i) {
        int best = i;
        for (int k = i + 1; k < n; ++k)
            if (v[k] < v[best]) best = k;
        if (best != i) {
            int t = v[i]; v[i] = v[best]; v[best] = t;
            t = m[i]; m[i] = m[best]; m[best] = t;
        }
    }
}

/* Hoare-style partition sort with an explicit range stack and a selection
 * sort below 21 elements; produces the same permutation as
 * coo2csr_lib.f90:191-246 (values are compared as 1-based column ids). */
static void argsort_columns(int *v, int *m, int n)
{
    enum { MAX_LEVELS = 300, SMALL = 20 };
    int lo[MAX_LEVELS + 1], hi[MAX_LEVELS + 1];
    for (int i = 0; i < n; ++i) m[i] = i;
    int top = 0;
    lo[0] = 0;
    hi[0] = n - 1;
    while (top >= 0) {
        int left = lo[top], right = hi[top];
        if (right - left < SMALL) {
            if (left < right) selection_sort_map(v + left, m + left, right - left + 1);
            --top;
            continue;
        }
        const int pivot = (v[left] + v[right]) / 2;
        --left;
        ++right;
        for (;;) {
            do { ++left; } while (v[left] < pivot);
            do { --right; } while (v[right] > pivot);
            if (left < right) {
                int t = v[left]; v[left] = v[right]; v[right] = t;
                t = m[left]; m[left] = m[right]; m[right] = t;
            } else {
                if (left == right) {
                    if (left == lo[top]) ++left; else --right;
                }
                break;
            }
        }
        if (top + 1 > MAX_LEVELS) { /* cannot happen for element rows; keep sorted anyway */
            selection_sort_map(v + lo[top], m + lo[top], hi[top] - lo[top] + 1);
            --top;
            continue;
        }
        hi[top + 1] = hi[top];
        lo[top + 1] = left;
        hi[top] = right;
        ++top;
    }
}

int sfh_coo2csr(int nnz, int node_num, const int *ai, const int *aj, const double *ax, int *bp,
                int *bj, double *bx, int *csr_size)
{
    if (nnz < 0 || node_num <= 0 || !bp || !bj || !bx) return SFH_E_ARG;
    int *cnt = (int *)calloc((size_t)node_num + 1, sizeof(int));
    int *pos = (int *)malloc(sizeof(int) * ((size_t)node_num + 1));
    int *tj = (int *)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    double *tx = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    int *kept = (int *)malloc(sizeof(int) * ((size_t)node_num + 1));
    if (!cnt || !pos || !tj || !tx || !kept) {
        free(cnt); free(pos); free(tj); free(tx); free(kept);
        return SFH_E_NOMEM;
    }
    for (int n = 0; n < nnz; ++n) {
        if (ai[n] < 0 || ai[n] >= node_num) {
            free(cnt); free(pos); free(tj); free(tx); free(kept);
            return SFH_E_ARG;
        }
        cnt[ai[n] + 1]++;
    }
    for (int i = 0; i < node_num; ++i) cnt[i + 1] += cnt[i];
    memcpy(pos, cnt, sizeof(int) * ((size_t)node_num + 1));
    /* stable scatter: each row keeps its COO order; columns stored 1-based */
    for (int n = 0; n < nnz; ++n) {
        const int d = pos[ai[n]]++;
        tj[d] = aj[n] + 1;
        tx[d] = ax[n];
    }
    /* per row: argsort, gather, sum duplicates, drop exact zeros (in place) */
#pragma omp parallel
    {
        int cap = 64;
        int *v = (int *)malloc(sizeof(int) * cap), *m = (int *)malloc(sizeof(int) * cap);
        double *xv = (double *)malloc(sizeof(double) * cap);
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < node_num; ++i) {
            const int r0 = cnt[i], len = cnt[i + 1] - cnt[i];
            if (len > cap) {
                cap = 2 * len;
                v = (int *)realloc(v, sizeof(int) * cap);
                m = (int *)realloc(m, sizeof(int) * cap);
                xv = (double *)realloc(xv, sizeof(double) * cap);
            }
            memcpy(v, tj + r0, sizeof(int) * (size_t)len);
            argsort_columns(v, m, len);
            for (int k = 0; k < len; ++k) xv[k] = tx[r0 + m[k]];
            int w = 0;
            for (int k = 0; k < len;) {
                const int col = v[k];
                double s = xv[k++];
                while (k < len && v[k] == col) s = s + xv[k++];
                if (s != 0.0) { /* -0.0 is dropped as well */
                    tj[r0 + w] = col - 1;
                    tx[r0 + w] = s;
                    ++w;
                }
            }
            kept[i] = w;
        }
        free(v); free(m); free(xv);
    }
    bp[0] = 0;
    for (int i = 0; i < node_num; ++i) {
        memmove(bj + bp[i], tj + cnt[i], sizeof(int) * (size_t)kept[i]);
        memmove(bx + bp[i], tx + cnt[i], sizeof(double) * (size_t)kept[i]);
        bp[i + 1] = bp[i] + kept[i];
    }
    if (csr_size) *csr_size = bp[node_num];
    free(cnt); free(pos); free(tj); free(tx); free(kept);
    return SFH_OK;
}

int sfh_csr_union(int node_num, int nmat, const int *const *bp, const int *const *bj,
                  const double *const *bx, int max_nnz, int *up, int *uj, double *const *ux)
{
    if (node_num <= 0 || nmat <= 0 || nmat > 16 || !up || !uj) return SFH_E_ARG;
    up[0] = 0;
    int w = 0;
    for (int i = 0; i < node_num; ++i) {
        int cur[16];
        for (int m = 0; m < nmat; ++m) cur[m] = bp[m][i];
        for (;;) {
            /* smallest pending column over all matrices */
            int col = -1;
            for (int m = 0; m < nmat; ++m)
                if (cur[m] < bp[m][i + 1] && (col < 0 || bj[m][cur[m]] < col)) col = bj[m][cur[m]];
            if (col < 0) break;
            if (w >= max_nnz) return SFH_E_ARG;
            uj[w] = col;
            for (int m = 0; m < nmat; ++m) {
                double val = 0.0;
                if (cur[m] < bp[m][i + 1] && bj[m][cur[m]] == col) val = bx[m][cur[m]++];
                if (ux && ux[m]) ux[m][w] = val;
            }
            ++w;
        }
        up[i + 1] = w;
    }
    return w;
}

/* The reference's Fortran entry point under its gfortran name, for callers
 * linked against coo2csr_lib.o (elastic_wave.c:241-313 call it seven times).
 * Everything by pointer, as Fortran passes it; semantics of
 * coo2csr_lib.f90:6-41: the outputs are cleared (Bp, and Bj / Bx over all
 * nnz entries), Ai and Aj are incremented in place and stay so, the canonical
 * CSR of the 0-based COO input is written 0-based, *csr_size receives the
 * kept count, and the unused tail of Bj reads -1 (cleared to 0, then shifted
 * to 0-based with the rest) while that of Bx reads 0.  The reference has no
 * error path; here a row index outside 0..Bp_size-2 or a failed allocation
 * prints a message to stderr and sets *csr_size = -1, with Ai / Aj untouched.
 * The reference's progress lines (write(*,*)) are not printed. */
void __coo2csr_lib_MOD_coo2csr_canonical(int *nnz, int *Bp_size, int *csr_size, int *Ai, int *Aj,
                                          double *Ax, int *Bp, int *Bj, double *Bx)
{
    const int n = *nnz, np1 = *Bp_size;
    for (int i = 0; i < np1; ++i) Bp[i] = 0;
    for (int k = 0; k < n; ++k) {
        Bj[k] = 0;
        Bx[k] = 0.0;
    }
    int kept = 0;
    const int rc = sfh_coo2csr(n, np1 - 1, Ai, Aj, Ax, Bp, Bj, Bx, &kept);
    if (rc != SFH_OK) {
        fprintf(stderr, "coo2csr_canonical: %s (nnz %d, Bp_size %d)\n",
                rc == SFH_E_NOMEM ? "out of memory" : "row index out of range", n, np1);
        *csr_size = -1;
        return;
    }
    for (int k = kept; k < n; ++k) {
        Bj[k] = -1;
        Bx[k] = 0.0;
    }
    for (int k = 0; k < n; ++k) {
        Ai[k] += 1;
        Aj[k] += 1;
    }
    *csr_size = kept;
}
