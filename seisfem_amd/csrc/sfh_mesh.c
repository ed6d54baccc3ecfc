/*
 * Structured meshes, station search and Dirichlet flags (host side).
 *
 * Node numbering, element node order and coordinate formulas follow the
 * reference generators so that node ids (and hence CSR row/column ids) are
 * identical:
 *   mesh/mesh_element_q4.c:76-93, mesh/mesh_element_q9.c:95-112,
 *   mesh/mesh_element_t3.c:75-100, mesh/mesh_element_t6.c:96-123,
 *   mesh/mesh_xy_q4.c:38-47 (same for t3), mesh/mesh_xy_q9.c:42-51 (same for t6).
 * Build with -ffp-contract=off: coordinates must round exactly as the reference.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include "seisfem_host.h"

int sfh_element_order(int type_code)
{
    switch (type_code) {
    case SFH_T3: return 3;
    case SFH_T6: return 6;
    case SFH_Q4: return 4;
    case SFH_Q9: return 9;
    default: return -1;
    }
}

int sfh_mesh_node_num(int type_code, int nelemx, int nelemy)
{
    if (nelemx <= 0 || nelemy <= 0) return -1;
    switch (type_code) {
    case SFH_T3: case SFH_Q4: return (nelemx + 1) * (nelemy + 1);
    case SFH_T6: case SFH_Q9: return (2 * nelemx + 1) * (2 * nelemy + 1);
    default: return -1;
    }
}

int sfh_mesh_element_num(int type_code, int nelemx, int nelemy)
{
    if (nelemx <= 0 || nelemy <= 0) return -1;
    switch (type_code) {
    case SFH_T3: case SFH_T6: return 2 * nelemx * nelemy;
    case SFH_Q4: case SFH_Q9: return nelemx * nelemy;
    default: return -1;
    }
}

int sfh_mesh_element(int type_code, int nelemx, int nelemy, int *en)
{
    if (!en || nelemx <= 0 || nelemy <= 0) return SFH_E_ARG;
    int e = 0;
    if (type_code == SFH_Q4 || type_code == SFH_T3) {
        /* corner lattice with (nelemx+1) nodes per line, 1-based */
        const int line = nelemx + 1;
        for (int j = 0; j < nelemy; ++j)
            for (int i = 0; i < nelemx; ++i) {
                int sw = 1 + i + j * line, se = sw + 1;
                int nw = sw + line, ne = nw + 1;
                if (type_code == SFH_Q4) {
                    int *o = en + 4 * e++;
                    o[0] = sw; o[1] = se; o[2] = ne; o[3] = nw;
                } else { /* two triangles split on the SE-NW diagonal */
                    int *a = en + 3 * e++;
                    a[0] = sw; a[1] = se; a[2] = nw;
                    int *b = en + 3 * e++;
                    b[0] = ne; b[1] = nw; b[2] = se;
                }
            }
        return SFH_OK;
    }
    if (type_code == SFH_Q9 || type_code == SFH_T6) {
        /* quadratic lattice with (2*nelemx+1) nodes per line, 1-based */
        const int line = 2 * nelemx + 1;
        for (int j = 0; j < nelemy; ++j)
            for (int i = 0; i < nelemx; ++i) {
                int sw = 1 + 2 * i + 2 * j * line;
                int s = sw + 1, se = sw + 2;
                int w = sw + line, c = w + 1, ee = w + 2;
                int nw = sw + 2 * line, n = nw + 1, ne = nw + 2;
                if (type_code == SFH_Q9) {
                    int *o = en + 9 * e++;
                    o[0] = sw; o[1] = se; o[2] = ne; o[3] = nw;
                    o[4] = s;  o[5] = ee; o[6] = n;  o[7] = w; o[8] = c;
                } else {
                    int *a = en + 6 * e++;
                    a[0] = sw; a[1] = se; a[2] = nw; a[3] = s; a[4] = c; a[5] = w;
                    int *b = en + 6 * e++;
                    b[0] = ne; b[1] = nw; b[2] = se; b[3] = n; b[4] = c; b[5] = ee;
                }
            }
        return SFH_OK;
    }
    return SFH_E_ARG;
}

int sfh_mesh_xy(int type_code, int nx, int ny, double xmin, double xmax, double ymin,
                double ymax, double *X, double *Y)
{
    if (!X || !Y || nx <= 0 || ny <= 0) return SFH_E_ARG;
    if (type_code == SFH_Q4 || type_code == SFH_T3) {
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                int k = i + j * (nx + 1);
                /* weighted end-point form, integer weights converted exactly */
                X[k] = ((double)(nx - i) * xmin + (double)i * xmax) / (1.0 * nx);
                Y[k] = ((double)(ny - j) * ymin + (double)j * ymax) / (1.0 * ny);
            }
        return SFH_OK;
    }
    if (type_code == SFH_Q9 || type_code == SFH_T6) {
        const int line = 2 * nx + 1;
        for (int j = 0; j <= 2 * ny; ++j)
            for (int i = 0; i < line; ++i) {
                int k = i + j * line;
                /* half-spacing lattice; the reference drops xmin/ymin here */
                X[k] = (double)i * (xmax - xmin) / (2.0 * nx);
                Y[k] = (double)j * (ymax - ymin) / (2.0 * ny);
            }
        return SFH_OK;
    }
    return SFH_E_ARG;
}

int sfh_node_location(int node_num, double edge_size, const double *X, const double *Y,
                      double x, double y)
{
    int found = -1;
    const double tol = edge_size / 2.0;
    for (int i = 0; i < node_num; ++i)
        if (fabs(X[i] - x) <= tol && fabs(Y[i] - y) <= tol)
            found = i; /* the last match wins */
    return found;
}

void sfh_station_line(int num, double x_first, double y_first, double x_last, double y_last,
                      double *x, double *y)
{
    for (int i = 0; i < num; ++i) {
        if (num == 1) {
            x[i] = x_first;
            y[i] = y_first;
        } else {
            x[i] = x_first + i * (x_last - x_first) / (1.0 * (num - 1));
            y[i] = y_first + i * (y_last - y_first) / (1.0 * (num - 1));
        }
    }
}

void sfh_dirichlet_flags(int node_num, const double *X, const double *Y, double edge_size,
                         double xmin, double xmax, double ymin, double ymax, int free_surface,
                         int *flag)
{
    const double q = edge_size / 4.0;
    for (int i = 0; i < node_num; ++i) {
        int on_side = (X[i] <= xmin + q) || (Y[i] <= ymin + q) || (Y[i] >= ymax - q);
        /* with a free surface the x = xmax side is left unclamped */
        if (!free_surface && X[i] >= xmax - q) on_side = 1;
        flag[i] = on_side ? 1 : 0;
    }
}

/* ---- locality order of the nodes (no counterpart in the reference) -------- */

/* Index of cell (x, y) along the Hilbert curve that fills a 2^16 x 2^16 grid:
 * at each level from the top, the quadrant gives two bits of the index and
 * the coordinates are reflected / swapped into the quadrant's frame. */
static uint64_t hilbert_index(uint32_t x, uint32_t y)
{
    uint64_t d = 0;
    for (uint32_t s = 1u << 15; s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1u : 0u;
        const uint32_t ry = (y & s) ? 1u : 0u;
        d += (uint64_t)s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = (s - 1) - (x & (s - 1));
                y = (s - 1) - (y & (s - 1));
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
        x &= s - 1;
        y &= s - 1;
    }
    return d;
}

typedef struct {
    uint64_t key;
    int node;
} sfh_key_node;

static int cmp_key_node(const void *a, const void *b)
{
    const sfh_key_node *p = (const sfh_key_node *)a, *q = (const sfh_key_node *)b;
    if (p->key != q->key) return p->key < q->key ? -1 : 1;
    return (p->node > q->node) - (p->node < q->node);
}

int sfh_locality_order(int node_num, const double *x, const double *y, int *order)
{
    if (node_num <= 0 || !x || !y || !order) return SFH_E_ARG;
    double x0 = x[0], x1 = x[0], y0 = y[0], y1 = y[0];
    for (int i = 1; i < node_num; ++i) {
        if (x[i] < x0) x0 = x[i];
        if (x[i] > x1) x1 = x[i];
        if (y[i] < y0) y0 = y[i];
        if (y[i] > y1) y1 = y[i];
    }
    /* one scale for both axes keeps the cells square */
    const double span = fmax(x1 - x0, y1 - y0);
    const double sc = span > 0.0 ? 65535.0 / span : 0.0;
    sfh_key_node *kn = (sfh_key_node *)malloc(sizeof(sfh_key_node) * (size_t)node_num);
    if (!kn) return SFH_E_NOMEM;
    for (int i = 0; i < node_num; ++i) {
        const uint32_t cx = (uint32_t)fmin(65535.0, fmax(0.0, (x[i] - x0) * sc + 0.5));
        const uint32_t cy = (uint32_t)fmin(65535.0, fmax(0.0, (y[i] - y0) * sc + 0.5));
        kn[i].key = hilbert_index(cx, cy);
        kn[i].node = i;
    }
    qsort(kn, (size_t)node_num, sizeof(sfh_key_node), cmp_key_node);
    for (int i = 0; i < node_num; ++i) order[i] = kn[i].node;
    free(kn);
    return SFH_OK;
}
