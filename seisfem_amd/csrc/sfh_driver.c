/*
 * Drop-in driver: the reference's elastic_wave() and main() on top of the GPU
 * loop, reading and writing the reference's files.
 *
 *   elastic_wave(...)   same 26 arguments as time_evolution/elastic_wave.c:1-3.
 *                       Setup as :136-313 (host C, bit-identical arrays), the
 *                       shot loop :446-818 on the GPU with all shots batched as
 *                       right-hand sides, outputs in ./outputfile/ with the
 *                       reference's "%f   " / "%f\t" rows.
 *   seisfem_run(dir)    seisfem/seisfem.c:28-299 for a run directory holding
 *                       par.txt (and mesh/ for exterior meshes).
 *
 * Only the masslump solver is supported (solver_code 3); pardiso / mgmres are
 * not part of this library and return an error instead of running.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "seisfem_gpu.h"
#include "seisfem_host.h"

/* extra control of the drop-in, set before calling elastic_wave()/seisfem_run() */
static int g_precision = SFG_FP64;
static int g_binary = 0;
static int g_last_status = 0;

void sfh_driver_options(int precision, int binary_output)
{
    g_precision = precision;
    g_binary = binary_output;
}

int sfh_driver_status(void) { return g_last_status; }

static int type_code_of(const char *type)
{
    if (!type) return -1;
    if (!strcmp(type, "T3")) return SFH_T3;
    if (!strcmp(type, "T6")) return SFH_T6;
    if (!strcmp(type, "Q4")) return SFH_Q4;
    if (!strcmp(type, "Q9")) return SFH_Q9;
    return -1;
}

static void write_rows(FILE *f, const double *v, long rows, long cols, const char *fmt)
{
    for (long r = 0; r < rows; ++r) {
        for (long c = 0; c < cols; ++c) fprintf(f, fmt, v[r * cols + c]);
        fprintf(f, "\n");
    }
}

/* per-node model: velocity_and_density.txt (exterior mesh) or the reference's
 * built-in homogeneous model (model_elastic_parameter.c:34-45), which is also
 * written out as the reference does (:213-219) */
static int read_or_make_model(int node_num, int use_exterior_mesh, double *rho, double *vp, double *vs,
                              double *c[4])
{
    if (use_exterior_mesh == 1) {
        FILE *f = fopen("./mesh/velocity_and_density.txt", "r");
        if (!f) return SFH_E_IO;
        for (int i = 0; i < node_num; ++i)
            if (fscanf(f, "%lf %lf %lf %lf %lf %lf %lf", &rho[i], &vp[i], &vs[i], &c[0][i], &c[1][i],
                       &c[2][i], &c[3][i]) != 7) {
                fclose(f);
                return SFH_E_IO;
            }
        fclose(f);
        return SFH_OK;
    }
    for (int i = 0; i < node_num; ++i) {
        const double RHO = 2200.0, VP = 2000.0, VS = 1154.7;
        const double miu = VS * VS * RHO;
        const double lamda = VP * VP * RHO - 2 * miu;
        rho[i] = RHO;
        vp[i] = VP;
        vs[i] = VS;
        c[0][i] = lamda + 2 * miu;
        c[1][i] = lamda;
        c[2][i] = lamda + 2 * miu;
        c[3][i] = miu;
    }
    FILE *f = fopen("./mesh/velocity_and_density.txt", "w");
    if (f) {
        for (int i = 0; i < node_num; ++i)
            fprintf(f, "%f\t%f\t%f\t%f\t%f\t%f\t%f\n", rho[i], vp[i], vs[i], c[0][i], c[1][i], c[2][i], c[3][i]);
        fclose(f);
    }
    return SFH_OK;
}

#define FREE_ALL()                                                                                 \
    do {                                                                                           \
        for (int m_ = 0; m_ < 7; ++m_) { free(bp[m_]); free(bj[m_]); free(bx[m_]); }               \
        for (int m_ = 0; m_ < 7; ++m_) free(cx[m_]);                                               \
        for (int m_ = 0; m_ < 4; ++m_) free(c[m_]);                                                \
        for (int m_ = 0; m_ < 6; ++m_) free(pm[m_]);                                               \
        free(ci); free(cj); free(rho); free(vp); free(vs); free(mass_lump); free(flag);            \
        free(X); free(Y); free(seis_u); free(seis_w); free(snapU); free(snapW);                    \
    } while (0)

void elastic_wave(char *type, int node_num, int element_num, int element_order, int *element_node,
                  double **node_xy, int nnz, int csr_p_size, int step, double dt, double f0, double t0,
                  double edge_size, double xmin, double xmax, double ymin, double ymax, int pml_nx,
                  int pml_ny, int src_num, int *src_node, int rec_num, int *rec_node, char *solver,
                  int use_exterior_mesh, int free_surface_code)
{
    int *bp[7] = {0}, *bj[7] = {0}, *ci = NULL, *cj = NULL, *flag = NULL;
    double *bx[7] = {0}, *cx[7] = {0}, *c[4] = {0}, *pm[6] = {0};
    double *rho = NULL, *vp = NULL, *vs = NULL, *mass_lump = NULL, *X = NULL, *Y = NULL;
    double *seis_u = NULL, *seis_w = NULL, *snapU = NULL, *snapW = NULL;
    sfg_operator *op = NULL;
    sfg_medium *md = NULL;
    sfg_run *run = NULL;
    int rc = SFH_OK;
    const int tc = type_code_of(type);
    g_last_status = 0;
    if (tc < 0 || sfh_element_order(tc) != element_order || csr_p_size != node_num + 1 ||
        nnz != element_num * element_order * element_order) {
        fprintf(stderr, "ELASTIC_WAVE - unsupported element type or inconsistent sizes\n");
        g_last_status = SFH_E_ARG;
        return;
    }
    if (!solver || strcmp(solver, "masslump") != 0) {
        fprintf(stderr, "ELASTIC_WAVE - only the masslump solver runs on the GPU (got \"%s\")\n",
                solver ? solver : "(null)");
        g_last_status = SFH_E_ARG;
        return;
    }
    if (step < 3) {
        g_last_status = SFH_E_ARG;
        return;
    }
    const long n = node_num, steps = step - 2;
    X = (double *)malloc(sizeof(double) * n);
    Y = (double *)malloc(sizeof(double) * n);
    rho = (double *)malloc(sizeof(double) * n);
    vp = (double *)malloc(sizeof(double) * n);
    vs = (double *)malloc(sizeof(double) * n);
    mass_lump = (double *)malloc(sizeof(double) * n);
    flag = (int *)malloc(sizeof(int) * n);
    for (int k = 0; k < 4; ++k) c[k] = (double *)malloc(sizeof(double) * n);
    for (int k = 0; k < 6; ++k) pm[k] = (double *)malloc(sizeof(double) * n);
    ci = (int *)malloc(sizeof(int) * (size_t)nnz);
    cj = (int *)malloc(sizeof(int) * (size_t)nnz);
    for (int m = 0; m < 7; ++m) {
        cx[m] = (double *)malloc(sizeof(double) * (size_t)nnz);
        bp[m] = (int *)malloc(sizeof(int) * (n + 1));
        bj[m] = (int *)malloc(sizeof(int) * (size_t)nnz);
        bx[m] = (double *)malloc(sizeof(double) * (size_t)nnz);
    }
    memcpy(X, node_xy[0], sizeof(double) * n);
    memcpy(Y, node_xy[1], sizeof(double) * n);

    /* setup, elastic_wave.c:136-313 */
    sfh_dirichlet_flags(node_num, X, Y, edge_size, xmin, xmax, ymin, ymax, free_surface_code, flag);
    const int use_xmax = 1 - free_surface_code;
    rc = read_or_make_model(node_num, use_exterior_mesh, rho, vp, vs, c);
    if (rc) goto out;
    double vp_max = vp[0];
    for (long i = 0; i < n; ++i)
        if (vp[i] > vp_max) vp_max = vp[i];
    abc_mpml(node_num, element_num, element_order, element_node, node_xy, pml_nx, pml_ny, edge_size, xmin,
             xmax, ymin, ymax, vp_max, 1, use_xmax, 1, 1, pm[0], pm[1], pm[2], pm[3], pm[4], pm[5]);
    if (sfg_last_error()[0]) { rc = SFH_E_ARG; goto out; }
    if ((rc = sfh_lumped_mass(tc, node_num, element_num, element_node, X, Y, rho, mass_lump))) goto out;
    if ((rc = sfh_assemble_coo(tc, node_num, element_num, element_node, X, Y, rho, ci, cj, cx))) goto out;
    for (int m = 0; m < 7; ++m) {
        int sz = 0;
        if ((rc = sfh_coo2csr(nnz, node_num, ci, cj, cx[m], bp[m], bj[m], bx[m], &sz))) goto out;
    }

    /* the shot loop, elastic_wave.c:446-818, all shots at once */
    {
        const int *P[7], *J[7];
        const double *V[7], *C4[4], *PM[6];
        for (int m = 0; m < 7; ++m) { P[m] = bp[m]; J[m] = bj[m]; V[m] = bx[m]; }
        for (int k = 0; k < 4; ++k) C4[k] = c[k];
        for (int k = 0; k < 6; ++k) PM[k] = pm[k];
        if ((rc = sfg_operator_create(&op, node_num, P, J, V))) goto out;
        if ((rc = sfg_medium_create(&md, op, C4, mass_lump, flag, PM))) goto out;
        sfg_shot_desc d;
        memset(&d, 0, sizeof d);
        d.precision = g_precision;
        d.nshots = src_num;
        d.src_node = src_node;
        d.max_steps = (int)steps;
        d.rec_num = rec_num;
        d.rec_node = rec_node;
        d.dt = dt;
        d.f0 = f0;
        d.t0 = t0;
        d.amplitude = 1.0e10;  /* elastic_wave.c:526 */
        d.angle_deg = 90.0;    /* elastic_wave.c:66 */
        d.use_graph = 1;
        if ((rc = sfg_run_create(&run, op, md, &d))) goto out;

        FILE **fu = (FILE **)calloc((size_t)src_num, sizeof(FILE *));
        FILE **fw = (FILE **)calloc((size_t)src_num, sizeof(FILE *));
        char name[256];
        for (int s = 0; s < src_num; ++s) {
            printf("\n ######## Shot num: %d ########\n", s + 1);
            snprintf(name, sizeof name, "./outputfile/wavefield_u_shot_%d.txt", s + 1);
            fu[s] = fopen(name, "w");
            snprintf(name, sizeof name, "./outputfile/wavefield_w_shot_%d.txt", s + 1);
            fw[s] = fopen(name, "w");
        }
        snapU = (double *)malloc(sizeof(double) * n);
        snapW = (double *)malloc(sizeof(double) * n);
        for (int s = 0; s < src_num; ++s) { /* initial zero row, :503-507 */
            for (long i = 0; i < n; ++i) {
                if (fu[s]) fprintf(fu[s], "%f\t", 0.0);
                if (fw[s]) fprintf(fw[s], "%f\t", 0.0);
            }
            if (fu[s]) fprintf(fu[s], "\n");
            if (fw[s]) fprintf(fw[s], "\n");
        }
        long done = 0;
        for (long t = 0; t < steps && rc >= 0; ++t) {
            const long it = t + 2;
            if ((it + 1) % 200 != 0) continue; /* snapshot marks, :797 */
            rc = sfg_run_advance(run, (int)(t + 1 - done));
            if (rc < 0) break;
            done = t + 1;
            for (int s = 0; s < src_num && rc >= 0; ++s) {
                rc = sfg_run_snapshot(run, s, snapU, snapW);
                if (rc < 0) break;
                if (fu[s]) write_rows(fu[s], snapU, 1, n, "%f\t");
                if (fw[s]) write_rows(fw[s], snapW, 1, n, "%f\t");
            }
        }
        if (rc >= 0) rc = sfg_run_advance(run, (int)(steps - done));
        for (int s = 0; s < src_num; ++s) {
            if (fu[s]) fclose(fu[s]);
            if (fw[s]) fclose(fw[s]);
        }
        free(fu);
        free(fw);
        if (rc < 0) {
            if (rc == SFG_E_ENERGY) fprintf(stderr, "\nELASTIC_MPML - Fatal error!\nEnergy exceeds maximum value!\n");
            goto out;
        }
        rc = SFH_OK;
        const size_t per = (size_t)steps * (rec_num > 0 ? rec_num : 1);
        seis_u = (double *)malloc(sizeof(double) * per * src_num);
        seis_w = (double *)malloc(sizeof(double) * per * src_num);
        if ((rc = sfg_run_seismograms(run, seis_u, seis_w))) goto out;
        for (int s = 0; s < src_num; ++s) {
            snprintf(name, sizeof name, "./outputfile/seismogram_u_shot_%d.txt", s + 1);
            FILE *f = fopen(name, "w");
            if (f) { write_rows(f, seis_u + per * s, steps, rec_num, "%f   "); fclose(f); }
            snprintf(name, sizeof name, "./outputfile/seismogram_w_shot_%d.txt", s + 1);
            f = fopen(name, "w");
            if (f) { write_rows(f, seis_w + per * s, steps, rec_num, "%f   "); fclose(f); }
            if (g_binary) { /* full-precision companion files */
                snprintf(name, sizeof name, "./outputfile/seismogram_uw_shot_%d.f64", s + 1);
                f = fopen(name, "wb");
                if (f) {
                    fwrite(seis_u + per * s, sizeof(double), per, f);
                    fwrite(seis_w + per * s, sizeof(double), per, f);
                    fclose(f);
                }
            }
        }
        printf("\nTime iteration end!\n");
    }
out:
    if (rc) {
        fprintf(stderr, "ELASTIC_WAVE - failed (%d): %s\n", rc, sfg_last_error());
        g_last_status = rc;
    }
    sfg_run_destroy(run);
    sfg_medium_destroy(md);
    sfg_operator_destroy(op);
    FREE_ALL();
}

/* ---- par.txt driver (seisfem/seisfem.c:28-299) ---------------------------- */

static int scan_key(FILE *f, const char *key, const char *fmt, void *out)
{
    char line[512];
    rewind(f);
    const size_t kl = strlen(key);
    while (fgets(line, sizeof line, f)) {
        const char *p = line;
        while (*p == ' ' || *p == '\t') ++p;
        if (strncmp(p, key, kl) == 0 && (p[kl] == ' ' || p[kl] == '=' || p[kl] == '\t')) {
            const char *eq = strchr(p, '=');
            if (eq && sscanf(eq + 1, fmt, out) == 1) return 0;
        }
    }
    return -1;
}

int seisfem_run(const char *workdir)
{
    if (workdir && chdir(workdir) != 0) return SFH_E_IO;
    FILE *f = fopen("par.txt", "r");
    if (!f) return SFH_E_IO;
    int use_ext = 0, type_code = 0, nelemx = 0, nelemy = 0, src_num = 0, rec_num = 0, step = 0, solver_code = 0,
        free_surface = 0;
    double edge = 0, f0 = 0, t0 = 0, sxf = 0, syf = 0, sxl = 0, syl = 0, rxf = 0, ryf = 0, rxl = 0, ryl = 0, dt = 0;
    int bad = 0;
    bad |= scan_key(f, "use_exterior_mesh", "%d", &use_ext);
    bad |= scan_key(f, "type_code", "%d", &type_code);
    bad |= scan_key(f, "nelemx", "%d", &nelemx);
    bad |= scan_key(f, "nelemy", "%d", &nelemy);
    bad |= scan_key(f, "edge_size", "%lf", &edge);
    bad |= scan_key(f, "f0", "%lf", &f0);
    bad |= scan_key(f, "t0", "%lf", &t0);
    bad |= scan_key(f, "src_num", "%d", &src_num);
    bad |= scan_key(f, "src_x_first", "%lf", &sxf);
    bad |= scan_key(f, "src_y_first", "%lf", &syf);
    bad |= scan_key(f, "src_x_last", "%lf", &sxl);
    bad |= scan_key(f, "src_y_last", "%lf", &syl);
    bad |= scan_key(f, "rec_num", "%d", &rec_num);
    bad |= scan_key(f, "rec_x_first", "%lf", &rxf);
    bad |= scan_key(f, "rec_y_first", "%lf", &ryf);
    bad |= scan_key(f, "rec_x_last", "%lf", &rxl);
    bad |= scan_key(f, "rec_y_last", "%lf", &ryl);
    bad |= scan_key(f, "dt", "%lf", &dt);
    bad |= scan_key(f, "step", "%d", &step);
    bad |= scan_key(f, "solver_code", "%d", &solver_code);
    bad |= scan_key(f, "free_surface", "%d", &free_surface);
    fclose(f);
    if (bad) return SFH_E_IO;
    const int order = sfh_element_order(type_code);
    if (order < 0 || solver_code != 3 || src_num < 1 || rec_num < 1) return SFH_E_ARG;
    const char *tname = type_code == SFH_T3 ? "T3" : type_code == SFH_T6 ? "T6" : type_code == SFH_Q4 ? "Q4" : "Q9";

    int node_num = 0, element_num = 0;
    int *en = NULL;
    double *xy[2] = {NULL, NULL};
    double xmin = 0, xmax = 0, ymin = 0, ymax = 0;
    if (use_ext == 1) {
        FILE *fn = fopen("./mesh/element_and_node_num.txt", "r");
        if (!fn) return SFH_E_IO;
        if (fscanf(fn, "element_num = %d\n", &element_num) != 1 || fscanf(fn, "node_num = %d\n", &node_num) != 1) {
            fclose(fn);
            return SFH_E_IO;
        }
        fclose(fn);
        en = (int *)malloc(sizeof(int) * (size_t)element_num * order);
        xy[0] = (double *)malloc(sizeof(double) * node_num);
        xy[1] = (double *)malloc(sizeof(double) * node_num);
        FILE *fe = fopen("./mesh/element_node.txt", "r"), *fx = fopen("./mesh/node_xy.txt", "r");
        int ok = fe && fx;
        for (long k = 0; ok && k < (long)element_num * order; ++k) ok = fscanf(fe, "%d", &en[k]) == 1;
        for (int i = 0; ok && i < node_num; ++i) ok = fscanf(fx, "%lf %lf", &xy[0][i], &xy[1][i]) == 2;
        if (fe) fclose(fe);
        if (fx) fclose(fx);
        if (!ok) { free(en); free(xy[0]); free(xy[1]); return SFH_E_IO; }
        for (int i = 0; i < node_num; ++i) { /* bbox starts at 0, seisfem.c:199-214 */
            if (xy[0][i] < xmin) xmin = xy[0][i];
            if (xy[0][i] > xmax) xmax = xy[0][i];
            if (xy[1][i] < ymin) ymin = xy[1][i];
            if (xy[1][i] > ymax) ymax = xy[1][i];
        }
    } else {
        xmax = edge * nelemx;
        ymax = edge * nelemy;
        node_num = sfh_mesh_node_num(type_code, nelemx, nelemy);
        element_num = sfh_mesh_element_num(type_code, nelemx, nelemy);
        if (node_num <= 0 || element_num <= 0) return SFH_E_ARG;
        en = (int *)malloc(sizeof(int) * (size_t)element_num * order);
        xy[0] = (double *)malloc(sizeof(double) * node_num);
        xy[1] = (double *)malloc(sizeof(double) * node_num);
        sfh_mesh_element(type_code, nelemx, nelemy, en);
        sfh_mesh_xy(type_code, nelemx, nelemy, xmin, xmax, ymin, ymax, xy[0], xy[1]);
    }
    int *rec = (int *)malloc(sizeof(int) * rec_num), *src = (int *)malloc(sizeof(int) * src_num);
    double *rx = (double *)malloc(sizeof(double) * rec_num), *ry = (double *)malloc(sizeof(double) * rec_num);
    double *sx = (double *)malloc(sizeof(double) * src_num), *sy = (double *)malloc(sizeof(double) * src_num);
    sfh_station_line(rec_num, rxf, ryf, rxl, ryl, rx, ry);
    sfh_station_line(src_num, sxf, syf, sxl, syl, sx, sy);
    int status = SFH_OK;
    FILE *fr = fopen("./outputfile/receiver_station.txt", "w");
    for (int i = 0; i < rec_num; ++i) {
        rec[i] = sfh_node_location(node_num, edge, xy[0], xy[1], rx[i], ry[i]);
        if (rec[i] < 0) status = SFH_E_NOTFOUND;
        else if (fr) fprintf(fr, "%f\t%f\n", xy[0][rec[i]], xy[1][rec[i]]);
    }
    if (fr) fclose(fr);
    FILE *fs = fopen("./outputfile/source_location.txt", "w");
    for (int i = 0; i < src_num; ++i) {
        src[i] = sfh_node_location(node_num, edge, xy[0], xy[1], sx[i], sy[i]);
        if (src[i] < 0) status = SFH_E_NOTFOUND;
        else if (fs) fprintf(fs, "%lf\t%lf\n", xy[0][src[i]], xy[1][src[i]]);
    }
    if (fs) fclose(fs);
    if (status == SFH_OK) {
        elastic_wave((char *)tname, node_num, element_num, order, en, xy, element_num * order * order,
                     node_num + 1, step, dt, f0, t0, edge, xmin, xmax, ymin, ymax, 20, 20, src_num, src, rec_num,
                     rec, (char *)"masslump", use_ext, free_surface);
        status = g_last_status;
    }
    free(en); free(xy[0]); free(xy[1]); free(rec); free(src); free(rx); free(ry); free(sx); free(sy);
    return status;
}
