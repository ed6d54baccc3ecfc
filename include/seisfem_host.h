/*
 * seisfem_host.h -- host-side setup that feeds the GPU time loop.
 *
 * These routines rebuild, on the host, exactly the arrays the reference's
 * elastic_wave() consumes: structured meshes, element COO matrices, the HRZ
 * lumped mass, the canonical CSR matrices (bit-identical to the reference's
 * Fortran coo2csr_canonical), station nodes and Dirichlet flags.  All arrays
 * use the reference layouts: 1-based element_node (element-major), node_xy as
 * two separate coordinate arrays, 0-based CSR with ascending columns.
 *
 * Every function returns 0 on success and a negative SFH_E* code on error;
 * none of them calls exit().
 */
#ifndef SEISFEM_HOST_H
#define SEISFEM_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* element type codes, as in par.txt type_code (reference mesh/element_type.c) */
enum { SFH_T3 = 1, SFH_T6 = 2, SFH_Q4 = 4, SFH_Q9 = 5 };

/* error codes */
enum {
    SFH_OK = 0,
    SFH_E_ARG = -1,        /* bad argument (unknown element type, size, NULL) */
    SFH_E_AREA = -2,       /* zero-area element (reference prints + exit(1)) */
    SFH_E_IO = -3,         /* file missing / malformed */
    SFH_E_NOMEM = -4,
    SFH_E_NOTFOUND = -5    /* station outside the mesh */
};

/* element order (nodes per element), or -1.  Ref: mesh/mesh_element_order.c */
int sfh_element_order(int type_code);
/* node count of a structured nelemx x nelemy grid.  Ref: mesh/mesh_node_num.c */
int sfh_mesh_node_num(int type_code, int nelemx, int nelemy);
/* element count of a structured grid.  Ref: mesh/mesh_element_num.c */
int sfh_mesh_element_num(int type_code, int nelemx, int nelemy);

/* Structured connectivity, 1-based, element-major [element_num*order].
 * Ref: mesh/mesh_element_{t3,t6,q4,q9}.c */
int sfh_mesh_element(int type_code, int nelemx, int nelemy, int *element_node);
/* Structured coordinates on [xmin,xmax]x[ymin,ymax].
 * Ref: mesh/mesh_xy_{t3,t6,q4,q9}.c (the quadratic generators ignore xmin/ymin
 * exactly as the reference does). */
int sfh_mesh_xy(int type_code, int nelemx, int nelemy, double xmin, double xmax,
                double ymin, double ymax, double *node_x, double *node_y);

/* Element COO matrices for the 7 operators in reference order
 *   0: M  (consistent mass, rho sampled at the row node)
 *   1: S1 = dphidx*dphidx   2: S2 = dphidy*dphidy   3: S3 = dphidx*dphidy
 *   4: S4 = dphidy*dphidx   5: S5 = phi*dphidx      6: S6 = phi*dphidy
 * coo_x[m] (m = 0..6, NULL to skip) receives element_num*order^2 values at
 * coo_index = e*order^2 + iq*order + jq; coo_i/coo_j (may be NULL) receive
 * the 0-based row/column node.  Ref: assemble/mass_sparse_*.c (lumpflag=0),
 * assemble/stiffness_sparse_*.c (stif_type 1..6). */
int sfh_assemble_coo(int type_code, int node_num, int element_num, const int *element_node,
                     const double *node_x, const double *node_y, const double *rho,
                     int *coo_i, int *coo_j, double *const coo_x[7]);

/* HRZ lumped mass [node_num]: diagonal mass scaled by the global ratio
 * sum(all entries)/sum(diagonal entries).  Ref: assemble/mass_sparse_*.c, lumpflag=1. */
int sfh_lumped_mass(int type_code, int node_num, int element_num, const int *element_node,
                    const double *node_x, const double *node_y, const double *rho,
                    double *mass_lump);

/* COO -> canonical CSR (count, scatter in COO order, per-row argsort replaying
 * the reference's quicksort / selection sort, sequential duplicate sum, drop of
 * exact zeros).  Output 0-based.  bj/bx must hold nnz_coo entries; *csr_size
 * receives the kept count.  Inputs are not modified (the reference adds 1 to
 * them in place).  Ref: sparse_matrix/coo2csr_lib.f90:6-161,191-267. */
int sfh_coo2csr(int nnz_coo, int node_num, const int *ai, const int *aj, const double *ax,
                int *bp, int *bj, double *bx, int *csr_size);

/* Drop-in for the reference's Fortran module procedure under its gfortran
 * symbol (sparse_matrix/coo2csr_lib.f90:6-41, called by pointer at
 * time_evolution/elastic_wave.c:241-313), so a caller without a Fortran
 * compiler links this library instead of coo2csr_lib.o.  Same arguments and
 * effects: clears Bp (Bp_size entries) and Bj / Bx (nnz entries), adds 1 to
 * Ai and Aj in place, writes the canonical CSR 0-based (the sfh_coo2csr
 * result, bit for bit) with the unused tail of Bj at -1 and of Bx at 0, and
 * sets *csr_size.  A row index outside 0..*Bp_size-2 (where the reference
 * writes out of bounds) prints to stderr and sets *csr_size = -1. */
void __coo2csr_lib_MOD_coo2csr_canonical(int *nnz, int *Bp_size, int *csr_size, int *Ai, int *Aj,
                                          double *Ax, int *Bp, int *Bj, double *Bx);

/* Node search: the LAST node within edge_size/2 in x and y, or -1.
 * Ref: source_receiver/node_location.c:1-28 */
int sfh_node_location(int node_num, double edge_size, const double *node_x,
                      const double *node_y, double x, double y);
/* Linear station spacing.  Ref: source_receiver/set_rec_and_src.c */
void sfh_station_line(int num, double x_first, double y_first, double x_last, double y_last,
                      double *x, double *y);

/* Dirichlet clamp flags (1 = clamp), zero-initialised.
 * Ref: time_evolution/elastic_wave.c:136-155 */
void sfh_dirichlet_flags(int node_num, const double *node_x, const double *node_y,
                         double edge_size, double xmin, double xmax, double ymin, double ymax,
                         int free_surface, int *flag);

/* Locality order of the nodes: node ids sorted along a Hilbert curve over
 * their coordinates (quantised to a 2^16 x 2^16 grid of square cells, ties by
 * id).  Stepping rows in this order keeps the rows in flight on the GPU a
 * compact patch whose neighbours stay in L2 (sfg_shot_desc.row_order).  No
 * counterpart in the reference, which steps rows in node order. */
int sfh_locality_order(int node_num, const double *x, const double *y, int *order);

/* Union of several CSR patterns with the values of each matrix scattered
 * onto it (0.0 where a matrix has no entry).  up must hold node_num+1 ints;
 * uj/ux_m must hold max_nnz entries.  Returns the union nnz or a negative code. */
int sfh_csr_union(int node_num, int nmat, const int *const *bp, const int *const *bj,
                  const double *const *bx, int max_nnz, int *up, int *uj, double *const *ux);

#ifdef __cplusplus
}
#endif
#endif
