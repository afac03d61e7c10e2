/*
 * dg_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference solver's algorithm (penghaorui/Discontinuous-Galerkin-method),
 * used as the checker for the HIP path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (libdgx) never links, loads or calls this code.
 *
 * Layout follows the reference: element-major arrays [e*Np + i], per-element Dx/Dz [Tri*36],
 * Q [Tri*30], A/B [Tri*5], 1-based `long` indices.  Every function cites the reference file:line
 * it restates and keeps the reference's arithmetic order, so that with the same compiler flags
 * (gcc -O2, no FMA contraction) the results are bit-identical to the reference's own functions.
 *
 * Parity status of each function (how it is pinned) is listed in the oracle's test file
 * tests/test_oracle.py; the extensions marked [EXT] have no reference counterpart and are pinned
 * by closed forms (Lagrange basis, point location) and by their reduction to the reference case.
 */
#ifndef DG_ORACLE_H
#define DG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NP 6 /* nodes per P2 triangle, dg.h:3 */
#define ORC_NFP 3 /* nodes per face, N+1 with N=2 (dg.h:2) */

/* Sponge geometry: dg.h:8-9,11-12 (bound_x, bound_z, model_w, model_h), the taper parameter beta
 * passed at DG_leap_Frog_OpenMP.c:287 and the value of Pi at dg.h:6. */
typedef struct {
    double bound_x, bound_z, model_w, model_h, beta, pi;
} orc_sponge_params;

/* ---- file formats (fun_save_to_file.c) ---- */
long orc_parse_double(const char *buf, long len, double *m, long rows, long cols);
long orc_parse_long(const char *buf, long len, long *m, long rows, long cols);

/* ---- setup (fun_DG.c) ---- */
void orc_edge_length(long tri, const double *point, const long *triangle, double *length);
double orc_indiameter_min(long tri, const double *point, const long *triangle);
double orc_time_step(long tri, const double *point, const long *triangle, const double *vp, int order);
void orc_jacobian(long tri, const double *point, const long *triangle, double *jacobi_edge, double *jacobi);
void orc_normals(long tri, const double *point, const long *triangle, double *n1, double *n2, double *n3);
void orc_neighbor_edge(long tri, const long *neigh_raw, const long *triangle, long *neighbor, long *edge,
                       long *edge_flag);
void orc_ibe(long *ibe);
void orc_neighbor_nodes(long tri, const long *triangle, const long *edge, const long *neighbor, const long *ibe,
                        long *node_neighbor);
long orc_absorb_count(long tri, const double *point, const long *triangle, const orc_sponge_params *sp);
long orc_absorbing_fun(long tri, const double *point, const long *triangle, const double *fekete,
                       const orc_sponge_params *sp, double *absorb_fun);
double orc_ricker(long step, double dt, double t0, double f0, double pi);
double orc_ricker_delay(double dt, double f0, double pi);
int orc_get_source(long tri, long point_num, const double *point, const long *triangle, double x, double z,
                   long *src_elem, long *src_node);
int orc_get_receiver(long tri, long point_num, const double *point, const long *triangle, double x, double z,
                     long *rec_elem, long *rec_node);
void orc_source_vector(const double *V, long node, double jacobi, double *fun);
void orc_elastic_ab(long tri, const double *lambda, const double *miu, const double *den, double *A, double *B);
void orc_central_flux(long tri, const double *jacobi, const double *jacobi_edge, const double *A, const double *B,
                      const double *n1, const double *n2, const double *n3, double *Q);
void orc_dx_dz(long tri, const double *point, const long *triangle, const double *Vr, const double *Vs,
               const double *inverV, double *Dx, double *Dz);
void orc_all_node(long tri, const double *point, const long *triangle, const double *fekete, double *all_node);

/* ---- time-step kernels (OpenMP_Function.c) ---- */
void orc_absorb(long tri, double *Vx, double *Vz, double *Txx, double *Tzz, double *Txz, const double *absorb_fun);
void orc_source_kernel(double *Vz, long src_elem, const double *fun, double w);
void orc_lu_velocity(long tri, double dt, double *Vx, double *Vz, const double *Txx, const double *Tzz,
                     const double *Txz, const double *A, const double *B, const double *Dx, const double *Dz,
                     const double *Q, const double *S, const long *ibe, const long *node_neighbor,
                     const long *neighbor, const long *edge_flag);
void orc_lu_stress(long tri, double dt, double *Txx, double *Tzz, double *Txz, const double *Vx, const double *Vz,
                   const double *A, const double *B, const double *Dx, const double *Dz, const double *Q,
                   const double *S, const long *ibe, const long *node_neighbor, const long *neighbor,
                   const long *edge_flag);

/* ---- [EXT] extensions the workloads need (no reference counterpart) ---- */
void orc_p2_basis(double r, double s, double *phi);
long orc_locate(long tri, const double *point, const long *triangle, double x, double z, double *r, double *s);
void orc_source_vector_at(const double *V, const double *phi, double jacobi, double *fun);

/* ---- whole-run composition (DG_leap_Frog_OpenMP.c:397-446) ---- */
typedef struct {
    long elem;      /* 1-based element */
    int kind;       /* 0 = vertical force (adds to Vz), 1 = explosive (adds to Txx and Tzz) */
    double fun[6];  /* per-node source vector */
} orc_source;

typedef struct {
    long elem;         /* 1-based element */
    double weight[6];  /* interpolation weights over the element's 6 nodes */
} orc_receiver;

typedef struct {
    long tri;
    double dt;
    /* setup arrays in the reference's layout, built by the functions above */
    const double *A, *B, *Dx, *Dz, *Q, *S, *absorb_fun;
    const long *ibe, *node_neighbor, *neighbor, *edge_flag;
    long nsrc;
    const orc_source *src;
    const double *wavelet; /* [nsteps] */
    long nrec;
    const orc_receiver *rec;
} orc_problem;

/* Advance fields by nsteps starting at step index `step0`; traces rec_x/rec_z are [nsteps][nrec]. */
void orc_record(const orc_problem *p, const double *Vx, const double *Vz, double *rec_x, double *rec_z);
void orc_run(const orc_problem *p, long step0, long nsteps, double *Vx, double *Vz, double *Txx, double *Tzz,
             double *Txz, double *rec_x, double *rec_z);

/* One step, exposed so tests can time and compose it. */
void orc_step(const orc_problem *p, long step, double *Vx, double *Vz, double *Txx, double *Tzz, double *Txz);

int orc_num_threads(void);
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
