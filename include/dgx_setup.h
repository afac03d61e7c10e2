/*
 * dgx_setup.h -- host-side setup and I/O of libdgx.so: the reference's input formats and the
 * once-per-run quantities its setup computes on the host (transcendental ones included), so that a C
 * driver can prepare a dgx_problem exactly as the reference's main does.
 * All functions are plain C, allocate nothing they return, and report errors as negative DGX_E*.
 */
#ifndef DGX_SETUP_H
#define DGX_SETUP_H

#include "dgx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* MATLAB "save -ascii" reader with the reference's digit arithmetic (ReadDouble,
 * fun_save_to_file.c:22-113; ReadInt, fun_save_to_file.c:116-204), streaming, without the
 * reference's 1e8-byte limit.  Returns the number of values read (rows*cols) or a DGX_E code. */
long dgx_read_ascii(const char *path, double *m, long rows, long cols);
long dgx_read_ascii_long(const char *path, long *m, long rows, long cols);
/* Reads the 8 operator files of DG_leap_Frog_OpenMP.c:52-81 from a directory. */
int dgx_read_operators(const char *dir, dgx_operators *ops);
/* Header row counts of .ele / .node (Calculate_Tri, Calculate_Point_Num, fun_DG.c:70-87). */
long dgx_mesh_count(const char *path);
/* Mesh_Input (fun_DG.c:90-119): Point [np*2], Triangle [tri*3], neighbor [tri*3], attribute [tri]. */
int dgx_read_mesh(const char *ele, const char *node, const char *neigh, long tri, long point_num, double *point,
                  long *triangle, long *neighbor, long *attribute);

/* Binary mesh (no counterpart in the reference, whose text reader cannot hold files over 1e8 bytes,
 * fun_save_to_file.c:36,130): little-endian, "DGXMESH1", int64 tri, int64 point_num, then float64
 * point [np*2], int64 triangle [tri*3], int64 neighbor [tri*3] (the .neigh columns, -1 = none), int64
 * attribute [tri].  Read and written in chunks, so a mesh never needs a second in-memory copy. */
int dgx_write_mesh_binary(const char *path, long tri, long point_num, const double *point, const long *triangle,
                          const long *neighbor, const long *attribute);
int dgx_mesh_binary_counts(const char *path, long *tri, long *point_num);
int dgx_read_mesh_binary(const char *path, long tri, long point_num, double *point, long *triangle, long *neighbor,
                         long *attribute);

/* Writers: Save_Data "%0.14f " rows (fun_save_to_file.c:233-255), Save_Record (fun_save_to_file.c:260-281),
 * and a lossless little-endian float64 twin of either. */
int dgx_save_data(const char *path, const double *m, long rows, long cols);
int dgx_save_record(const char *path, const double *m, long rows);
int dgx_save_binary(const char *path, const double *m, long n);
/* Physical coordinates of all element nodes, [tri*6*2] (set_All_node, fun_DG.c:839-861). */
void dgx_all_node(long tri, const double *point, const long *triangle, const double *fekete, double *all_node);

/* Sponge: interior box [bound_x, bound_x+model_w] x [bound_z, bound_z+model_h], taper
 * 0.5+0.5*cos(beta*d*pi/L) (Absorbing_Fun, fun_DG.c:906-1006; box and widths dg.h:8-12). */
typedef struct {
    double bound_x, bound_z, model_w, model_h, beta, pi;
} dgx_sponge_params;
/* Fills absorb_fun [tri*6].  Returns the number of elements with a factor != 1, or DGX_EMESH when an
 * element has a vertex outside the box but not its first vertex (the reference would read past its
 * absorbing-element list for such a mesh, fun_DG.c:202-281). */
long dgx_sponge(long tri, const double *point, const long *triangle, const double *fekete,
                const dgx_sponge_params *sp, double *absorb_fun);

/* CFL step dt = 0.99 * min in-circle diameter / max(vp) / (2N+1), N = 2 (Calculate_Time_Step,
 * fun_DG.c:179-199). */
double dgx_cfl_dt(long tri, const double *point, const long *triangle, const double *vp);
/* Ricker wavelet at t = dt*step (Ricker_Source, fun_DG.c:1279-1286) and its delay
 * (Ricker_Delay_Time, fun_DG.c:1289-1309); pi is a parameter because the reference uses 3.1415. */
double dgx_ricker(long step, double dt, double t0, double f0, double pi);
double dgx_ricker_delay(double dt, double f0, double pi);
int dgx_wavelet(long nsteps, double dt, double t0, double f0, double pi, double *w);

/* Point location accelerator over a mesh (uniform grid of element bounding boxes). */
typedef struct dgx_locator dgx_locator;
int dgx_locator_create(long tri, long point_num, const double *point, const long *triangle, dgx_locator **out);
void dgx_locator_destroy(dgx_locator *loc);
/* Source at (x,z): on a mesh vertex (|dx|,|dz| < 1e-3) it is the reference's Get_Source
 * (fun_DG.c:284-353: first such vertex, first element holding it); elsewhere the first element
 * (in element order) containing the point.  Fills elem and fun = (V V^T phi)/J. */
int dgx_locate_source(const dgx_locator *loc, const dgx_operators *ops, double x, double z, int kind, int shot,
                      long wavelet, dgx_source *out);
/* Receivers: on a vertex (< 1e-2) the reference's Get_Receiver (fun_DG.c:356-418: last such vertex,
 * last element holding it, unit weight); elsewhere the first containing element and the P2 Lagrange
 * weights of the point. */
int dgx_locate_receivers(const dgx_locator *loc, long n, const double *xz, dgx_receiver *out);
/* The same two searches on the current HIP device (a vertex match over all vertices, the element holding it, and
 * the smallest element containing the point among those the grid files under the point's cell), for meshes
 * and receiver sets too large for the host search to be cheap.  The results equal dgx_locate_source and
 * dgx_locate_receivers on a locator of the same mesh bit for bit.  Sources: kind, shot and wavelet per point. */
int dgx_locate_receivers_device(long tri, long point_num, const double *point, const long *triangle, long n,
                                const double *xz, dgx_receiver *out);
int dgx_locate_sources_device(long tri, long point_num, const double *point, const long *triangle,
                              const dgx_operators *ops, long n, const double *xz, const int *kind, const int *shot,
                              const long *wavelet, dgx_source *out);
/* The reference's per-element operators, as its setup forms them on the host: Dx, Dz [tri*36] (Dx_Dz,
 * fun_DG.c:1137-1192), the face flux coefficients Q [tri*30] (Central_flux, fun_DG.c:1091-1136, from the
 * normals and edge Jacobians of fun_DG.c:421-554) and the material factors A, B [tri*5] (Elastic_A_B,
 * fun_DG.c:1071-1090), with the reference's expressions in its order, so that they equal its arrays bit for
 * bit.  DGX_FP64_REFERENCE steps with Q, A and B from here and forms Dx and Dz in its kernels with the same
 * operations; the caller needs these arrays only to inspect them. */
int dgx_reference_operators(long tri, const double *point, const long *triangle, const double *lambda,
                            const double *miu, const double *den, const dgx_operators *ops, double *Dx, double *Dz,
                            double *Q, double *A, double *B);
/* P2 Lagrange basis at reference coordinates (r,s) in the node order of FeketeNew.txt. */
void dgx_p2_basis(double r, double s, double *phi);

#ifdef __cplusplus
}
#endif
#endif
