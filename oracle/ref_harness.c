/*
 * ref_harness.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Builds the reference's own solver functions, compiled from their sources where they lie under the
 * reference checkout (the recipe in oracle/Makefile passes -I<reference>), into oracle/_ref/libref.so.
 * Nothing of the reference is copied: the three translation units are #included by path exactly as
 * the reference's driver includes them (DG_leap_Frog_OpenMP.c:4,17-19), with the reference's own dg.h.
 * The reference functions keep their names and are called directly from the tests (ctypes).
 *
 * ref_case() below is this harness's own driver: it calls the reference functions in the order of
 * the reference's main (DG_leap_Frog_OpenMP.c:190-446) on in-memory mesh arrays instead of the
 * hard-coded file paths, with two deviations that avoid the reference's undefined behaviour:
 *   - the premature NeighborNodes call on uninitialised Edge/Neighbor (DG_leap_Frog_OpenMP.c:193)
 *     is skipped; its result is overwritten at DG_leap_Frog_OpenMP.c:324 anyway;
 *   - Edge is zero-initialised (it is never written on boundary faces, fun_DG.c:589-717), and the
 *     element array carries a 3-entry front pad because Neighbor_Edge reads Triangle[-3..-1] on
 *     boundary faces before testing the neighbour (fun_DG.c:584-587);
 *   - the partition-boundary list is sized by Tri (see the comment at its allocation).
 */
#include "dg.h"
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
#include <unistd.h>
#include <sys/types.h>
#include <sys/stat.h>
#include <fcntl.h>
#include <errno.h>
#include <string.h>
#include <sys/time.h>
#include <malloc.h>
#include <omp.h>
#include "fun_save_to_file.c"
#include "fun_DG.c"
#include "OpenMP_Function.c"

/* Compile-time configuration of the reference build, so the tests can mirror it. */
long ref_macro(int which)
{
    switch (which) {
    case 0: return N;
    case 1: return Np;
    case 2: return TIME;
    case 3: return Fp;
    case 4: return Channel;
    case 5: return bound_x;
    case 6: return bound_z;
    case 7: return Source_tri;
    case 8: return model_w;
    case 9: return model_h;
    case 10: return Process_Num;
    }
    return -1;
}
double ref_pi(void) { return Pi; }

typedef struct {
    /* outputs of the setup, all in the reference's layout (caller allocates) */
    double *Dx, *Dz, *Q, *A, *B, *absorb_fun, *jacobi, *jacobi_edge, *n1, *n2, *n3, *fun;
    long *neighbor, *edge_flag, *node_neighbor, *edge;
    double dt, t0;
    long source[2], receiver[2], absorb_num, inside_num;
} ref_setup_out;

/*
 * Run the reference's setup and `nsteps` time steps on an in-memory mesh.  Materials are chosen by
 * attribute exactly like DG_leap_Frog_OpenMP.c:236-265 unless lambda/miu/den/vp are given.
 * snap[k] receives all five fields after step (k+1)*snap_every - 1 (Vx,Vz,Txx,Tzz,Txz, each Tri*Np).
 */
int ref_case(const char *opdir, long Tri, long Point_Num, const double *point_in, const long *triangle_in,
             const long *neigh_in, const long *attribute, const long *epart_in, const double *lambda_in,
             const double *miu_in, const double *den_in, const double *vp_in, double src_x, double src_z,
             double rec_x_pos, double rec_z_pos, long nsteps, long snap_every, ref_setup_out *o, double *rec_x,
             double *rec_z, double *snap)
{
    char path[4096];
    static double FeketeNew[Np][2], S1[Np][N + 1], S2[Np][N + 1], S3[Np][N + 1], V[Np][Np], Vr[Np][Np],
        Vs[Np][Np], inverV[Np][Np];
#define LOADM(name, arr, r, c)                                                                                 \
    snprintf(path, sizeof(path), "%s/%s", opdir, name);                                                      \
    ReadDouble(path, *arr, r, c);
    LOADM("FeketeNew.txt", FeketeNew, Np, 2);
    LOADM("S1.txt", S1, Np, N + 1);
    LOADM("S2.txt", S2, Np, N + 1);
    LOADM("S3.txt", S3, Np, N + 1);
    LOADM("V.txt", V, Np, Np);
    LOADM("Vr.txt", Vr, Np, Np);
    LOADM("Vs.txt", Vs, Np, Np);
    LOADM("inverV.txt", inverV, Np, Np);
#undef LOADM

    long i, j, k, n, num;
    double sum;
    double *Point = (double *)malloc(sizeof(double) * Point_Num * 2);
    memcpy(Point, point_in, sizeof(double) * Point_Num * 2);
    long *tri_pad = (long *)calloc(Tri * 3 + 3, sizeof(long));
    long *Triangle = tri_pad + 3;
    memcpy(Triangle, triangle_in, sizeof(long) * Tri * 3);
    long *neighbor = (long *)malloc(sizeof(long) * Tri * 3);
    memcpy(neighbor, neigh_in, sizeof(long) * Tri * 3);

    double *Vp = (double *)malloc(sizeof(double) * Tri), *Den = (double *)malloc(sizeof(double) * Tri);
    double *Miu = (double *)malloc(sizeof(double) * Tri), *Lambda = (double *)malloc(sizeof(double) * Tri);
    long *Node_Neighbor = (long *)calloc(Tri * 3 * (N + 1), sizeof(long));
    long *Neighbor = (long *)calloc(Tri * 3, sizeof(long));
    long *Edge = (long *)calloc(Tri * 3, sizeof(long));
    long *Edge_flag = (long *)calloc(Tri * 3, sizeof(long));
    double *absorb_fun = (double *)malloc(sizeof(double) * Tri * Np);
    long *Neighbor_Mark = (long *)malloc(sizeof(long) * Tri * 3);
    long Series[Process_Num], Series_order[Process_Num], Bound_Series[Process_Num];
    long *Triangle_part = (long *)malloc(sizeof(long) * Tri);
    double *Jacobi = (double *)malloc(sizeof(double) * Tri), *Jacobi_Edge = (double *)malloc(sizeof(double) * Tri * 3);
    double *Q = (double *)malloc(sizeof(double) * Tri * 30);
    double *A = (double *)malloc(sizeof(double) * Tri * 5), *B = (double *)malloc(sizeof(double) * Tri * 5);
    double *Dx = (double *)malloc(sizeof(double) * Tri * Np * Np), *Dz = (double *)malloc(sizeof(double) * Tri * Np * Np);
    double *S = (double *)malloc(sizeof(double) * 3 * Np * (N + 1));
    long ibe[3 * (N + 1)];
    double *normal1 = (double *)malloc(sizeof(double) * Tri * 2), *normal2 = (double *)malloc(sizeof(double) * Tri * 2),
           *normal3 = (double *)malloc(sizeof(double) * Tri * 2);
    double fun[Source_tri * Np];
    long *epart = (long *)malloc(sizeof(long) * Tri);
    memcpy(epart, epart_in, sizeof(long) * Tri);

    /* DG_leap_Frog_OpenMP.c:202-203 */
    Normals(Tri, Point, Triangle, normal1, normal2, normal3);
    Jacobian(Tri, Jacobi_Edge, Jacobi, Point, Triangle);

    /* DG_leap_Frog_OpenMP.c:207-217 */
    double source_position[2] = {src_x, src_z};
    long source[Source_tri * 3];
    memset(source, 0, sizeof(source));
    Get_Source(Tri, Point_Num, Point, Triangle, source, source_position);
    double receiver_position[2] = {rec_x_pos, rec_z_pos};
    long receiver[Channel * 3];
    memset(receiver, 0, sizeof(receiver));
    Get_Receiver(Tri, Point_Num, Point, Triangle, receiver, receiver_position);
    if (source[0] < 1 || receiver[0] < 1) return -1;

    /* DG_leap_Frog_OpenMP.c:220-233 */
    for (num = 0; num < Source_tri; num++)
        for (i = 0; i < Np; i++) {
            n = source[num * 2 + 0] - 1;
            k = source[num * 2 + 1] - 1;
            sum = 0;
            for (j = 0; j < Np; j++) sum = sum + V[i][j] * V[k][j];
            fun[num * Np + i] = sum / Jacobi[n];
        }

    /* DG_leap_Frog_OpenMP.c:236-265 */
    double Cp1 = 1500, Cs1 = 0, den1 = 1000, miu1 = den1 * Cs1 * Cs1, lambda1 = den1 * Cp1 * Cp1 - 2 * miu1;
    double Cp2 = 3000, Cs2 = 1800, den2 = 2200, miu2 = den2 * Cs2 * Cs2, lambda2 = den2 * Cp2 * Cp2 - 2 * miu2;
    for (i = 0; i < Tri; i++) {
        if (lambda_in) {
            Vp[i] = vp_in[i];
            Lambda[i] = lambda_in[i];
            Miu[i] = miu_in[i];
            Den[i] = den_in[i];
            continue;
        }
        if (attribute[i] == 1 || attribute[i] == 2) {
            Vp[i] = Cp1; Lambda[i] = lambda1; Miu[i] = miu1; Den[i] = den1;
        }
        if (attribute[i] == 3 || attribute[i] == 4) {
            Vp[i] = Cp2; Lambda[i] = lambda2; Miu[i] = miu2; Den[i] = den2;
        }
    }

    /* DG_leap_Frog_OpenMP.c:268-287 */
    double dt = Calculate_Time_Step(Tri, Point, Triangle, Vp);
    double t0 = Ricker_Delay_Time(dt);
    long Absorb_Num = Get_Absorb_Square_Num(Tri, Point, Triangle);
    long Inside_Num = Get_Inside_Square_Num(Tri, Point, Triangle);
    long *absorbing_triangle = (long *)calloc(Absorb_Num + 1, sizeof(long));
    long *inside_triangle = (long *)calloc(Inside_Num + 1, sizeof(long));
    Get_absorb_square(Tri, Point, Triangle, absorbing_triangle, inside_triangle);
    for (i = 0; i < Absorb_Num; i++)
        if (absorbing_triangle[i] < 1) return -2; /* list shorter than its count: reference UB */
    Absorbing_Fun(Tri, Absorb_Num, 0.1, absorb_fun, absorbing_triangle, Point, Triangle, FeketeNew);

    /* DG_leap_Frog_OpenMP.c:290-305 */
    Elastic_A_B(Tri, A, B, Lambda, Miu, Den);
    for (j = 0; j < Np; j++)
        for (n = 0; n < N + 1; n++) {
            S[0 * Np * (N + 1) + j * (N + 1) + n] = S1[j][n];
            S[1 * Np * (N + 1) + j * (N + 1) + n] = S2[j][n];
            S[2 * Np * (N + 1) + j * (N + 1) + n] = S3[j][n];
        }
    Neighbor_Edge(Tri, Neighbor, neighbor, Edge, Triangle, Edge_flag);
    IBE(ibe);

    /* DG_leap_Frog_OpenMP.c:324-343 */
    NeighborNodes(Tri, Node_Neighbor, Triangle, Edge, Neighbor, ibe);
    Dx_Dz(Tri, Dx, Dz, Vr, Vs, inverV, Triangle, Point);
    Neighbor_Partition_Mark(Tri, Neighbor, epart, Neighbor_Mark);
    long Boundary_Num = Calculate_Boundary_Num(Tri, Neighbor_Mark);
    /* Calculate_Boundary_Num (fun_DG.c:761-794) does not count elements whose three neighbours all lie
     * in another partition, but Partition_Boundary_Neighbor (fun_DG.c:815-824) lists them: size the list
     * by Tri so such partitions do not overrun it. */
    long *Boundary_Triangle = (long *)calloc(Tri + 1, sizeof(long));
    (void)Boundary_Num;
    Partition_Boundary_Neighbor(Tri, Neighbor, epart, Neighbor_Mark, Series, Series_order, Bound_Series,
                                Triangle_part, Boundary_Triangle);
    Central_flux(Tri, Jacobi, Jacobi_Edge, Q, A, B, normal1, normal2, normal3);

    if (o) {
#define OUT(dst, src, n) if (o->dst) memcpy(o->dst, src, sizeof(*src) * (n));
        OUT(Dx, Dx, Tri * 36); OUT(Dz, Dz, Tri * 36); OUT(Q, Q, Tri * 30); OUT(A, A, Tri * 5); OUT(B, B, Tri * 5);
        OUT(absorb_fun, absorb_fun, Tri * Np); OUT(jacobi, Jacobi, Tri); OUT(jacobi_edge, Jacobi_Edge, Tri * 3);
        OUT(n1, normal1, Tri * 2); OUT(n2, normal2, Tri * 2); OUT(n3, normal3, Tri * 2); OUT(fun, fun, Np);
        OUT(neighbor, Neighbor, Tri * 3); OUT(edge_flag, Edge_flag, Tri * 3);
        OUT(node_neighbor, Node_Neighbor, Tri * 9); OUT(edge, Edge, Tri * 3);
#undef OUT
        o->dt = dt;
        o->t0 = t0;
        o->source[0] = source[0]; o->source[1] = source[1];
        o->receiver[0] = receiver[0]; o->receiver[1] = receiver[1];
        o->absorb_num = Absorb_Num;
        o->inside_num = Inside_Num;
    }

    /* DG_leap_Frog_OpenMP.c:382-446, without the timer and the text snapshots */
    double *Vx1 = (double *)calloc(Tri * Np, sizeof(double)), *Vz1 = (double *)calloc(Tri * Np, sizeof(double));
    double *Txx1 = (double *)calloc(Tri * Np, sizeof(double)), *Tzz1 = (double *)calloc(Tri * Np, sizeof(double));
    double *Txz1 = (double *)calloc(Tri * Np, sizeof(double));
    double *record_x = (double *)calloc((nsteps > TIME ? nsteps : TIME) * Channel, sizeof(double));
    double *record_z = (double *)calloc((nsteps > TIME ? nsteps : TIME) * Channel, sizeof(double));
    long nsnap = 0;
    int m = 0;
    for (long time = 0; time < nsteps; time++) {
        double w = Ricker_Source(time, dt, t0);
        Absorb_Boundary(Tri, Vx1, Vz1, Txx1, Tzz1, Txz1, absorb_fun, Triangle_part, Series, Series_order);
        source_kernel(Vz1, source, fun, w, m);
        LU_Velocity(Tri, dt, Vx1, Vz1, Txx1, Tzz1, Txz1, Triangle_part, Neighbor_Mark, Series, Series_order, A, B,
                    Dx, Dz, Q, S, ibe, Node_Neighbor, Neighbor, Edge_flag);
        LU_Stress_Elastic(Tri, dt, Txx1, Tzz1, Txz1, Vx1, Vz1, Triangle_part, Neighbor_Mark, Series, Series_order,
                          A, B, Dx, Dz, Q, S, ibe, Node_Neighbor, Neighbor, Edge_flag);
        Record(Vx1, Vz1, time, record_x, record_z, receiver);
        if (snap && snap_every > 0 && (time + 1) % snap_every == 0) {
            double *dst = snap + nsnap * 5 * Tri * Np;
            memcpy(dst + 0 * Tri * Np, Vx1, sizeof(double) * Tri * Np);
            memcpy(dst + 1 * Tri * Np, Vz1, sizeof(double) * Tri * Np);
            memcpy(dst + 2 * Tri * Np, Txx1, sizeof(double) * Tri * Np);
            memcpy(dst + 3 * Tri * Np, Tzz1, sizeof(double) * Tri * Np);
            memcpy(dst + 4 * Tri * Np, Txz1, sizeof(double) * Tri * Np);
            nsnap++;
        }
    }
    if (rec_x) memcpy(rec_x, record_x, sizeof(double) * nsteps);
    if (rec_z) memcpy(rec_z, record_z, sizeof(double) * nsteps);

    free(Point); free(tri_pad); free(neighbor); free(Vp); free(Den); free(Miu); free(Lambda);
    free(Node_Neighbor); free(Neighbor); free(Edge); free(Edge_flag); free(absorb_fun); free(Neighbor_Mark);
    free(Triangle_part); free(Jacobi); free(Jacobi_Edge); free(Q); free(A); free(B); free(Dx); free(Dz); free(S);
    free(normal1); free(normal2); free(normal3); free(epart); free(absorbing_triangle); free(inside_triangle);
    free(Boundary_Triangle); free(Vx1); free(Vz1); free(Txx1); free(Tzz1); free(Txz1); free(record_x);
    free(record_z);
    return 0;
}

/*
 * Time only the reference's step loop (absorb, source, velocity, stress, record) for `nsteps` steps
 * on the given setup arrays; used as the CPU baseline.  Returns wall seconds.
 */
double ref_time_steps(long Tri, double dt, const double *A, const double *B, const double *Dx, const double *Dz,
                      const double *Q, const double *S, long *ibe, long *Node_Neighbor, long *Neighbor,
                      long *Edge_flag, double *absorb_fun, long src_elem, const double *fun, long rec_elem,
                      long rec_node, long nsteps, double t0)
{
    long Series[Process_Num], Series_order[Process_Num];
    long *Triangle_part = (long *)malloc(sizeof(long) * Tri);
    long *Neighbor_Mark = (long *)calloc(Tri * 3, sizeof(long));
    long per = (Tri + Process_Num - 1) / Process_Num, start = 0;
    for (int m = 0; m < Process_Num; m++) {
        long cnt = (start + per <= Tri) ? per : Tri - start;
        Series[m] = cnt;
        Series_order[m] = start;
        start += cnt;
    }
    for (long e = 0; e < Tri; e++) Triangle_part[e] = e + 1;
    double *Vx1 = (double *)calloc(Tri * Np, sizeof(double)), *Vz1 = (double *)calloc(Tri * Np, sizeof(double));
    double *Txx1 = (double *)calloc(Tri * Np, sizeof(double)), *Tzz1 = (double *)calloc(Tri * Np, sizeof(double));
    double *Txz1 = (double *)calloc(Tri * Np, sizeof(double));
    double *record_x = (double *)calloc(nsteps > TIME ? nsteps : TIME, sizeof(double));
    double *record_z = (double *)calloc(nsteps > TIME ? nsteps : TIME, sizeof(double));
    long source[3] = {src_elem, 0, 0}, receiver[3] = {rec_elem, rec_node, 0};
    double funv[Np];
    memcpy(funv, fun, sizeof(funv));
    struct timeval a, b;
    gettimeofday(&a, NULL);
    for (long time = 0; time < nsteps; time++) {
        double w = Ricker_Source(time, dt, t0);
        Absorb_Boundary(Tri, Vx1, Vz1, Txx1, Tzz1, Txz1, absorb_fun, Triangle_part, Series, Series_order);
        source_kernel(Vz1, source, funv, w, 0);
        LU_Velocity(Tri, dt, Vx1, Vz1, Txx1, Tzz1, Txz1, Triangle_part, Neighbor_Mark, Series, Series_order,
                    (double *)A, (double *)B, (double *)Dx, (double *)Dz, (double *)Q, (double *)S, ibe,
                    Node_Neighbor, Neighbor, Edge_flag);
        LU_Stress_Elastic(Tri, dt, Txx1, Tzz1, Txz1, Vx1, Vz1, Triangle_part, Neighbor_Mark, Series, Series_order,
                          (double *)A, (double *)B, (double *)Dx, (double *)Dz, (double *)Q, (double *)S, ibe,
                          Node_Neighbor, Neighbor, Edge_flag);
        Record(Vx1, Vz1, time, record_x, record_z, receiver);
    }
    gettimeofday(&b, NULL);
    double sec = (b.tv_sec - a.tv_sec) + 1e-6 * (b.tv_usec - a.tv_usec);
    free(Triangle_part); free(Neighbor_Mark); free(Vx1); free(Vz1); free(Txx1); free(Tzz1); free(Txz1);
    free(record_x); free(record_z);
    return sec;
}
