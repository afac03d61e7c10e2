/*
 * dg_oracle.c -- TEST INFRASTRUCTURE ONLY (see dg_oracle.h).
 *
 * Clean-room CPU restatement of the reference's 2-D nodal P2 velocity-stress DG solver.  It keeps the
 * reference's data layout and, where the reference fixes an evaluation order, the same order, so that
 * this file and the reference functions compiled with the same flags agree bit for bit.  The element
 * loops are OpenMP `parallel for`, like OpenMP_Function.c; every element writes only its own 6 nodes,
 * so results do not depend on the thread count (OpenMP_Function.c:118-121,180-181,298-302).
 */
#include "dg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NP ORC_NP
#define NFP ORC_NFP

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------------------------------
 * Text formats.  MATLAB `save -ascii` fields "   d.ddddddde+XX".  The reference does not call strtod:
 * it locates each 'e', takes the sign from 10 bytes earlier, sums the 8 mantissa digits with
 * pow(10,k) weights (last digit first) and scales by pow(10, exponent)
 * (fun_save_to_file.c:55-111).  Integers use the same scan with the digits weighted by the exponent
 * (fun_save_to_file.c:149-202).  Rows advance at '\n'.  Returns the number of values stored.
 * ---------------------------------------------------------------------------------------------- */
long orc_parse_double(const char *s, long len, double *m, long rows, long cols)
{
    long want = rows * cols, got = 0, row = 0, col = 0;
    long sign = 1, esign = 1;
    for (long i = 0; i < len && got < want; i++) {
        if (s[i] == 'e' && i >= 10 && i + 3 < len) {
            if (s[i - 10] == ' ') sign = 1;
            if (s[i - 10] == '-') sign = -1;
            if (s[i + 1] == '+') esign = 1;
            if (s[i + 1] == '-') esign = -1;
            double v = (double)(s[i - 9] - '0');
            long k = -7;
            for (long j = i - 1; j >= i - 7; j--) { /* fun_save_to_file.c:86-90 */
                v = v + (s[j] - '0') * pow(10, k);
                k++;
            }
            long p = esign * ((s[i + 2] - '0') * 10 + (s[i + 3] - '0'));
            v = v * pow(10, p) * sign; /* fun_save_to_file.c:93 */
            m[row * cols + col] = v;
            col++;
            got++;
        }
        if (got == want) break;
        if (s[i] == '\n') {
            row++;
            col = 0;
        }
    }
    return got;
}

long orc_parse_long(const char *s, long len, long *m, long rows, long cols)
{
    long want = rows * cols, got = 0, row = 0, col = 0;
    long sign = 1, esign = 1;
    for (long i = 0; i < len && got < want; i++) {
        if (s[i] == 'e' && i >= 10 && i + 3 < len) {
            if (s[i - 10] == ' ') sign = 1;
            if (s[i - 10] == '-') sign = -1;
            if (s[i + 1] == '+') esign = 1;
            if (s[i + 1] == '-') esign = -1;
            long p = esign * ((s[i + 2] - '0') * 10 + (s[i + 3] - '0'));
            long *dst = &m[row * cols + col];
            *dst = (long)((s[i - 9] - '0') * pow(10, p) * sign); /* fun_save_to_file.c:178 */
            long k = p - 1;
            for (long j = i - 7; j <= i - 7 + p - 1; j++) { /* fun_save_to_file.c:181-185 */
                *dst = (long)(*dst + (s[j] - '0') * pow(10, k));
                k--;
            }
            col++;
            got++;
        }
        if (got == want) break;
        if (s[i] == '\n') {
            row++;
            col = 0;
        }
    }
    return got;
}

/* ------------------------------------------------------------------------------------------------
 * Geometry (fun_DG.c:122-199, 421-554).
 * ---------------------------------------------------------------------------------------------- */
static inline void tri_vertices(const double *point, const long *triangle, long e, double P[3][2])
{
    for (int v = 0; v < 3; v++) {
        long p = triangle[e * 3 + v] - 1;
        P[v][0] = point[p * 2 + 0];
        P[v][1] = point[p * 2 + 1];
    }
}

/* L0=|v0v1|, L1=|v1v2|, L2=|v0v2| (fun_DG.c:140-142) */
void orc_edge_length(long tri, const double *point, const long *triangle, double *length)
{
    for (long e = 0; e < tri; e++) {
        double P[3][2];
        tri_vertices(point, triangle, e, P);
        length[e * 3 + 0] = sqrt(pow(P[0][0] - P[1][0], 2) + pow(P[0][1] - P[1][1], 2));
        length[e * 3 + 1] = sqrt(pow(P[1][0] - P[2][0], 2) + pow(P[1][1] - P[2][1], 2));
        length[e * 3 + 2] = sqrt(pow(P[0][0] - P[2][0], 2) + pow(P[0][1] - P[2][1], 2));
    }
}

/* Smallest in-circle diameter 2*Area/s over all elements, Heron's area (fun_DG.c:147-176). */
double orc_indiameter_min(long tri, const double *point, const long *triangle)
{
    double *len = (double *)malloc(sizeof(double) * tri * 3);
    orc_edge_length(tri, point, triangle, len);
    double best = 0;
    for (long e = 0; e < tri; e++) {
        double a = len[e * 3 + 0], b = len[e * 3 + 1], c = len[e * 3 + 2];
        double s = 0.5 * (a + b + c);
        double S = sqrt(s * (s - a) * (s - b) * (s - c));
        double d = 2 * S / s;
        if (e == 0 || d < best) best = d;
    }
    free(len);
    return best;
}

/* dt = 0.99 * min diameter / max Vp / (2N+1) (fun_DG.c:179-199) */
double orc_time_step(long tri, const double *point, const long *triangle, const double *vp, int order)
{
    double vmax = vp[0];
    for (long e = 1; e < tri; e++)
        if (vp[e] > vmax) vmax = vp[e];
    double dmin = orc_indiameter_min(tri, point, triangle);
    return dmin / vmax / (2 * order + 1) * 0.99;
}

/* J = xr*ys - xs*yr with the half-edge vectors; Jacobi_Edge = L/2 (fun_DG.c:436-452) */
void orc_jacobian(long tri, const double *point, const long *triangle, double *jacobi_edge, double *jacobi)
{
    double *len = (double *)malloc(sizeof(double) * tri * 3);
    orc_edge_length(tri, point, triangle, len);
    for (long e = 0; e < tri; e++) {
        double P[3][2];
        tri_vertices(point, triangle, e, P);
        double xr = (P[1][0] - P[0][0]) * 0.5, yr = (P[1][1] - P[0][1]) * 0.5;
        double xs = (P[2][0] - P[0][0]) * 0.5, ys = (P[2][1] - P[0][1]) * 0.5;
        jacobi[e] = xr * ys - xs * yr;
        for (int j = 0; j < 3; j++) jacobi_edge[e * 3 + j] = len[e * 3 + j] / 2;
    }
    free(len);
}

/*
 * Outward unit normals through the in-circle tangent points (fun_DG.c:456-554): the in-centre A is
 * the side-length-weighted vertex mean; the tangent point on the edge that starts at vertex Pk lies
 * |A-Pk|*cos(angle_k/2) along that edge; the normal is (B-A)/|B-A|.  Edge order: P1P2, P2P3, P3P1.
 */
void orc_normals(long tri, const double *point, const long *triangle, double *n1, double *n2, double *n3)
{
    double *len = (double *)malloc(sizeof(double) * tri * 3);
    orc_edge_length(tri, point, triangle, len);
    const double ux[2] = {1, 0}, uy[2] = {0, 1};
    for (long e = 0; e < tri; e++) {
        double P[3][2];
        tri_vertices(point, triangle, e, P);
        double L1 = len[e * 3 + 0], L2 = len[e * 3 + 1], L3 = len[e * 3 + 2];
        double ang[3];
        ang[0] = acos((L1 * L1 + L3 * L3 - L2 * L2) / (2 * L1 * L3)); /* fun_DG.c:470-472 */
        ang[1] = acos((L1 * L1 + L2 * L2 - L3 * L3) / (2 * L1 * L2));
        ang[2] = acos((L2 * L2 + L3 * L3 - L1 * L1) / (2 * L2 * L3));
        double A[2];
        for (int j = 0; j < 2; j++)
            A[j] = (len[e * 3 + 1] * P[0][j] + len[e * 3 + 2] * P[1][j] + len[e * 3 + 0] * P[2][j]) /
                   (len[e * 3 + 1] + len[e * 3 + 2] + len[e * 3 + 0]);
        double *out[3] = {n1, n2, n3};
        for (int k = 0; k < 3; k++) {
            const double *Ps = P[k], *Pe = P[(k + 1) % 3];
            double Lk = len[e * 3 + k];
            double d = sqrt(pow(A[0] - Ps[0], 2) + pow(A[1] - Ps[1], 2)) * cos(0.5 * ang[k]);
            double AB[2];
            for (int j = 0; j < 2; j++) {
                double B = Ps[j] + (d / Lk) * (Pe[j] - Ps[j]);
                AB[j] = B - A[j];
            }
            double nrm = sqrt(pow(AB[0], 2) + pow(AB[1], 2));
            out[k][e * 2 + 0] = (AB[0] * ux[0] + AB[1] * ux[1]) / nrm;
            out[k][e * 2 + 1] = (AB[0] * uy[0] + AB[1] * uy[1]) / nrm;
        }
    }
    free(len);
}

/* ------------------------------------------------------------------------------------------------
 * Topology (fun_DG.c:557-739, 1009-1068, 1195-1276).
 * Face j of element e runs v_j -> v_{j+1} for j = 0,1 and v0 -> v2 for j = 2.  Triangle's .neigh
 * column c is the element opposite vertex c, so face 0 <- column 2, face 1 <- column 0,
 * face 2 <- column 1 (fun_DG.c:576-578).  Edge[e][j] is the 1-based face of the neighbour holding
 * the same two vertices; it stays 0 on the outer boundary, where Neighbor becomes 1 and
 * Edge_flag 0 (fun_DG.c:720-738).
 * ---------------------------------------------------------------------------------------------- */
static int face_of(const long *v, long a, long b)
{
    /* 1-based face index holding vertex ids {a,b}: {v0,v1}->1, {v1,v2}->2, {v0,v2}->3 */
    int has0 = (v[0] == a || v[0] == b), has1 = (v[1] == a || v[1] == b), has2 = (v[2] == a || v[2] == b);
    if (a == b) return 0;
    if (has0 && has1) return 1;
    if (has1 && has2) return 2;
    if (has0 && has2) return 3;
    return 0;
}

void orc_neighbor_edge(long tri, const long *neigh_raw, const long *triangle, long *neighbor, long *edge,
                       long *edge_flag)
{
    static const int col_of_face[3] = {2, 0, 1};
    for (long e = 0; e < tri; e++) {
        const long *v = &triangle[e * 3];
        long fa[3] = {v[0], v[1], v[2]}, fb[3] = {v[1], v[2], v[0]};
        for (int j = 0; j < 3; j++) {
            long nb = neigh_raw[e * 3 + col_of_face[j]];
            if (nb == -1) nb = 0;
            neighbor[e * 3 + j] = nb;
            edge[e * 3 + j] = 0;
            if (nb) edge[e * 3 + j] = face_of(&triangle[(nb - 1) * 3], fa[j], fb[j]);
        }
    }
    for (long e = 0; e < tri * 3; e++) {
        if (neighbor[e] == 0) neighbor[e] = 1;
        edge_flag[e] = edge[e] ? 1 : 0;
    }
}

/* Local face nodes in face order, P2 (fun_DG.c:1046-1057). */
void orc_ibe(long *ibe)
{
    static const long tab[9] = {6, 4, 1, 1, 2, 3, 6, 5, 3};
    memcpy(ibe, tab, sizeof(tab));
}

/*
 * Node_Neighbor[e][j][i]: the neighbour's local node facing own face node ibe[j][i].  The
 * neighbour's face list is taken in its own ibe order, reversed when its face starts at a different
 * vertex than the own face (fun_DG.c:1207-1265).  Boundary faces get node 1 (fun_DG.c:1267-1272).
 */
void orc_neighbor_nodes(long tri, const long *triangle, const long *edge, const long *neighbor, const long *ibe,
                        long *node_neighbor)
{
    for (long e = 0; e < tri; e++) {
        const long *v = &triangle[e * 3];
        for (int j = 0; j < 3; j++) {
            long *out = &node_neighbor[e * 9 + j * 3];
            long m = edge[e * 3 + j] - 1;
            if (m == -1) {
                out[0] = out[1] = out[2] = 1;
                continue;
            }
            long nb = neighbor[e * 3 + j] - 1;
            long nb_start = (m == 2) ? triangle[nb * 3 + 0] : triangle[nb * 3 + m];
            long own_start = (j == 2) ? v[0] : v[j];
            for (int i = 0; i < 3; i++) out[i] = (own_start == nb_start) ? ibe[m * 3 + i] : ibe[m * 3 + 2 - i];
        }
    }
}

/* ------------------------------------------------------------------------------------------------
 * Sponge (fun_DG.c:202-281, 864-1006).  The reference counts "absorbing" elements as those with any
 * vertex strictly outside the interior box but lists only those whose FIRST vertex is outside; the
 * list tail is then read uninitialised.  orc_absorb_count returns the two counts' difference (0 means
 * the mesh is safe for the reference).  Node factors: 0.5+0.5*cos(beta*d*pi/L) in the four strips
 * (strict inequalities) and in the four corners within radial distance < bound_x of the inner
 * corner; the corner rules are applied after the strip rules; everything else stays 1.
 * ---------------------------------------------------------------------------------------------- */
static int vertex_outside(const double *p, const orc_sponge_params *sp)
{
    return p[0] > (sp->bound_x + sp->model_w) || p[0] < sp->bound_x || p[1] > (sp->bound_z + sp->model_h) ||
           p[1] < sp->bound_z;
}

long orc_absorb_count(long tri, const double *point, const long *triangle, const orc_sponge_params *sp)
{
    long any = 0, first = 0;
    for (long e = 0; e < tri; e++) {
        for (int v = 0; v < 3; v++) {
            if (vertex_outside(&point[(triangle[e * 3 + v] - 1) * 2], sp)) {
                any++;
                break;
            }
        }
        if (vertex_outside(&point[(triangle[e * 3 + 0] - 1) * 2], sp)) first++;
    }
    return any - first;
}

static void rs_to_xy(const double P[3][2], const double *fekete, double pos[NP][2])
{
    /* fun_DG.c:869-875 */
    for (int i = 0; i < NP; i++) {
        double r = fekete[i * 2 + 0], s = fekete[i * 2 + 1];
        for (int c = 0; c < 2; c++) pos[i][c] = -(r + s) / 2 * P[0][c] + (r + 1) / 2 * P[1][c] + (s + 1) / 2 * P[2][c];
    }
}

long orc_absorbing_fun(long tri, const double *point, const long *triangle, const double *fekete,
                       const orc_sponge_params *sp, double *absorb_fun)
{
    const double bx = sp->bound_x, bz = sp->bound_z, W = sp->model_w, H = sp->model_h;
    const double Lx = bx, Ly = bz, beta = sp->beta, Pi = sp->pi;
    long listed = 0;
    for (long k = 0; k < tri * NP; k++) absorb_fun[k] = 1;
    for (long e = 0; e < tri; e++) {
        if (!vertex_outside(&point[(triangle[e * 3 + 0] - 1) * 2], sp)) continue; /* fun_DG.c:266 */
        listed++;
        double P[3][2], pos[NP][2];
        tri_vertices(point, triangle, e, P);
        rs_to_xy(P, fekete, pos);
        for (int i = 0; i < NP; i++) {
            double x = pos[i][0], y = pos[i][1], d, dx, dy;
            double *f = &absorb_fun[e * NP + i];
            if (x < bx && (y > bz && y < bz + H)) { d = bx - x; *f = 0.5 + 0.5 * cos(beta * d * Pi / Lx); }
            if (x > bx + W && (y > bz && y < bz + H)) { d = x - bx - W; *f = 0.5 + 0.5 * cos(beta * d * Pi / Lx); }
            if (y < bz && (x > bx && x < bx + W)) { d = bz - y; *f = 0.5 + 0.5 * cos(beta * d * Pi / Ly); }
            if (y > bz + H && (x > bx && x < bx + W)) { d = y - bz - H; *f = 0.5 + 0.5 * cos(beta * d * Pi / Ly); }
            dx = fabs(x - bx); dy = fabs(y - bz); d = sqrt(dx * dx + dy * dy);
            if (x < bx && y < bz && d < bx) *f = 0.5 + 0.5 * cos(beta * d * Pi / Lx);
            dx = fabs(x - (bx + W)); dy = fabs(y - bz); d = sqrt(dx * dx + dy * dy);
            if (x > bx + W && y < bz && d < bx) *f = 0.5 + 0.5 * cos(beta * d * Pi / Lx);
            dx = fabs(x - (bx + W)); dy = fabs(y - (bz + H)); d = sqrt(dx * dx + dy * dy);
            if (x > bx + W && y > bz + H && d < bx) *f = 0.5 + 0.5 * cos(beta * d * Pi / Lx);
            dx = fabs(x - bx); dy = fabs(y - (bz + H)); d = sqrt(dx * dx + dy * dy);
            if (x < bx && y > bz + H && d < bx) *f = 0.5 + 0.5 * cos(beta * d * Pi / Lx);
        }
    }
    return listed;
}

/* ------------------------------------------------------------------------------------------------
 * Ricker wavelet and its delay (fun_DG.c:1279-1309).  t0 = 2.6 * the first sample time t2 at which
 * the undelayed wavelet is rising and below 1e-8 (i.e. just past its negative side lobe).
 * ---------------------------------------------------------------------------------------------- */
double orc_ricker(long step, double dt, double t0, double f0, double pi)
{
    double t = dt * step;
    return (1 - 2 * pi * pi * f0 * f0 * (t - t0) * (t - t0)) * exp(-pi * pi * f0 * f0 * (t - t0) * (t - t0));
}

double orc_ricker_delay(double dt, double f0, double pi)
{
    for (long T = 0;; T++) {
        double t1 = T * dt, t2 = (T + 1) * dt;
        double s1 = (1 - 2 * pi * pi * f0 * f0 * (t1) * (t1)) * exp(-pi * pi * f0 * f0 * (t1) * (t1));
        double s2 = (1 - 2 * pi * pi * f0 * f0 * (t2) * (t2)) * exp(-pi * pi * f0 * f0 * (t2) * (t2));
        if (s2 > s1 && s2 < 1e-8) return t2 * 2.6;
    }
}

/* ------------------------------------------------------------------------------------------------
 * Source and receiver on mesh vertices (fun_DG.c:284-418).  The source takes the first vertex within
 * 1e-3 and the first element holding it; the receiver takes the last vertex within 1e-2 and the
 * last element holding it.  Local node of a vertex: v0->6, v1->1, v2->3.
 * ---------------------------------------------------------------------------------------------- */
static const long vertex_node[3] = {6, 1, 3};

int orc_get_source(long tri, long point_num, const double *point, const long *triangle, double x, double z,
                   long *src_elem, long *src_node)
{
    long node = 0;
    for (long i = 0; i < point_num; i++) {
        if (fabs(point[i * 2] - x) < 1e-3 && fabs(point[i * 2 + 1] - z) < 1e-3) {
            node = i + 1;
            break;
        }
    }
    if (!node) return -1;
    for (long e = 0; e < tri; e++)
        for (int v = 0; v < 3; v++)
            if (triangle[e * 3 + v] == node) {
                *src_elem = e + 1;
                *src_node = vertex_node[v];
                return 0;
            }
    return -1;
}

int orc_get_receiver(long tri, long point_num, const double *point, const long *triangle, double x, double z,
                     long *rec_elem, long *rec_node)
{
    long node = 0;
    for (long i = 0; i < point_num; i++)
        if (fabs(point[i * 2] - x) < 1e-2 && fabs(point[i * 2 + 1] - z) < 1e-2) node = i + 1;
    if (!node) return -1;
    int found = 0;
    for (long e = 0; e < tri; e++)
        for (int v = 0; v < 3; v++)
            if (triangle[e * 3 + v] == node) {
                *rec_elem = e + 1;
                *rec_node = vertex_node[v];
                found = 1;
            }
    return found ? 0 : -1;
}

/* fun_i = sum_j V[i][j] V[k][j] / J  =  (M^-1 e_k)_i / J  (DG_leap_Frog_OpenMP.c:220-233) */
void orc_source_vector(const double *V, long node, double jacobi, double *fun)
{
    long k = node - 1;
    for (int i = 0; i < NP; i++) {
        double sum = 0;
        for (int j = 0; j < NP; j++) sum = sum + V[i * NP + j] * V[k * NP + j];
        fun[i] = sum / jacobi;
    }
}

/* ------------------------------------------------------------------------------------------------
 * Per-element operators (fun_DG.c:1071-1192).
 * ---------------------------------------------------------------------------------------------- */
void orc_elastic_ab(long tri, const double *lambda, const double *miu, const double *den, double *A, double *B)
{
    for (long e = 0; e < tri; e++) {
        A[e * 5 + 0] = 1 / den[e];
        A[e * 5 + 1] = 1 / den[e];
        A[e * 5 + 2] = lambda[e] + 2 * miu[e];
        A[e * 5 + 3] = lambda[e];
        A[e * 5 + 4] = miu[e];
        B[e * 5 + 0] = 1 / den[e];
        B[e * 5 + 1] = 1 / den[e];
        B[e * 5 + 2] = lambda[e];
        B[e * 5 + 3] = lambda[e] + 2 * miu[e];
        B[e * 5 + 4] = miu[e];
    }
}

/* Q[e][j][k][0|1] = 0.5 * n_(x|z) * (A|B)[e][k] * (L_j/2) / J; rows 1 and 4 swap nx and nz
 * (fun_DG.c:1118-1131). */
void orc_central_flux(long tri, const double *jacobi, const double *jacobi_edge, const double *A, const double *B,
                      const double *n1, const double *n2, const double *n3, double *Q)
{
    for (long e = 0; e < tri; e++) {
        for (int j = 0; j < 3; j++) {
            const double *n = (j == 0) ? &n1[e * 2] : (j == 1) ? &n2[e * 2] : &n3[e * 2];
            double nx = n[0], nz = n[1], je = jacobi_edge[e * 3 + j], J = jacobi[e];
            double *q = &Q[e * 30 + j * 10];
            const double *a = &A[e * 5], *b = &B[e * 5];
            q[0] = 0.5 * nx * a[0] * je / J;
            q[1] = 0.5 * nz * b[0] * je / J;
            q[2] = 0.5 * nz * a[1] * je / J;
            q[3] = 0.5 * nx * b[1] * je / J;
            q[4] = 0.5 * nx * a[2] * je / J;
            q[5] = 0.5 * nz * b[2] * je / J;
            q[6] = 0.5 * nx * a[3] * je / J;
            q[7] = 0.5 * nz * b[3] * je / J;
            q[8] = 0.5 * nz * a[4] * je / J;
            q[9] = 0.5 * nx * b[4] * je / J;
        }
    }
}

/* Dx = (rx Vr + sx Vs) inverV, Dz = (rz Vr + sz Vs) inverV, rx = zs/J, rz = -xs/J, sx = -zr/J,
 * sz = xr/J (fun_DG.c:1148-1191). */
void orc_dx_dz(long tri, const double *point, const long *triangle, const double *Vr, const double *Vs,
               const double *inverV, double *Dx, double *Dz)
{
#pragma omp parallel for schedule(static)
    for (long e = 0; e < tri; e++) {
        double P[3][2];
        tri_vertices(point, triangle, e, P);
        double Xr = (P[1][0] - P[0][0]) * 0.5, Zr = (P[1][1] - P[0][1]) * 0.5;
        double Xs = (P[2][0] - P[0][0]) * 0.5, Zs = (P[2][1] - P[0][1]) * 0.5;
        double J = Xr * Zs - Xs * Zr;
        double Rx = Zs / J, Rz = -Xs / J, Sx = -Zr / J, Sz = Xr / J;
        double gx[NP][NP], gz[NP][NP];
        for (int i = 0; i < NP; i++)
            for (int n = 0; n < NP; n++) {
                gx[i][n] = Rx * Vr[i * NP + n] + Sx * Vs[i * NP + n];
                gz[i][n] = Rz * Vr[i * NP + n] + Sz * Vs[i * NP + n];
            }
        for (int p = 0; p < NP; p++)
            for (int q = 0; q < NP; q++) {
                double f1 = 0, f2 = 0;
                for (int k = 0; k < NP; k++) {
                    f1 = f1 + gx[p][k] * inverV[k * NP + q];
                    f2 = f2 + gz[p][k] * inverV[k * NP + q];
                }
                Dx[e * 36 + p * NP + q] = f1;
                Dz[e * 36 + p * NP + q] = f2;
            }
    }
}

/* Physical node coordinates, x = -(r+s)/2 v0 + (r+1)/2 v1 + (s+1)/2 v2 (fun_DG.c:839-861). */
void orc_all_node(long tri, const double *point, const long *triangle, const double *fekete, double *all_node)
{
    for (long e = 0; e < tri; e++) {
        double P[3][2], pos[NP][2];
        tri_vertices(point, triangle, e, P);
        rs_to_xy(P, fekete, pos);
        for (int i = 0; i < NP; i++) {
            all_node[e * 12 + i * 2 + 0] = pos[i][0];
            all_node[e * 12 + i * 2 + 1] = pos[i][1];
        }
    }
}

/* ------------------------------------------------------------------------------------------------
 * Time-step kernels (OpenMP_Function.c:25-306).
 * ---------------------------------------------------------------------------------------------- */
void orc_absorb(long tri, double *Vx, double *Vz, double *Txx, double *Tzz, double *Txz, const double *absorb_fun)
{
#pragma omp parallel for schedule(static)
    for (long e = 0; e < tri; e++)
        for (int i = 0; i < NP; i++) {
            long id = e * NP + i;
            Vx[id] *= absorb_fun[id];
            Vz[id] *= absorb_fun[id];
            Txx[id] *= absorb_fun[id];
            Tzz[id] *= absorb_fun[id];
            Txz[id] *= absorb_fun[id];
        }
}

void orc_source_kernel(double *Vz, long src_elem, const double *fun, double w)
{
    for (int i = 0; i < NP; i++) Vz[NP * (src_elem - 1) + i] += fun[i] * w; /* OpenMP_Function.c:53-57 */
}

/*
 * Velocity from stress (OpenMP_Function.c:75-185):
 *   Vx += dt*[A0 Dx Txx + B0 Dz Txz + sum_j (Q_j00 S_j dTxx_j + Q_j01 S_j dTxz_j)]
 *   Vz += dt*[A1 Dx Txz + B1 Dz Tzz + sum_j (Q_j10 S_j dTzz_j + Q_j11 S_j dTxz_j)]
 * with the face jump dU_j[i] = flag_j * U[nb][NodeNb_j[i]] - U[e][ibe_j[i]].
 */
void orc_lu_velocity(long tri, double dt, double *Vx, double *Vz, const double *Txx, const double *Tzz,
                     const double *Txz, const double *A, const double *B, const double *Dx, const double *Dz,
                     const double *Q, const double *S, const long *ibe, const long *node_neighbor,
                     const long *neighbor, const long *edge_flag)
{
#pragma omp parallel for schedule(static)
    for (long e = 0; e < tri; e++) {
        const double *dx = &Dx[e * 36], *dz = &Dz[e * 36];
        const double *tx = &Txx[e * NP], *tz = &Tzz[e * NP], *tc = &Txz[e * NP];
        double vol_x[NP], vol_z[NP], fx[NP][3], fz[NP][3];
        for (int i = 0; i < NP; i++) {
            double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
            for (int n = 0; n < NP; n++) {
                s1 += dx[i * NP + n] * tx[n];
                s2 += dx[i * NP + n] * tc[n];
                s3 += dz[i * NP + n] * tz[n];
                s4 += dz[i * NP + n] * tc[n];
            }
            vol_x[i] = A[e * 5 + 0] * s1 + B[e * 5 + 0] * s4;
            vol_z[i] = A[e * 5 + 1] * s2 + B[e * 5 + 1] * s3;
        }
        for (int j = 0; j < 3; j++) {
            long nb = neighbor[e * 3 + j] - 1;
            double flag = (double)edge_flag[e * 3 + j];
            double jxx[3], jzz[3], jxz[3];
            for (int i = 0; i < 3; i++) {
                long o1 = NP * nb + node_neighbor[e * 9 + j * 3 + i] - 1;
                long o2 = NP * e + ibe[j * 3 + i] - 1;
                jxx[i] = Txx[o1] * flag - Txx[o2];
                jzz[i] = Tzz[o1] * flag - Tzz[o2];
                jxz[i] = Txz[o1] * flag - Txz[o2];
            }
            const double *q = &Q[e * 30 + j * 10];
            for (int i = 0; i < NP; i++) {
                double a = 0, b = 0, c = 0;
                for (int n = 0; n < 3; n++) {
                    double sv = S[j * 18 + i * 3 + n];
                    a += sv * jxx[n];
                    b += sv * jzz[n];
                    c += sv * jxz[n];
                }
                fx[i][j] = q[0] * a + q[1] * c;
                fz[i][j] = q[2] * b + q[3] * c;
            }
        }
        for (int i = 0; i < NP; i++) {
            long o = NP * e + i;
            Vx[o] = (vol_x[i] + fx[i][0] + fx[i][1] + fx[i][2]) * dt + Vx[o];
            Vz[o] = (vol_z[i] + fz[i][0] + fz[i][1] + fz[i][2]) * dt + Vz[o];
        }
    }
}

/*
 * Stress from velocity (OpenMP_Function.c:188-306):
 *   Txx += dt*[A2 Dx Vx + B2 Dz Vz + sum_j(Q_j20 S_j dVx + Q_j21 S_j dVz)]
 *   Tzz += dt*[A3 Dx Vx + B3 Dz Vz + sum_j(Q_j30 ... + Q_j31 ...)]
 *   Txz += dt*[A4 Dx Vz + B4 Dz Vx + sum_j(Q_j40 ... + Q_j41 ...)]
 */
void orc_lu_stress(long tri, double dt, double *Txx, double *Tzz, double *Txz, const double *Vx, const double *Vz,
                   const double *A, const double *B, const double *Dx, const double *Dz, const double *Q,
                   const double *S, const long *ibe, const long *node_neighbor, const long *neighbor,
                   const long *edge_flag)
{
#pragma omp parallel for schedule(static)
    for (long e = 0; e < tri; e++) {
        const double *dx = &Dx[e * 36], *dz = &Dz[e * 36];
        const double *vx = &Vx[e * NP], *vz = &Vz[e * NP];
        double vxx[NP], vzz[NP], vxz[NP], fxx[NP][3], fzz[NP][3], fxz[NP][3];
        for (int i = 0; i < NP; i++) {
            double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
            for (int n = 0; n < NP; n++) {
                s1 += dx[i * NP + n] * vx[n];
                s2 += dx[i * NP + n] * vz[n];
                s3 += dz[i * NP + n] * vx[n];
                s4 += dz[i * NP + n] * vz[n];
            }
            vxx[i] = A[e * 5 + 2] * s1 + B[e * 5 + 2] * s4;
            vzz[i] = A[e * 5 + 3] * s1 + B[e * 5 + 3] * s4;
            vxz[i] = A[e * 5 + 4] * s2 + B[e * 5 + 4] * s3;
        }
        for (int j = 0; j < 3; j++) {
            long nb = neighbor[e * 3 + j] - 1;
            double flag = (double)edge_flag[e * 3 + j];
            double jx[3], jz[3];
            for (int i = 0; i < 3; i++) {
                long o1 = NP * nb + node_neighbor[e * 9 + j * 3 + i] - 1;
                long o2 = NP * e + ibe[j * 3 + i] - 1;
                jx[i] = Vx[o1] * flag - Vx[o2];
                jz[i] = Vz[o1] * flag - Vz[o2];
            }
            const double *q = &Q[e * 30 + j * 10];
            for (int i = 0; i < NP; i++) {
                double a = 0, b = 0;
                for (int n = 0; n < 3; n++) {
                    double sv = S[j * 18 + i * 3 + n];
                    a += sv * jx[n];
                    b += sv * jz[n];
                }
                fxx[i][j] = q[4] * a + q[5] * b;
                fzz[i][j] = q[6] * a + q[7] * b;
                fxz[i][j] = q[8] * a + q[9] * b;
            }
        }
        for (int i = 0; i < NP; i++) {
            long o = NP * e + i;
            Txx[o] = (vxx[i] + fxx[i][0] + fxx[i][1] + fxx[i][2]) * dt + Txx[o];
            Tzz[o] = (vzz[i] + fzz[i][0] + fzz[i][1] + fzz[i][2]) * dt + Tzz[o];
            Txz[o] = (vxz[i] + fxz[i][0] + fxz[i][1] + fxz[i][2]) * dt + Txz[o];
        }
    }
}

/* ------------------------------------------------------------------------------------------------
 * [EXT] Point location and P2 interpolation.  Nodes (r,s): 1:(1,-1) 2:(0,0) 3:(-1,1) 4:(0,-1)
 * 5:(-1,0) 6:(-1,-1), i.e. v1, mid(v1v2), v2, mid(v0v1), mid(v0v2), v0 (FeketeNew.txt).  With
 * barycentric l0=-(r+s)/2, l1=(r+1)/2, l2=(s+1)/2 the Lagrange basis is l(2l-1) at a vertex and
 * 4 la lb at an edge midpoint.  A point on a vertex gives exactly the unit vector of that node.
 * ---------------------------------------------------------------------------------------------- */
void orc_p2_basis(double r, double s, double *phi)
{
    double l0 = -(r + s) / 2, l1 = (r + 1) / 2, l2 = (s + 1) / 2;
    phi[0] = l1 * (2 * l1 - 1);
    phi[1] = 4 * l1 * l2;
    phi[2] = l2 * (2 * l2 - 1);
    phi[3] = 4 * l0 * l1;
    phi[4] = 4 * l0 * l2;
    phi[5] = l0 * (2 * l0 - 1);
}

/* First element (in element order) whose closed triangle contains (x,z) within a relative tolerance;
 * returns its 1-based index and the reference coordinates, or 0. */
long orc_locate(long tri, const double *point, const long *triangle, double x, double z, double *r, double *s)
{
    for (long e = 0; e < tri; e++) {
        double P[3][2];
        tri_vertices(point, triangle, e, P);
        double xr = P[1][0] - P[0][0], zr = P[1][1] - P[0][1];
        double xs = P[2][0] - P[0][0], zs = P[2][1] - P[0][1];
        double det = xr * zs - xs * zr;
        double dx = x - P[0][0], dz = z - P[0][1];
        double a = (dx * zs - xs * dz) / det; /* l1 */
        double b = (xr * dz - dx * zr) / det; /* l2 */
        const double tol = 1e-10;
        if (a >= -tol && b >= -tol && a + b <= 1 + tol) {
            *r = 2 * a - 1;
            *s = 2 * b - 1;
            return e + 1;
        }
    }
    return 0;
}

/* fun_i = sum_k (V V^T)_{ik} phi_k / J ; reduces to orc_source_vector when phi is a unit vector. */
void orc_source_vector_at(const double *V, const double *phi, double jacobi, double *fun)
{
    double M[NP][NP];
    for (int i = 0; i < NP; i++)
        for (int k = 0; k < NP; k++) {
            double sum = 0;
            for (int j = 0; j < NP; j++) sum = sum + V[i * NP + j] * V[k * NP + j];
            M[i][k] = sum;
        }
    /* For a unit phi the sum adds exact zeros to M[i][k], so the result is bit-identical to
     * orc_source_vector. */
    for (int i = 0; i < NP; i++) {
        double acc = 0;
        for (int k = 0; k < NP; k++) acc = acc + M[i][k] * phi[k];
        fun[i] = acc / jacobi;
    }
}

/* ------------------------------------------------------------------------------------------------
 * Step composition (DG_leap_Frog_OpenMP.c:405-425): absorb -> source -> velocity -> stress ->
 * record.  [EXT] explosive sources add fun*w to Txx and Tzz at the source step, in the same place
 * where the reference adds its vertical force to Vz; receivers interpolate with 6 weights.
 * ---------------------------------------------------------------------------------------------- */
void orc_step(const orc_problem *p, long step, double *Vx, double *Vz, double *Txx, double *Tzz, double *Txz)
{
    double w = p->wavelet[step];
    if (p->absorb_fun) orc_absorb(p->tri, Vx, Vz, Txx, Tzz, Txz, p->absorb_fun);
    for (long k = 0; k < p->nsrc; k++) {
        const orc_source *s = &p->src[k];
        if (s->kind == 0) {
            orc_source_kernel(Vz, s->elem, s->fun, w);
        } else {
            orc_source_kernel(Txx, s->elem, s->fun, w);
            orc_source_kernel(Tzz, s->elem, s->fun, w);
        }
    }
    orc_lu_velocity(p->tri, p->dt, Vx, Vz, Txx, Tzz, Txz, p->A, p->B, p->Dx, p->Dz, p->Q, p->S, p->ibe,
                    p->node_neighbor, p->neighbor, p->edge_flag);
    orc_lu_stress(p->tri, p->dt, Txx, Tzz, Txz, Vx, Vz, p->A, p->B, p->Dx, p->Dz, p->Q, p->S, p->ibe,
                  p->node_neighbor, p->neighbor, p->edge_flag);
}

/* Record (OpenMP_Function.c:61-72): rec_x[k] = Vx at receiver k, for every receiver of the problem.  [EXT] the
 * value is sum_i weight[i] * V[elem][i] (P2 interpolation); a vertex receiver has one unit weight and reduces
 * to the reference's nodal pick Vx[6(e-1) + n - 1]. */
void orc_record(const orc_problem *p, const double *Vx, const double *Vz, double *rec_x, double *rec_z)
{
    for (long k = 0; k < p->nrec; k++) {
        const orc_receiver *r = &p->rec[k];
        const double *vx = &Vx[(r->elem - 1) * NP], *vz = &Vz[(r->elem - 1) * NP];
        double ax = 0, az = 0;
        for (int i = 0; i < NP; i++) {
            ax += r->weight[i] * vx[i];
            az += r->weight[i] * vz[i];
        }
        if (rec_x) rec_x[k] = ax;
        if (rec_z) rec_z[k] = az;
    }
}

void orc_run(const orc_problem *p, long step0, long nsteps, double *Vx, double *Vz, double *Txx, double *Tzz,
             double *Txz, double *rec_x, double *rec_z)
{
    for (long t = 0; t < nsteps; t++) {
        orc_step(p, step0 + t, Vx, Vz, Txx, Tzz, Txz);
        orc_record(p, Vx, Vz, rec_x ? rec_x + t * p->nrec : NULL, rec_z ? rec_z + t * p->nrec : NULL);
    }
}
