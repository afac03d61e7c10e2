/*
 * dgx_run -- C driver of the solver, the counterpart of the reference's main (DG_leap_Frog_OpenMP.c:32-453)
 * with the step loop on the GPU through libdgx.
 *
 * It reads the reference's inputs (Triangle .ele/.node/.neigh converted to MATLAB-ASCII, the
 * precomputed operator matrices), sets materials, sponge, dt, wavelet, source and receivers as the
 * reference's setup does, advances the wavefield on the device and writes the reference's outputs
 * (output_files/Vx<step>, Vz<step> every `snapshot` steps in Save_Data format, All_node.txt) plus the
 * seismograms record_x / record_z in Save_Record format (fun_save_to_file.c:260-281, which the
 * reference defines but never calls) and a lossless float64 twin (.bin) of every output.
 *
 * Usage: dgx_run <config file>
 * Config: "key value" lines, '#' comments.  Keys (defaults in brackets):
 *   mesh <prefix>            reads <prefix>.ele, .node, .neigh           [input_files/mesh/model.1]
 *   mesh_binary <path>       reads a binary mesh (dgx_write_mesh_binary) instead of the text files
 *   operators <dir>          precomputed_matrix directory               [input_files/precomputed_matrix]
 *   output <dir>             output directory (must exist)               [output_files]
 *   steps <n>                time steps                                  [2500, dg.h:4]
 *   snapshot <n>             snapshot cadence, 0 = none                  [100, DG_leap_Frog_OpenMP.c:433]
 *   f0 <Hz>                  Ricker peak frequency                       [15, dg.h:5]
 *   pi <value>               value of Pi in wavelet and sponge           [3.1415, dg.h:6]
 *   sponge <bx> <bz> <w> <h> <beta>                                      [100 100 1800 1800 0.1]
 *   material <attr> <vp> <vs> <rho>   per-attribute medium (repeatable)  [1,2 water; 3,4 rock as the reference]
 *   source <x> <z> <force|explosive>  (repeatable; all in one shot)      [1200 1200 force]
 *   receiver <x> <z>         one receiver (repeatable)                   [1600 1600]
 *   receiver_line <x0> <z0> <x1> <z1> <n>   n receivers on a segment
 *   precision <fp64|fp32|fp32c|reference>   fp32c: compensated FP32; reference: the reference's arithmetic,
 *                                  outputs bit-identical to its own loop (DGX_FP64_REFERENCE)   [fp64]
 *   step_kernel <auto|half|full|blocked>   two half-step kernels per step, one full-step kernel, or whole
 *                                  steps per launch (auto: blocked, half with precision reference)   [auto]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "dgx.h"
#include "dgx_setup.h"

#define MAXMAT 64
#define MAXSRC 64

typedef struct {
    char mesh[1024], mesh_bin[1024], ops[1024], out[1024];
    long steps, snapshot;
    double f0, pi;
    dgx_sponge_params sp;
    int nmat;
    long mat_attr[MAXMAT];
    double mat_vp[MAXMAT], mat_vs[MAXMAT], mat_rho[MAXMAT];
    int nsrc;
    double src_x[MAXSRC], src_z[MAXSRC];
    int src_kind[MAXSRC];
    long nrec, rec_cap;
    double *rec_xz;
    int precision;
    int step_kernel;
} config;

static double now(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

static void add_receiver(config *c, double x, double z)
{
    if (c->nrec == c->rec_cap) {
        c->rec_cap = c->rec_cap ? 2 * c->rec_cap : 64;
        c->rec_xz = (double *)realloc(c->rec_xz, sizeof(double) * 2 * c->rec_cap);
    }
    c->rec_xz[2 * c->nrec] = x;
    c->rec_xz[2 * c->nrec + 1] = z;
    c->nrec++;
}

static int read_config(const char *path, config *c)
{
    memset(c, 0, sizeof(*c));
    strcpy(c->mesh, "input_files/mesh/model.1");
    strcpy(c->ops, "input_files/precomputed_matrix");
    strcpy(c->out, "output_files");
    c->steps = 2500;
    c->snapshot = 100;
    c->f0 = 15;
    c->pi = 3.1415;
    c->sp = (dgx_sponge_params){100, 100, 1800, 1800, 0.1, 3.1415};
    int user_src = 0, user_rec = 0;
    FILE *f = fopen(path, "r");
    if (!f) { fprintf(stderr, "cannot open config %s\n", path); return -1; }
    char line[2048];
    while (fgets(line, sizeof(line), f)) {
        char *h = strchr(line, '#');
        if (h) *h = 0;
        char key[64], a[1024], kind[32];
        double v[5];
        long n;
        if (sscanf(line, "%63s", key) != 1) continue;
        if (!strcmp(key, "mesh") && sscanf(line, "%*s %1023s", a) == 1) strcpy(c->mesh, a);
        else if (!strcmp(key, "mesh_binary") && sscanf(line, "%*s %1023s", a) == 1) strcpy(c->mesh_bin, a);
        else if (!strcmp(key, "operators") && sscanf(line, "%*s %1023s", a) == 1) strcpy(c->ops, a);
        else if (!strcmp(key, "output") && sscanf(line, "%*s %1023s", a) == 1) strcpy(c->out, a);
        else if (!strcmp(key, "steps") && sscanf(line, "%*s %ld", &n) == 1) c->steps = n;
        else if (!strcmp(key, "snapshot") && sscanf(line, "%*s %ld", &n) == 1) c->snapshot = n;
        else if (!strcmp(key, "f0") && sscanf(line, "%*s %lf", &v[0]) == 1) c->f0 = v[0];
        else if (!strcmp(key, "pi") && sscanf(line, "%*s %lf", &v[0]) == 1) c->pi = c->sp.pi = v[0];
        else if (!strcmp(key, "sponge") && sscanf(line, "%*s %lf %lf %lf %lf %lf", &v[0], &v[1], &v[2], &v[3], &v[4]) == 5)
            c->sp = (dgx_sponge_params){v[0], v[1], v[2], v[3], v[4], c->pi};
        else if (!strcmp(key, "material") && sscanf(line, "%*s %ld %lf %lf %lf", &n, &v[0], &v[1], &v[2]) == 4 &&
                 c->nmat < MAXMAT) {
            c->mat_attr[c->nmat] = n; c->mat_vp[c->nmat] = v[0]; c->mat_vs[c->nmat] = v[1]; c->mat_rho[c->nmat] = v[2];
            c->nmat++;
        } else if (!strcmp(key, "source") && sscanf(line, "%*s %lf %lf %31s", &v[0], &v[1], kind) == 3 &&
                   c->nsrc < MAXSRC) {
            if (!user_src) { c->nsrc = 0; user_src = 1; }
            c->src_x[c->nsrc] = v[0]; c->src_z[c->nsrc] = v[1];
            c->src_kind[c->nsrc] = !strcmp(kind, "explosive") ? DGX_SRC_EXPLOSIVE : DGX_SRC_FORCE_Z;
            c->nsrc++;
        } else if (!strcmp(key, "receiver") && sscanf(line, "%*s %lf %lf", &v[0], &v[1]) == 2) {
            user_rec = 1;
            add_receiver(c, v[0], v[1]);
        } else if (!strcmp(key, "receiver_line") &&
                   sscanf(line, "%*s %lf %lf %lf %lf %ld", &v[0], &v[1], &v[2], &v[3], &n) == 5 && n > 0) {
            user_rec = 1;
            for (long k = 0; k < n; k++) {
                double t = n > 1 ? (double)k / (n - 1) : 0.0;
                add_receiver(c, v[0] + t * (v[2] - v[0]), v[1] + t * (v[3] - v[1]));
            }
        } else if (!strcmp(key, "precision") && sscanf(line, "%*s %31s", kind) == 1 &&
                   (!strcmp(kind, "fp64") || !strcmp(kind, "fp32") || !strcmp(kind, "fp32c") ||
                    !strcmp(kind, "reference")))
            c->precision = !strcmp(kind, "fp32")        ? DGX_FP32
                           : !strcmp(kind, "fp32c")     ? DGX_FP32_COMPENSATED
                           : !strcmp(kind, "reference") ? DGX_FP64_REFERENCE
                                                        : DGX_FP64;
        else if (!strcmp(key, "step_kernel") && sscanf(line, "%*s %31s", kind) == 1)
            c->step_kernel = !strcmp(kind, "half")      ? DGX_STEP_HALF
                             : !strcmp(kind, "full")    ? DGX_STEP_FULL
                             : !strcmp(kind, "blocked") ? DGX_STEP_BLOCKED
                                                        : DGX_STEP_AUTO;
        else {
            fprintf(stderr, "config: cannot parse line: %s", line);
            fclose(f);
            return -1;
        }
    }
    fclose(f);
    if (c->nmat == 0) { /* DG_leap_Frog_OpenMP.c:236-265 */
        long attrs[4] = {1, 2, 3, 4};
        for (int k = 0; k < 4; k++) {
            c->mat_attr[k] = attrs[k];
            c->mat_vp[k] = k < 2 ? 1500 : 3000;
            c->mat_vs[k] = k < 2 ? 0 : 1800;
            c->mat_rho[k] = k < 2 ? 1000 : 2200;
        }
        c->nmat = 4;
    }
    if (!user_src) { /* DG_leap_Frog_OpenMP.c:207 */
        c->nsrc = 1; c->src_x[0] = 1200; c->src_z[0] = 1200; c->src_kind[0] = DGX_SRC_FORCE_Z;
    }
    if (!user_rec) add_receiver(c, 1600, 1600); /* DG_leap_Frog_OpenMP.c:212 */
    return 0;
}

#define CHECK(call)                                                                                            \
    do {                                                                                                       \
        long rc_ = (long)(call);                                                                               \
        if (rc_ < 0) {                                                                                         \
            fprintf(stderr, "%s failed (%ld): %s\n", #call, rc_, dgx_last_error());                             \
            return 1;                                                                                          \
        }                                                                                                      \
    } while (0)

static int save_pair(const char *dir, const char *name, const double *m, long rows, long cols)
{
    char p[2048];
    snprintf(p, sizeof(p), "%s/%s", dir, name);
    if (dgx_save_data(p, m, rows, cols) < 0) return -1;
    snprintf(p, sizeof(p), "%s/%s.bin", dir, name);
    return dgx_save_binary(p, m, rows * cols);
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: %s <config>\n", argv[0]);
        return 2;
    }
    config c;
    if (read_config(argv[1], &c)) return 2;

    /* mesh (Calculate_Tri, Calculate_Point_Num, Mesh_Input; DG_leap_Frog_OpenMP.c:84-85,190) */
    char ele[1100], node[1100], neigh[1100];
    snprintf(ele, sizeof(ele), "%s.ele", c.mesh);
    snprintf(node, sizeof(node), "%s.node", c.mesh);
    snprintf(neigh, sizeof(neigh), "%s.neigh", c.mesh);
    long tri, np;
    if (c.mesh_bin[0]) {
        if (dgx_mesh_binary_counts(c.mesh_bin, &tri, &np)) tri = np = -1;
    } else {
        tri = dgx_mesh_count(ele);
        np = dgx_mesh_count(node);
    }
    if (tri <= 0 || np <= 0) {
        fprintf(stderr, "cannot read mesh %s: %s\n", c.mesh_bin[0] ? c.mesh_bin : c.mesh, dgx_last_error());
        return 1;
    }
    printf("Triangle number %ld\nPoint number %ld\n", tri, np);
    double *point = malloc(sizeof(double) * np * 2);
    long *triangle = malloc(sizeof(long) * tri * 3), *neighbor = malloc(sizeof(long) * tri * 3);
    long *attribute = malloc(sizeof(long) * tri);
    if (c.mesh_bin[0]) CHECK(dgx_read_mesh_binary(c.mesh_bin, tri, np, point, triangle, neighbor, attribute));
    else CHECK(dgx_read_mesh(ele, node, neigh, tri, np, point, triangle, neighbor, attribute));
    dgx_operators ops;
    CHECK(dgx_read_operators(c.ops, &ops));

    /* materials by attribute (DG_leap_Frog_OpenMP.c:236-265) */
    double *lam = malloc(sizeof(double) * tri), *miu = malloc(sizeof(double) * tri);
    double *den = malloc(sizeof(double) * tri), *vp = malloc(sizeof(double) * tri);
    for (long e = 0; e < tri; e++) {
        int k = -1;
        for (int m = 0; m < c.nmat; m++)
            if (c.mat_attr[m] == attribute[e]) k = m;
        if (k < 0) { fprintf(stderr, "element %ld: no material for attribute %ld\n", e + 1, attribute[e]); return 1; }
        den[e] = c.mat_rho[k];
        miu[e] = den[e] * c.mat_vs[k] * c.mat_vs[k];
        lam[e] = den[e] * c.mat_vp[k] * c.mat_vp[k] - 2 * miu[e];
        vp[e] = c.mat_vp[k];
    }

    /* dt, delay, wavelet, sponge (DG_leap_Frog_OpenMP.c:268-287) */
    double dt = dgx_cfl_dt(tri, point, triangle, vp);
    double t0 = dgx_ricker_delay(dt, c.f0, c.pi);
    printf("step dt  %f delay t0 %f\n", dt, t0);
    double *wav = malloc(sizeof(double) * c.steps);
    CHECK(dgx_wavelet(c.steps, dt, t0, c.f0, c.pi, wav));
    double *absorb = malloc(sizeof(double) * tri * 6);
    CHECK(dgx_sponge(tri, point, triangle, ops.FeketeNew, &c.sp, absorb));

    /* source and receivers (Get_Source, Get_Receiver, DG_leap_Frog_OpenMP.c:207-233), searched on the GPU */
    dgx_source src[MAXSRC];
    double src_xz[2 * MAXSRC];
    int src_shot[MAXSRC];
    long src_wav[MAXSRC];
    for (int k = 0; k < c.nsrc; k++) {
        src_xz[2 * k] = c.src_x[k];
        src_xz[2 * k + 1] = c.src_z[k];
        src_shot[k] = 0;
        src_wav[k] = 0;
    }
    CHECK(dgx_locate_sources_device(tri, np, point, triangle, &ops, c.nsrc, src_xz, c.src_kind, src_shot, src_wav, src));
    dgx_receiver *rec = malloc(sizeof(dgx_receiver) * (c.nrec ? c.nrec : 1));
    CHECK(dgx_locate_receivers_device(tri, np, point, triangle, c.nrec, c.rec_xz, rec));
    for (int k = 0; k < c.nsrc; k++) printf("source[%d]  element %ld\n", k, src[k].elem);

    /* node coordinates for plotting (set_All_node + Save_Data, DG_leap_Frog_OpenMP.c:363-367) */
    double *all_node = malloc(sizeof(double) * tri * 12);
    dgx_all_node(tri, point, triangle, ops.FeketeNew, all_node);
    if (save_pair(c.out, "All_node.txt", all_node, tri * 6, 2)) { fprintf(stderr, "cannot write to %s\n", c.out); return 1; }

    dgx_problem p = {tri, np, point, triangle, neighbor, lam, miu, den, absorb, dt, c.precision, 1, c.steps, 1, wav,
                     c.nsrc, src, c.nrec, rec, DGX_LAYOUT_ELEMENT, 0, c.step_kernel};
    dgx_solver *s;
    CHECK(dgx_create(&p, &ops, &s));

    /* time loop (DG_leap_Frog_OpenMP.c:405-446): step on the device, snapshots at the cadence.  A snapshot
     * is begun when its step is reached and collected (and written) after the next block of steps has been
     * enqueued, so the download and the text output overlap the stepping. */
    double *vx = malloc(sizeof(double) * tri * 6), *vz = malloc(sizeof(double) * tri * 6);
    double t_io = 0;
    long done = 0, pending = -1; /* step of the snapshot in flight, -1 = none */
    double t_start = now();
    while (done < c.steps || pending >= 0) {
        long n = c.snapshot > 0 ? c.snapshot - done % c.snapshot : c.steps - done;
        if (n > c.steps - done) n = c.steps - done;
        if (n > 0) CHECK(dgx_step(s, n));
        if (pending >= 0) {
            double a = now();
            char name[64];
            CHECK(dgx_snapshot_wait(s, vx, vz, NULL, NULL, NULL));
            snprintf(name, sizeof(name), "Vx%ld", pending);
            CHECK(save_pair(c.out, name, vx, tri * 6, 1));
            snprintf(name, sizeof(name), "Vz%ld", pending);
            CHECK(save_pair(c.out, name, vz, tri * 6, 1));
            t_io += now() - a;
            pending = -1;
        }
        if (n == 0) break;
        done += n;
        if (c.snapshot > 0 && done % c.snapshot == 0) {
            CHECK(dgx_snapshot_begin(s, 0, 3)); /* Vx, Vz */
            pending = done;
        }
        printf("time:%ld,\tenqueued\n", done - 1);
    }
    CHECK(dgx_synchronize(s));
    double t_loop = now() - t_start;

    /* seismograms: record_x/record_z are [steps][nrec]; Save_Record writes one column per receiver */
    double *rx = malloc(sizeof(double) * c.steps * (c.nrec ? c.nrec : 1));
    double *rz = malloc(sizeof(double) * c.steps * (c.nrec ? c.nrec : 1));
    CHECK(dgx_get_traces(s, 0, rx, rz));
    double *col = malloc(sizeof(double) * c.steps);
    for (long k = 0; k < c.nrec; k++) {
        char name[64];
        for (long t = 0; t < c.steps; t++) col[t] = rx[t * c.nrec + k];
        snprintf(name, sizeof(name), c.nrec == 1 ? "record_x" : "record_x%ld", k);
        CHECK(save_pair(c.out, name, col, c.steps, 1));
        for (long t = 0; t < c.steps; t++) col[t] = rz[t * c.nrec + k];
        snprintf(name, sizeof(name), c.nrec == 1 ? "record_z" : "record_z%ld", k);
        CHECK(save_pair(c.out, name, col, c.steps, 1));
    }
    double dofs = (double)tri * 30.0 * c.steps;
    printf("steps %ld  time loop %.3f s with snapshot output overlapped (%.3e DOF-updates/s)  host time in "
           "snapshot collection and output %.3f s\n", c.steps, t_loop, dofs / t_loop, t_io);
    dgx_destroy(s);
    return 0;
}
