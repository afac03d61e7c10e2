/*
 * dgx_setup.c -- host setup and I/O for libdgx (declared in include/dgx_setup.h).
 *
 * Compiled with gcc -O2 and no -march, like the reference's OpenMP build, so the few quantities that
 * the reference computes on the host with libm (lengths via pow/sqrt, the CFL step, the Ricker
 * samples, the sponge cosines) come out bit-identical to the reference's when the same steps are
 * taken.  Those steps are kept where the reference's result depends on them; the data structures
 * around them (streaming reader, grid locator, vertex/element incidence) are this library's own.
 */
#include "dgx_setup.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void dgx_set_error(const char *fmt, ...);

/* ------------------------------------------------------------------------------------------------
 * MATLAB "save -ascii" text.  Each value is a 16-byte field "   d.ddddddde+XX" ("  -d..." when
 * negative).  The value is rebuilt from its digits the way the reference does it (lead digit, then
 * the seven fraction digits from the last to the first with weights 10^-7..10^-1, then times
 * 10^exp times the sign, fun_save_to_file.c:83-93), not with strtod, so both readers yield the same
 * doubles.  Integers weight the digits by the exponent (fun_save_to_file.c:177-185).
 * ---------------------------------------------------------------------------------------------- */
static char *slurp(const char *path, long *len)
{
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    if (fseek(f, 0, SEEK_END) != 0) { fclose(f); return NULL; }
    long n = ftell(f);
    if (n < 0) { fclose(f); return NULL; }
    rewind(f);
    char *buf = (char *)malloc((size_t)n + 1);
    if (!buf) { fclose(f); return NULL; }
    if (n > 0 && fread(buf, 1, (size_t)n, f) != (size_t)n) { free(buf); fclose(f); return NULL; }
    fclose(f);
    buf[n] = 0;
    *len = n;
    return buf;
}

static double field_double(const char *s, long i, long *sign)
{
    if (s[i - 10] == ' ') *sign = 1;
    else if (s[i - 10] == '-') *sign = -1;
    long esign = (s[i + 1] == '-') ? -1 : 1;
    double v = (double)(s[i - 9] - '0');
    for (long j = i - 1, k = -7; j >= i - 7; j--, k++) v = v + (s[j] - '0') * pow(10, k);
    long p = esign * ((s[i + 2] - '0') * 10 + (s[i + 3] - '0'));
    return v * pow(10, p) * (*sign);
}

static long field_long(const char *s, long i, long *sign)
{
    if (s[i - 10] == ' ') *sign = 1;
    else if (s[i - 10] == '-') *sign = -1;
    long esign = (s[i + 1] == '-') ? -1 : 1;
    long p = esign * ((s[i + 2] - '0') * 10 + (s[i + 3] - '0'));
    long v = (long)((s[i - 9] - '0') * pow(10, p) * (*sign));
    for (long j = i - 7, k = p - 1; j <= i - 8 + p; j++, k--) v = (long)(v + (s[j] - '0') * pow(10, k));
    return v;
}

/* Scan text, calling back per value with its (row, col).  Returns values seen (capped at want). */
typedef void (*field_cb)(const char *s, long i, long row, long col, void *ctx);
static long scan_fields(const char *s, long len, long want, field_cb cb, void *ctx)
{
    long got = 0, row = 0, col = 0;
    for (long i = 0; i < len && got < want; i++) {
        char c = s[i];
        if (c == 'e' && i >= 10 && i + 3 < len) {
            cb(s, i, row, col, ctx);
            col++;
            got++;
        } else if (c == '\n') {
            row++;
            col = 0;
        }
    }
    return got;
}

typedef struct { double *m; long rows, cols; long sign; } dctx;
typedef struct { long *m; long rows, cols; long sign; } lctx;
static void put_double(const char *s, long i, long row, long col, void *c)
{
    dctx *d = (dctx *)c;
    double v = field_double(s, i, &d->sign);
    if (row < d->rows && col < d->cols) d->m[row * d->cols + col] = v;
}
static void put_long(const char *s, long i, long row, long col, void *c)
{
    lctx *d = (lctx *)c;
    long v = field_long(s, i, &d->sign);
    if (row < d->rows && col < d->cols) d->m[row * d->cols + col] = v;
}

long dgx_read_ascii(const char *path, double *m, long rows, long cols)
{
    long len = 0;
    char *buf = slurp(path, &len);
    if (!buf) { dgx_set_error("cannot read %s", path); return DGX_EIO; }
    dctx c = {m, rows, cols, 1};
    long got = scan_fields(buf, len, rows * cols, put_double, &c);
    free(buf);
    if (got != rows * cols) { dgx_set_error("%s: %ld of %ld values", path, got, rows * cols); return DGX_EIO; }
    return got;
}

long dgx_read_ascii_long(const char *path, long *m, long rows, long cols)
{
    long len = 0;
    char *buf = slurp(path, &len);
    if (!buf) { dgx_set_error("cannot read %s", path); return DGX_EIO; }
    lctx c = {m, rows, cols, 1};
    long got = scan_fields(buf, len, rows * cols, put_long, &c);
    free(buf);
    if (got != rows * cols) { dgx_set_error("%s: %ld of %ld values", path, got, rows * cols); return DGX_EIO; }
    return got;
}

int dgx_read_operators(const char *dir, dgx_operators *ops)
{
    struct { const char *name; double *dst; long r, c; } tab[] = {
        {"V.txt", ops->V, 6, 6},         {"inverV.txt", ops->inverV, 6, 6}, {"Vr.txt", ops->Vr, 6, 6},
        {"Vs.txt", ops->Vs, 6, 6},       {"S1.txt", ops->S1, 6, 3},         {"S2.txt", ops->S2, 6, 3},
        {"S3.txt", ops->S3, 6, 3},       {"FeketeNew.txt", ops->FeketeNew, 6, 2},
    };
    char path[4096];
    for (size_t k = 0; k < sizeof(tab) / sizeof(tab[0]); k++) {
        snprintf(path, sizeof(path), "%s/%s", dir, tab[k].name);
        long rc = dgx_read_ascii(path, tab[k].dst, tab[k].r, tab[k].c);
        if (rc < 0) return (int)rc;
    }
    return DGX_OK;
}

long dgx_mesh_count(const char *path)
{
    long v = 0;
    long rc = dgx_read_ascii_long(path, &v, 1, 1);
    return rc < 0 ? rc : v;
}

int dgx_read_mesh(const char *ele, const char *node, const char *neigh, long tri, long point_num, double *point,
                  long *triangle, long *neighbor, long *attribute)
{
    long *E = (long *)malloc(sizeof(long) * (tri + 1) * 5);
    long *G = (long *)malloc(sizeof(long) * (tri + 1) * 4);
    double *P = (double *)malloc(sizeof(double) * (point_num + 1) * 4);
    int rc = DGX_OK;
    if (!E || !G || !P) { rc = DGX_ENOMEM; goto done; }
    if (dgx_read_ascii_long(ele, E, tri + 1, 5) < 0 || dgx_read_ascii_long(neigh, G, tri + 1, 4) < 0 ||
        dgx_read_ascii(node, P, point_num + 1, 4) < 0) {
        rc = DGX_EIO;
        goto done;
    }
    /* row 0 is the header; columns: id, v1, v2, v3, attribute / id, n1, n2, n3 / id, x, z, marker */
    for (long e = 0; e < tri; e++) {
        for (int j = 0; j < 3; j++) {
            triangle[e * 3 + j] = E[(e + 1) * 5 + 1 + j];
            neighbor[e * 3 + j] = G[(e + 1) * 4 + 1 + j];
        }
        if (attribute) attribute[e] = E[(e + 1) * 5 + 4];
    }
    for (long i = 0; i < point_num; i++) {
        point[i * 2 + 0] = P[(i + 1) * 4 + 1];
        point[i * 2 + 1] = P[(i + 1) * 4 + 2];
    }
done:
    free(E); free(G); free(P);
    return rc;
}

/* ------------------------------------------------------------------------------------------------
 * Binary mesh: header + four arrays, moved through a fixed-size buffer.
 * ---------------------------------------------------------------------------------------------- */
static const char MESH_MAGIC[8] = {'D', 'G', 'X', 'M', 'E', 'S', 'H', '1'};

static int write_chunked(FILE *f, const void *src, size_t elem, size_t n, int widen_long)
{
    /* widen_long: src holds C longs, the file holds int64 */
    if (!widen_long || sizeof(long) == 8) return fwrite(src, elem, n, f) == n ? 0 : -1;
    int64_t buf[4096];
    const long *l = (const long *)src;
    for (size_t i = 0; i < n; i += 4096) {
        size_t k = n - i < 4096 ? n - i : 4096;
        for (size_t j = 0; j < k; j++) buf[j] = (int64_t)l[i + j];
        if (fwrite(buf, 8, k, f) != k) return -1;
    }
    return 0;
}

static int read_chunked(FILE *f, void *dst, size_t elem, size_t n, int narrow_long)
{
    if (!narrow_long || sizeof(long) == 8) return fread(dst, elem, n, f) == n ? 0 : -1;
    int64_t buf[4096];
    long *l = (long *)dst;
    for (size_t i = 0; i < n; i += 4096) {
        size_t k = n - i < 4096 ? n - i : 4096;
        if (fread(buf, 8, k, f) != k) return -1;
        for (size_t j = 0; j < k; j++) l[i + j] = (long)buf[j];
    }
    return 0;
}

int dgx_write_mesh_binary(const char *path, long tri, long point_num, const double *point, const long *triangle,
                          const long *neighbor, const long *attribute)
{
    if (!path || tri <= 0 || point_num <= 0 || !point || !triangle || !neighbor) return DGX_EARG;
    FILE *f = fopen(path, "wb");
    if (!f) { dgx_set_error("cannot write %s", path); return DGX_EIO; }
    int64_t hdr[2] = {tri, point_num};
    int bad = fwrite(MESH_MAGIC, 1, 8, f) != 8 || fwrite(hdr, 8, 2, f) != 2;
    bad = bad || write_chunked(f, point, 8, (size_t)point_num * 2, 0);
    bad = bad || write_chunked(f, triangle, sizeof(long), (size_t)tri * 3, 1);
    bad = bad || write_chunked(f, neighbor, sizeof(long), (size_t)tri * 3, 1);
    if (!bad) {
        if (attribute) {
            bad = write_chunked(f, attribute, sizeof(long), (size_t)tri, 1);
        } else {
            int64_t one[1024];
            for (int j = 0; j < 1024; j++) one[j] = 1;
            for (long i = 0; i < tri && !bad; i += 1024) {
                size_t k = (size_t)(tri - i < 1024 ? tri - i : 1024);
                bad = fwrite(one, 8, k, f) != k;
            }
        }
    }
    if (fclose(f) != 0) bad = 1;
    if (bad) { dgx_set_error("write error on %s", path); return DGX_EIO; }
    return DGX_OK;
}

static FILE *open_mesh_binary(const char *path, long *tri, long *point_num)
{
    FILE *f = fopen(path, "rb");
    if (!f) { dgx_set_error("cannot read %s", path); return NULL; }
    char magic[8];
    int64_t hdr[2];
    if (fread(magic, 1, 8, f) != 8 || memcmp(magic, MESH_MAGIC, 8) != 0 || fread(hdr, 8, 2, f) != 2 || hdr[0] <= 0 ||
        hdr[1] <= 0) {
        dgx_set_error("%s is not a binary mesh file", path);
        fclose(f);
        return NULL;
    }
    *tri = (long)hdr[0];
    *point_num = (long)hdr[1];
    return f;
}

int dgx_mesh_binary_counts(const char *path, long *tri, long *point_num)
{
    if (!path || !tri || !point_num) return DGX_EARG;
    FILE *f = open_mesh_binary(path, tri, point_num);
    if (!f) return DGX_EIO;
    fclose(f);
    return DGX_OK;
}

int dgx_read_mesh_binary(const char *path, long tri, long point_num, double *point, long *triangle, long *neighbor,
                         long *attribute)
{
    if (!path || !point || !triangle || !neighbor) return DGX_EARG;
    long t, np;
    FILE *f = open_mesh_binary(path, &t, &np);
    if (!f) return DGX_EIO;
    if (t != tri || np != point_num) {
        dgx_set_error("%s holds %ld elements and %ld points, not %ld and %ld", path, t, np, tri, point_num);
        fclose(f);
        return DGX_EARG;
    }
    int bad = read_chunked(f, point, 8, (size_t)np * 2, 0) || read_chunked(f, triangle, sizeof(long), (size_t)t * 3, 1) ||
              read_chunked(f, neighbor, sizeof(long), (size_t)t * 3, 1);
    if (!bad && attribute) bad = read_chunked(f, attribute, sizeof(long), (size_t)t, 1);
    fclose(f);
    if (bad) { dgx_set_error("%s is truncated", path); return DGX_EIO; }
    return DGX_OK;
}

/* ------------------------------------------------------------------------------------------------
 * Writers.
 * ---------------------------------------------------------------------------------------------- */
int dgx_save_data(const char *path, const double *m, long rows, long cols)
{
    FILE *f = fopen(path, "wb");
    if (!f) { dgx_set_error("cannot write %s", path); return DGX_EIO; }
    for (long i = 0; i < rows; i++) {
        for (long j = 0; j < cols; j++) fprintf(f, "%0.14f ", m[i * cols + j]);
        fputc('\n', f);
    }
    return fclose(f) == 0 ? DGX_OK : DGX_EIO;
}

int dgx_save_record(const char *path, const double *m, long rows) { return dgx_save_data(path, m, rows, 1); }

int dgx_save_binary(const char *path, const double *m, long n)
{
    FILE *f = fopen(path, "wb");
    if (!f) { dgx_set_error("cannot write %s", path); return DGX_EIO; }
    size_t w = fwrite(m, sizeof(double), (size_t)n, f);
    int ok = (fclose(f) == 0) && (w == (size_t)n);
    return ok ? DGX_OK : DGX_EIO;
}

/* Node coordinates: x = -(r+s)/2 v0 + (r+1)/2 v1 + (s+1)/2 v2, evaluated in this order so the
 * coordinates the sponge tests see are the reference's (fun_DG.c:857-858, 873-874). */
static void node_xy(const double v[3][2], const double *fekete, int i, double *xy)
{
    double r = fekete[i * 2 + 0], s = fekete[i * 2 + 1];
    for (int c = 0; c < 2; c++) xy[c] = -(r + s) / 2 * v[0][c] + (r + 1) / 2 * v[1][c] + (s + 1) / 2 * v[2][c];
}

static void element_vertices(const double *point, const long *triangle, long e, double v[3][2])
{
    for (int k = 0; k < 3; k++) {
        long p = triangle[e * 3 + k] - 1;
        v[k][0] = point[p * 2 + 0];
        v[k][1] = point[p * 2 + 1];
    }
}

void dgx_all_node(long tri, const double *point, const long *triangle, const double *fekete, double *all_node)
{
    for (long e = 0; e < tri; e++) {
        double v[3][2];
        element_vertices(point, triangle, e, v);
        for (int i = 0; i < 6; i++) node_xy(v, fekete, i, &all_node[e * 12 + i * 2]);
    }
}

/* ------------------------------------------------------------------------------------------------
 * Sponge.  As a function of a node's position (x,z) relative to the interior box, with d the
 * distance to the box (to its corner point in the corner regions):
 *   side strips (open intervals along the box side): f = 0.5 + 0.5 cos(beta d pi / L_side)
 *   corner regions, d < bound_x:                     f = 0.5 + 0.5 cos(beta d pi / bound_x)
 *   otherwise f = 1, also for corner nodes with d >= bound_x and nodes on a strip's end line.
 * Only elements whose first vertex lies outside the box are visited (fun_DG.c:266), which equals
 * "any vertex outside" for the meshes this library accepts (checked below).
 * ---------------------------------------------------------------------------------------------- */
static int outside_box(const double *p, const dgx_sponge_params *sp)
{
    return p[0] > (sp->bound_x + sp->model_w) || p[0] < sp->bound_x || p[1] > (sp->bound_z + sp->model_h) ||
           p[1] < sp->bound_z;
}

static double taper(double d, double L, const dgx_sponge_params *sp)
{
    return 0.5 + 0.5 * cos(sp->beta * d * sp->pi / L);
}

static double sponge_at(double x, double z, const dgx_sponge_params *sp)
{
    const double x0 = sp->bound_x, x1 = sp->bound_x + sp->model_w;
    const double z0 = sp->bound_z, z1 = sp->bound_z + sp->model_h;
    const int in_x = (x > x0 && x < x1), in_z = (z > z0 && z < z1);
    if (x < x0 && in_z) return taper(x0 - x, sp->bound_x, sp);
    if (x > x1 && in_z) return taper(x - x0 - sp->model_w, sp->bound_x, sp);
    if (z < z0 && in_x) return taper(z0 - z, sp->bound_z, sp);
    if (z > z1 && in_x) return taper(z - z0 - sp->model_h, sp->bound_z, sp);
    if ((x < x0 || x > x1) && (z < z0 || z > z1)) {
        double cx = (x < x0) ? x0 : x1, cz = (z < z0) ? z0 : z1;
        double dx = fabs(x - cx), dz = fabs(z - cz);
        double d = sqrt(dx * dx + dz * dz);
        if (d < sp->bound_x) return taper(d, sp->bound_x, sp);
    }
    return 1.0;
}

long dgx_sponge(long tri, const double *point, const long *triangle, const double *fekete,
                const dgx_sponge_params *sp, double *absorb_fun)
{
    long touched = 0;
    for (long e = 0; e < tri; e++) {
        const long *v = &triangle[e * 3];
        int first = outside_box(&point[(v[0] - 1) * 2], sp);
        int any = first || outside_box(&point[(v[1] - 1) * 2], sp) || outside_box(&point[(v[2] - 1) * 2], sp);
        if (any && !first) {
            dgx_set_error("element %ld has a vertex outside the sponge box but not its first vertex; rotate its "
                          "vertex list (the reference's absorbing-element list would be short)", e + 1);
            return DGX_EMESH;
        }
        double vx[3][2];
        element_vertices(point, triangle, e, vx);
        int any_damped = 0;
        for (int i = 0; i < 6; i++) {
            double f = 1.0;
            if (first) {
                double xy[2];
                node_xy(vx, fekete, i, xy);
                f = sponge_at(xy[0], xy[1], sp);
            }
            absorb_fun[e * 6 + i] = f;
            any_damped |= (f != 1.0);
        }
        touched += any_damped;
    }
    return touched;
}

/* ------------------------------------------------------------------------------------------------
 * CFL step and Ricker wavelet.
 * ---------------------------------------------------------------------------------------------- */
double dgx_cfl_dt(long tri, const double *point, const long *triangle, const double *vp)
{
    double vmax = vp[0], dmin = 0;
    for (long e = 1; e < tri; e++) vmax = vp[e] > vmax ? vp[e] : vmax;
    for (long e = 0; e < tri; e++) {
        double v[3][2];
        element_vertices(point, triangle, e, v);
        double a = sqrt(pow(v[0][0] - v[1][0], 2) + pow(v[0][1] - v[1][1], 2));
        double b = sqrt(pow(v[1][0] - v[2][0], 2) + pow(v[1][1] - v[2][1], 2));
        double c = sqrt(pow(v[0][0] - v[2][0], 2) + pow(v[0][1] - v[2][1], 2));
        double s = 0.5 * (a + b + c);
        double area = sqrt(s * (s - a) * (s - b) * (s - c));
        double d = 2 * area / s; /* in-circle diameter */
        if (e == 0 || d < dmin) dmin = d;
    }
    return dmin / vmax / (2 * 2 + 1) * 0.99;
}

double dgx_ricker(long step, double dt, double t0, double f0, double pi)
{
    double t = dt * step, u = t - t0;
    return (1 - 2 * pi * pi * f0 * f0 * u * u) * exp(-pi * pi * f0 * f0 * u * u);
}

double dgx_ricker_delay(double dt, double f0, double pi)
{
    /* first step T+1 at which the undelayed wavelet rises while still below 1e-8 */
    double prev = dgx_ricker(0, dt, 0.0, f0, pi);
    for (long T = 0; T < 100000000L; T++) {
        double t2 = (T + 1) * dt;
        double cur = dgx_ricker(T + 1, dt, 0.0, f0, pi);
        if (cur > prev && cur < 1e-8) return t2 * 2.6;
        prev = cur;
    }
    return 0.0;
}

int dgx_wavelet(long nsteps, double dt, double t0, double f0, double pi, double *w)
{
    if (nsteps < 0 || !w) return DGX_EARG;
    for (long n = 0; n < nsteps; n++) w[n] = dgx_ricker(n, dt, t0, f0, pi);
    return DGX_OK;
}

/* ------------------------------------------------------------------------------------------------
 * Point location: uniform grid over the mesh bounding box; each cell lists, in increasing element
 * order, the elements whose bounding box overlaps it, and the vertices inside it.
 * ---------------------------------------------------------------------------------------------- */
struct dgx_locator {
    long tri, point_num;
    const double *point;
    const long *triangle;
    double x0, z0, h;
    long nx, nz;
    long *cell_start, *cell_elem;    /* element CSR per cell */
    long *vcell_start, *vcell_vert;  /* vertex CSR per cell */
    long *inc_start, *inc_elem;      /* vertex -> elements holding it, increasing */
};

static long clampl(long v, long lo, long hi) { return v < lo ? lo : (v > hi ? hi : v); }

static void cell_of(const dgx_locator *L, double x, double z, long *ix, long *iz)
{
    *ix = clampl((long)floor((x - L->x0) / L->h), 0, L->nx - 1);
    *iz = clampl((long)floor((z - L->z0) / L->h), 0, L->nz - 1);
}

int dgx_locator_create(long tri, long point_num, const double *point, const long *triangle, dgx_locator **out)
{
    if (tri <= 0 || point_num <= 0 || !point || !triangle || !out) return DGX_EARG;
    dgx_locator *L = (dgx_locator *)calloc(1, sizeof(*L));
    if (!L) return DGX_ENOMEM;
    L->tri = tri; L->point_num = point_num; L->point = point; L->triangle = triangle;
    double xmin = point[0], xmax = point[0], zmin = point[1], zmax = point[1];
    for (long i = 1; i < point_num; i++) {
        xmin = fmin(xmin, point[i * 2]); xmax = fmax(xmax, point[i * 2]);
        zmin = fmin(zmin, point[i * 2 + 1]); zmax = fmax(zmax, point[i * 2 + 1]);
    }
    double area = (xmax - xmin) * (zmax - zmin);
    double h = sqrt((area > 0 ? area : 1.0) / (double)tri) * 2.0;
    if (!(h > 0)) h = 1.0;
    L->h = h;
    L->x0 = xmin - 1e-6 * h;
    L->z0 = zmin - 1e-6 * h;
    L->nx = (long)((xmax - L->x0) / h) + 1;
    L->nz = (long)((zmax - L->z0) / h) + 1;
    long ncell = L->nx * L->nz;
    L->cell_start = (long *)calloc(ncell + 1, sizeof(long));
    L->vcell_start = (long *)calloc(ncell + 1, sizeof(long));
    L->inc_start = (long *)calloc(point_num + 1, sizeof(long));
    if (!L->cell_start || !L->vcell_start || !L->inc_start) { dgx_locator_destroy(L); return DGX_ENOMEM; }
    /* pass 1: counts */
    for (int pass = 0; pass < 2; pass++) {
        long *fill = NULL, *vfill = NULL, *ifill = NULL;
        if (pass == 1) {
            for (long c = 0; c < ncell; c++) {
                L->cell_start[c + 1] += L->cell_start[c];
                L->vcell_start[c + 1] += L->vcell_start[c];
            }
            for (long p = 0; p < point_num; p++) L->inc_start[p + 1] += L->inc_start[p];
            L->cell_elem = (long *)malloc(sizeof(long) * (L->cell_start[ncell] + 1));
            L->vcell_vert = (long *)malloc(sizeof(long) * (L->vcell_start[ncell] + 1));
            L->inc_elem = (long *)malloc(sizeof(long) * (L->inc_start[point_num] + 1));
            fill = (long *)calloc(ncell, sizeof(long));
            vfill = (long *)calloc(ncell, sizeof(long));
            ifill = (long *)calloc(point_num, sizeof(long));
            if (!L->cell_elem || !L->vcell_vert || !L->inc_elem || !fill || !vfill || !ifill) {
                free(fill); free(vfill); free(ifill);
                dgx_locator_destroy(L);
                return DGX_ENOMEM;
            }
        }
        for (long e = 0; e < tri; e++) {
            double bx0 = 1e300, bx1 = -1e300, bz0 = 1e300, bz1 = -1e300;
            for (int k = 0; k < 3; k++) {
                long p = triangle[e * 3 + k] - 1;
                if (p < 0 || p >= point_num) { dgx_set_error("element %ld: vertex id out of range", e + 1); dgx_locator_destroy(L); return DGX_EMESH; }
                bx0 = fmin(bx0, point[p * 2]); bx1 = fmax(bx1, point[p * 2]);
                bz0 = fmin(bz0, point[p * 2 + 1]); bz1 = fmax(bz1, point[p * 2 + 1]);
                if (pass == 0) L->inc_start[p + 1]++;
                else L->inc_elem[L->inc_start[p] + ifill[p]++] = e;
            }
            long ix0, iz0, ix1, iz1;
            cell_of(L, bx0 - 1e-9 * h, bz0 - 1e-9 * h, &ix0, &iz0);
            cell_of(L, bx1 + 1e-9 * h, bz1 + 1e-9 * h, &ix1, &iz1);
            for (long iz = iz0; iz <= iz1; iz++)
                for (long ix = ix0; ix <= ix1; ix++) {
                    long c = iz * L->nx + ix;
                    if (pass == 0) L->cell_start[c + 1]++;
                    else L->cell_elem[L->cell_start[c] + fill[c]++] = e;
                }
        }
        for (long p = 0; p < point_num; p++) {
            long ix, iz;
            cell_of(L, point[p * 2], point[p * 2 + 1], &ix, &iz);
            long c = iz * L->nx + ix;
            if (pass == 0) L->vcell_start[c + 1]++;
            else L->vcell_vert[L->vcell_start[c] + vfill[c]++] = p;
        }
        free(fill); free(vfill); free(ifill);
    }
    *out = L;
    return DGX_OK;
}

void dgx_locator_destroy(dgx_locator *L)
{
    if (!L) return;
    free(L->cell_start); free(L->cell_elem); free(L->vcell_start); free(L->vcell_vert);
    free(L->inc_start); free(L->inc_elem);
    free(L);
}

/* Vertex within tol of (x,z) in both coordinates; first (lowest id) or last (highest id). */
static long match_vertex(const dgx_locator *L, double x, double z, double tol, int last)
{
    long ix, iz, best = -1;
    cell_of(L, x, z, &ix, &iz);
    for (long cz = iz - 1; cz <= iz + 1; cz++)
        for (long cx = ix - 1; cx <= ix + 1; cx++) {
            if (cx < 0 || cz < 0 || cx >= L->nx || cz >= L->nz) continue;
            long c = cz * L->nx + cx;
            for (long k = L->vcell_start[c]; k < L->vcell_start[c + 1]; k++) {
                long p = L->vcell_vert[k];
                if (fabs(L->point[p * 2] - x) < tol && fabs(L->point[p * 2 + 1] - z) < tol)
                    if (best < 0 || (last ? p > best : p < best)) best = p;
            }
        }
    return best;
}

/* Barycentric test with the same formulas and tolerance as the oracle's locator. */
static int inside_elem(const double *point, const long *triangle, long e, double x, double z, double *r, double *s)
{
    double v[3][2];
    element_vertices(point, triangle, e, v);
    double xr = v[1][0] - v[0][0], zr = v[1][1] - v[0][1];
    double xs = v[2][0] - v[0][0], zs = v[2][1] - v[0][1];
    double det = xr * zs - xs * zr;
    double dx = x - v[0][0], dz = z - v[0][1];
    double a = (dx * zs - xs * dz) / det, b = (xr * dz - dx * zr) / det;
    const double tol = 1e-10;
    if (a >= -tol && b >= -tol && a + b <= 1 + tol) {
        *r = 2 * a - 1;
        *s = 2 * b - 1;
        return 1;
    }
    return 0;
}

static int inside(const dgx_locator *L, long e, double x, double z, double *r, double *s)
{
    return inside_elem(L->point, L->triangle, e, x, z, r, s);
}

static long locate_point(const dgx_locator *L, double x, double z, double *r, double *s)
{
    long ix, iz;
    cell_of(L, x, z, &ix, &iz);
    long c = iz * L->nx + ix;
    for (long k = L->cell_start[c]; k < L->cell_start[c + 1]; k++)
        if (inside(L, L->cell_elem[k], x, z, r, s)) return L->cell_elem[k];
    return -1;
}

void dgx_p2_basis(double r, double s, double *phi)
{
    double l0 = -(r + s) / 2, l1 = (r + 1) / 2, l2 = (s + 1) / 2;
    phi[0] = l1 * (2 * l1 - 1); /* node 1: v1 */
    phi[1] = 4 * l1 * l2;       /* node 2: mid v1v2 */
    phi[2] = l2 * (2 * l2 - 1); /* node 3: v2 */
    phi[3] = 4 * l0 * l1;       /* node 4: mid v0v1 */
    phi[4] = 4 * l0 * l2;       /* node 5: mid v0v2 */
    phi[5] = l0 * (2 * l0 - 1); /* node 6: v0 */
}

static const int node_of_vertex[3] = {5, 0, 2}; /* v0 -> node 6, v1 -> node 1, v2 -> node 3 */

static int vertex_slot(const long *tv, long p)
{
    for (int k = 0; k < 3; k++)
        if (tv[k] == p + 1) return k;
    return -1;
}

static double element_jacobian(const double *point, const long *triangle, long e)
{
    double v[3][2];
    element_vertices(point, triangle, e, v);
    double xr = (v[1][0] - v[0][0]) * 0.5, yr = (v[1][1] - v[0][1]) * 0.5;
    double xs = (v[2][0] - v[0][0]) * 0.5, ys = (v[2][1] - v[0][1]) * 0.5;
    return xr * ys - xs * yr; /* fun_DG.c:441-447 */
}

/* Source vector of a located source: element e (0-based) and, on a vertex, that vertex (0-based), else -1
 * and the point (x, z) inside e.  Shared by the host and device locators (dgx_locate_sources_device). */
int dgx_source_at(const double *point, const long *triangle, const dgx_operators *ops, long e, long vertex,
                  double x, double z, int kind, int shot, long wavelet, dgx_source *out)
{
    double phi[6] = {0, 0, 0, 0, 0, 0};
    if (vertex >= 0) {
        int k = vertex_slot(&triangle[e * 3], vertex);
        if (k < 0) { dgx_set_error("locator: element %ld does not hold vertex %ld", e + 1, vertex + 1); return DGX_EMESH; }
        phi[node_of_vertex[k]] = 1.0;
    } else {
        double r, s;
        if (!inside_elem(point, triangle, e, x, z, &r, &s)) {
            dgx_set_error("locator: (%g, %g) is not inside element %ld", x, z, e + 1);
            return DGX_EMESH;
        }
        dgx_p2_basis(r, s, phi);
    }
    double J = element_jacobian(point, triangle, e);
    for (int i = 0; i < 6; i++) {
        double acc = 0;
        for (int k = 0; k < 6; k++) {
            double m = 0; /* (V V^T)_{ik}, summed like DG_leap_Frog_OpenMP.c:227-230 */
            for (int j = 0; j < 6; j++) m = m + ops->V[i * 6 + j] * ops->V[k * 6 + j];
            acc = acc + m * phi[k];
        }
        out->fun[i] = acc / J;
    }
    out->elem = e + 1;
    out->kind = kind;
    out->shot = shot;
    out->wavelet = wavelet;
    return DGX_OK;
}

/* Receiver weights of a located receiver (element e, vertex or -1, point (x, z)); shared like dgx_source_at. */
int dgx_receiver_at(const double *point, const long *triangle, long e, long vertex, double x, double z,
                    dgx_receiver *out)
{
    for (int i = 0; i < 6; i++) out->weight[i] = 0;
    out->elem = e + 1;
    if (vertex >= 0) {
        int k = vertex_slot(&triangle[e * 3], vertex);
        if (k < 0) { dgx_set_error("locator: element %ld does not hold vertex %ld", e + 1, vertex + 1); return DGX_EMESH; }
        out->weight[node_of_vertex[k]] = 1.0;
        return DGX_OK;
    }
    double r, s;
    if (!inside_elem(point, triangle, e, x, z, &r, &s)) {
        dgx_set_error("locator: (%g, %g) is not inside element %ld", x, z, e + 1);
        return DGX_EMESH;
    }
    dgx_p2_basis(r, s, out->weight);
    return DGX_OK;
}

int dgx_locate_source(const dgx_locator *L, const dgx_operators *ops, double x, double z, int kind, int shot,
                      long wavelet, dgx_source *out)
{
    if (!L || !ops || !out) return DGX_EARG;
    long e = -1, vertex = -1;
    long p = match_vertex(L, x, z, 1e-3, 0);
    if (p >= 0 && L->inc_start[p] < L->inc_start[p + 1]) {
        e = L->inc_elem[L->inc_start[p]]; /* first element holding the vertex */
        vertex = p;
    } else {
        double r, s;
        e = locate_point(L, x, z, &r, &s);
        if (e < 0) { dgx_set_error("source (%g, %g) lies outside the mesh", x, z); return DGX_EARG; }
    }
    return dgx_source_at(L->point, L->triangle, ops, e, vertex, x, z, kind, shot, wavelet, out);
}

int dgx_locate_receivers(const dgx_locator *L, long n, const double *xz, dgx_receiver *out)
{
    if (!L || (n > 0 && (!xz || !out))) return DGX_EARG;
    for (long k = 0; k < n; k++) {
        double x = xz[k * 2], z = xz[k * 2 + 1];
        long e = -1, vertex = -1;
        long p = match_vertex(L, x, z, 1e-2, 1);
        if (p >= 0 && L->inc_start[p] < L->inc_start[p + 1]) {
            e = L->inc_elem[L->inc_start[p + 1] - 1]; /* last element holding the vertex */
            vertex = p;
        } else {
            double r, s;
            e = locate_point(L, x, z, &r, &s);
            if (e < 0) { dgx_set_error("receiver %ld at (%g, %g) lies outside the mesh", k, x, z); return DGX_EARG; }
        }
        int rc = dgx_receiver_at(L->point, L->triangle, e, vertex, x, z, &out[k]);
        if (rc) return rc;
    }
    return DGX_OK;
}

/* ---- the reference's per-element operators, for DGX_FP64_REFERENCE ----
 * The reference forms these once per run on the host (Jacobian, Angle, Normals, Elastic_A_B, Central_flux,
 * Dx_Dz; fun_DG.c:421-554 and 1071-1192) and its step loops read them per element.  The
 * expressions and their order below are the reference's, so that a caller's arrays equal its arrays bit for
 * bit on the same machine (same libm for sqrt, pow, acos and cos; no fused multiply-adds). */

/* The three vertices of element e (1-based ids in triangle). */
static void ref_vertices(const double *point, const long *triangle, long e, double P[3][2])
{
    for (int v = 0; v < 3; v++) {
        const long p = triangle[e * 3 + v] - 1;
        P[v][0] = point[p * 2 + 0];
        P[v][1] = point[p * 2 + 1];
    }
}

/* Edge lengths L0 = |v0 v1|, L1 = |v1 v2|, L2 = |v0 v2| (EdgeLength, fun_DG.c:122-146). */
static void ref_lengths(const double P[3][2], double L[3])
{
    L[0] = sqrt(pow(P[0][0] - P[1][0], 2) + pow(P[0][1] - P[1][1], 2));
    L[1] = sqrt(pow(P[1][0] - P[2][0], 2) + pow(P[1][1] - P[2][1], 2));
    L[2] = sqrt(pow(P[0][0] - P[2][0], 2) + pow(P[0][1] - P[2][1], 2));
}

int dgx_reference_operators(long tri, const double *point, const long *triangle, const double *lambda,
                            const double *miu, const double *den, const dgx_operators *ops, double *Dx, double *Dz,
                            double *Q, double *A, double *B)
{
    if (tri <= 0 || !point || !triangle || !lambda || !miu || !den || !ops || !Dx || !Dz || !Q || !A || !B)
        return DGX_EARG;
    const double ux[2] = {1, 0}, uz[2] = {0, 1};
    for (long e = 0; e < tri; e++) {
        double P[3][2], len[3];
        ref_vertices(point, triangle, e, P);
        ref_lengths(P, len);
        /* Jacobian and half edge lengths (Jacobian, fun_DG.c:421-455) */
        const double xr = (P[1][0] - P[0][0]) * 0.5, yr = (P[1][1] - P[0][1]) * 0.5;
        const double xs = (P[2][0] - P[0][0]) * 0.5, ys = (P[2][1] - P[0][1]) * 0.5;
        const double J = xr * ys - xs * yr;
        /* outward unit normals through the in-circle's tangent points (Angle and Normals, fun_DG.c:456-554): the in-centre Ac, the
         * angle at each face's first vertex, the tangent point Bt on the face, and n = (Bt - Ac) / |Bt - Ac| */
        const double L1 = len[0], L2 = len[1], L3 = len[2];
        double ang[3];
        ang[0] = acos((L1 * L1 + L3 * L3 - L2 * L2) / (2 * L1 * L3));
        ang[1] = acos((L1 * L1 + L2 * L2 - L3 * L3) / (2 * L1 * L2));
        ang[2] = acos((L2 * L2 + L3 * L3 - L1 * L1) / (2 * L2 * L3));
        double Ac[2];
        for (int j = 0; j < 2; j++)
            Ac[j] = (len[1] * P[0][j] + len[2] * P[1][j] + len[0] * P[2][j]) / (len[1] + len[2] + len[0]);
        double nrm[3][2];
        for (int k = 0; k < 3; k++) {
            const double *Ps = P[k], *Pe = P[(k + 1) % 3];
            const double d = sqrt(pow(Ac[0] - Ps[0], 2) + pow(Ac[1] - Ps[1], 2)) * cos(0.5 * ang[k]);
            double AB[2];
            for (int j = 0; j < 2; j++) {
                const double Bt = Ps[j] + (d / len[k]) * (Pe[j] - Ps[j]);
                AB[j] = Bt - Ac[j];
            }
            const double n = sqrt(pow(AB[0], 2) + pow(AB[1], 2));
            nrm[k][0] = (AB[0] * ux[0] + AB[1] * ux[1]) / n;
            nrm[k][1] = (AB[0] * uz[0] + AB[1] * uz[1]) / n;
        }
        /* material factors (Elastic_A_B, fun_DG.c:1071-1090) */
        double *a = &A[e * 5], *b = &B[e * 5];
        a[0] = 1 / den[e];
        a[1] = 1 / den[e];
        a[2] = lambda[e] + 2 * miu[e];
        a[3] = lambda[e];
        a[4] = miu[e];
        b[0] = 1 / den[e];
        b[1] = 1 / den[e];
        b[2] = lambda[e];
        b[3] = lambda[e] + 2 * miu[e];
        b[4] = miu[e];
        /* face flux coefficients 0.5 n (A|B) (L/2) / J (Central_flux, fun_DG.c:1091-1136) */
        for (int j = 0; j < 3; j++) {
            const double nx = nrm[j][0], nz = nrm[j][1], je = len[j] / 2;
            double *q = &Q[e * 30 + j * 10];
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
        /* Dx = (rx Vr + sx Vs) inverV, Dz = (rz Vr + sz Vs) inverV (Dx_Dz, fun_DG.c:1137-1192) */
        const double Rx = ys / J, Rz = -xs / J, Sx = -yr / J, Sz = xr / J;
        double gx[6][6], gz[6][6];
        for (int i = 0; i < 6; i++)
            for (int n = 0; n < 6; n++) {
                gx[i][n] = Rx * ops->Vr[i * 6 + n] + Sx * ops->Vs[i * 6 + n];
                gz[i][n] = Rz * ops->Vr[i * 6 + n] + Sz * ops->Vs[i * 6 + n];
            }
        for (int p = 0; p < 6; p++)
            for (int q = 0; q < 6; q++) {
                double f1 = 0, f2 = 0;
                for (int k = 0; k < 6; k++) {
                    f1 = f1 + gx[p][k] * ops->inverV[k * 6 + q];
                    f2 = f2 + gz[p][k] * ops->inverV[k * 6 + q];
                }
                Dx[e * 36 + p * 6 + q] = f1;
                Dz[e * 36 + p * 6 + q] = f2;
            }
    }
    return DGX_OK;
}
