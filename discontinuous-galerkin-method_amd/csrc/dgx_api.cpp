/*
 * dgx_api.cpp -- host side of the solver C-ABI (include/dgx.h): device-resident mesh and fields,
 * launch sequencing and the step graph.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "dgx.h"
#include "dgx_kernels.h"
#include "dgx_p2ops.h"
#include "dgx_setup.h"

extern "C" {
hipError_t dgx_launch_f64(const void *args, int nshots, hipStream_t st, int which);
hipError_t dgx_launch_f32(const void *args, int nshots, hipStream_t st, int which);
hipError_t dgx_launch_source(const void *args, int fp32, int nsrc, double w, hipStream_t st);
hipError_t dgx_launch_record(const void *args, int fp32, int t, int nshots, int nrec, hipStream_t st);
hipError_t dgx_launch_advance(int *step_base, int n, hipStream_t st);
hipError_t dgx_noalloc_scratch(double **ptr);
hipError_t dgx_launch_unpermute(const void *f, const void *lo, int fp32, const int *perm, double *out, long tri, long spad,
                                int il, long shot, hipStream_t st);
hipError_t dgx_launch_copy(const void *src, void *dst, long long n32, int variant, hipStream_t st);
hipError_t dgx_launch_batched_f64(const void *args, hipStream_t st, int which);
int dgx_batched_shape_ok(int shot_tile, int wg_waves);
hipError_t dgx_launch_batched_f32(const void *args, hipStream_t st, int which);
hipError_t dgx_launch_reference(const void *args, const double *ops, const double *mats, int which, int nshots,
                                int nsrc, int nrec, int t, double w, hipStream_t st);
}

/* ---- error reporting ---- */
static thread_local char g_err[1024] = "";

extern "C" void dgx_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char *dgx_last_error(void) { return g_err; }

#define HIP_TRY(call)                                                                                          \
    do {                                                                                                       \
        hipError_t e_ = (call);                                                                                \
        if (e_ != hipSuccess) {                                                                                \
            dgx_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__);          \
            return DGX_EHIP;                                                                                   \
        }                                                                                                      \
    } while (0)

namespace {

enum { K_VEL_FUSED = 0, K_STRESS_FUSED = 1, K_VEL = 2, K_STRESS = 3, K_ABSORB = 4, K_FULL = 5, K_BLOCK = 6 };
const int kGraphSteps = 32;    /* steps captured per graph replay (half-step and full-step kernels) */
const int kGraphLaunches = 16; /* launches captured per graph replay of the blocked kernel (even: the graph
                                * ends on the field copy it starts from) */

template <typename T>
hipError_t upload(T **dst, const std::vector<T> &src)
{
    *dst = nullptr;
    if (src.empty()) return hipSuccess;
    hipError_t e = hipMalloc((void **)dst, sizeof(T) * src.size());
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice);
}

/* Hilbert-curve index of a point in [0,1)^2 on a 2^16 x 2^16 grid.  Consecutive indices are always
 * adjacent cells (unlike the Z-order curve, which jumps at every power-of-two boundary), so elements
 * that are close in this order are close in space. */
uint32_t hilbert(double u, double v)
{
    const uint32_t n = 1u << 16;
    uint32_t x = (uint32_t)std::min(65535.0, std::max(0.0, u * 65536.0));
    uint32_t y = (uint32_t)std::min(65535.0, std::max(0.0, v * 65536.0));
    uint32_t d = 0;
    for (uint32_t s = n / 2; s > 0; s /= 2) {
        uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += s * s * ((3u * rx) ^ ry);
        if (ry == 0) { /* rotate the quadrant */
            if (rx == 1) {
                x = s - 1 - (x & (s - 1)) + (x & ~(s - 1));
                y = s - 1 - (y & (s - 1)) + (y & ~(s - 1));
                x &= (s - 1);
                y &= (s - 1);
            } else {
                x &= (s - 1);
                y &= (s - 1);
            }
            std::swap(x, y);
        } else {
            x &= (s - 1);
            y &= (s - 1);
        }
    }
    return d;
}

inline long long fidx(long long q, int i) { return ((q >> 6) * 6 + i) * 64 + (q & 63); }

} // namespace

#include "dgx_internal.h"

struct dgx_solver {
    int fp32 = 0;
    int comp = 0;                    /* compensated FP32: a low word per value (blocked kernel) */
    int exact = 0;                   /* DGX_FP64_REFERENCE: the reference's arrays and order (dgx_reference.hip) */
    int layout = DGX_LAYOUT_ELEMENT;
    int shot_tile = 0;               /* shot-innermost: shots per lane group */
    int wg_waves = 4;                /* shot-innermost: wavefronts per workgroup */
    int il = 0;                      /* shot-innermost: log2 of the elements interleaved per node row */
    long spad = 1;                   /* shot-innermost: padded shot count */
    std::vector<int> tune_tiles, tune_waves; /* autotuner candidates and their ms per step */
    std::vector<double> tune_ms;
    long tri = 0, tri_pad = 0, nshots = 1, nsteps = 0, nrec = 0, nsrc = 0;
    long step = 0;          /* steps completed */
    long base_host = 0;     /* mirror of the device step counter */
    int nper = 0;
    std::vector<long> perm; /* device element -> caller element (0-based) */
    std::vector<long> inv;  /* caller element -> device element */
    int device = 0;         /* the HIP device current at dgx_create */
    int *d_perm = nullptr;  /* device element -> caller element, on the device */
    hipStream_t stream = nullptr;
    hipGraphExec_t graph[2] = {nullptr, nullptr}; /* step graphs, by which field copy holds the state */
    int full = 0;           /* one full-step kernel per step (element layout) */
    int blocked = 0;        /* blocked kernel: bkmax steps per launch (element layout) */
    int bkmax = 0, bthreads = 0;
    long bown = 0, blanes = 0; /* own elements and lanes over all tiles of the blocked kernel */
    int *blevels = nullptr;
    int4 *blrec = nullptr;
    int4 *blnb = nullptr;
    int cur = 0;            /* full step: copy of the fields holding the current state */
    long vrows = 0;         /* element layout: rows per field and shot (elements, then ring rows) */
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    /* device memory */
    void *fields = nullptr;  /* 5 * nshots * vrows*6 reals (element layout) */
    void *fields2 = nullptr; /* full step: the second copy */
    int4 *nb2 = nullptr, *tiles = nullptr;
    int *ring = nullptr;
    long ntile = 0;
    int4 *nbr = nullptr, *vtx = nullptr;
    double2 *pt = nullptr;
    void *mat = nullptr;    /* rho_inv | lam | mu, tri_pad reals each */
    void *absorb = nullptr; /* tri_pad*6 reals */
    dgx_aux *aux = nullptr;
    dgx_src_dev *src = nullptr;
    dgx_rec_dev *rec = nullptr;
    double *wav = nullptr;
    double *trace = nullptr;
    int *step_base = nullptr;
    double *refops = nullptr; /* DGX_FP64_REFERENCE: Q | 1/rho | lambda + 2 mu | lambda | mu per element, blocked by 64 */
    double refmats[108] = {};  /* DGX_FP64_REFERENCE: Vr | Vs | inverV */
    /* argument blocks */
    dgx::StepArgs<double> a64;
    dgx::StepArgs<float> a32;
    std::vector<dgx_src_dev> src_host;
    /* snapshots (allocated at the first one): the fields of one shot in the caller's element order as
     * doubles on the device, their pinned host copy, and the stream that moves them while stepping goes on */
    double *d_snap = nullptr, *h_snap = nullptr;
    hipStream_t copy_stream = nullptr;
    hipEvent_t snap_ready = nullptr, snap_done = nullptr;
    int snap_mask = 0; /* fields of the snapshot in flight; 0 = none */

    const void *args() const { return fp32 ? (const void *)&a32 : (const void *)&a64; }
    size_t real_size() const { return fp32 ? 4 : 8; }
    /* the copy holding the current state */
    void *state() const { return cur ? fields2 : fields; }
};

/* Full step: the copy just written becomes the current state. */
static void swap_copies(dgx_solver *s)
{
    s->cur ^= 1;
    for (int c = 0; c < 5; c++) {
        std::swap(s->a64.f[c], s->a64.g[c]);
        std::swap(s->a32.f[c], s->a32.g[c]);
        std::swap(s->a32.fl[c], s->a32.gl[c]);
    }
}

/*
 * Tiles of the full-step kernel: runs of consecutive elements (locality order) grown while the run plus
 * its ring (elements outside the run that share a face with it) fit one element per thread of a
 * workgroup.  tiles[t] = {first, own, first ring entry, ring entries}; ring = element ids, tile after
 * tile; nb2[q] = {row of face j's neighbour velocities in the output copy, code}: the neighbour itself
 * inside the tile, row vrow0 + its ring entry outside it, -1 on the boundary.
 */
static void build_tiles(long tri, const std::vector<int4> &nbr, long vrow0, std::vector<int4> &tiles,
                        std::vector<int> &ring, std::vector<int4> &nb2)
{
    std::vector<int> stamp(tri, -1), pos(tri, -1);
    nb2.assign(tri, int4{-1, -1, -1, 0});
    long q0 = 0;
    int t = 0;
    while (q0 < tri) {
        long q = q0;
        long nring = 0;
        std::vector<int> cand;
        while (q < tri && q - q0 < DGX_BLOCK) {
            const int nq[3] = {nbr[q].x, nbr[q].y, nbr[q].z};
            long add = 0;
            for (int j = 0; j < 3; j++) {
                const int e = nq[j];
                if (e >= 0 && !(e >= q0 && e <= q) && stamp[e] != t) add++;
            }
            const long in_ring = stamp[q] == t ? 1 : 0;
            if ((q - q0 + 1) + nring + add - in_ring > DGX_BLOCK) break;
            if (in_ring) { stamp[q] = -1; nring--; }
            for (int j = 0; j < 3; j++) {
                const int e = nq[j];
                if (e >= 0 && !(e >= q0 && e <= q) && stamp[e] != t) {
                    stamp[e] = t;
                    cand.push_back(e);
                    nring++;
                }
            }
            q++;
        }
        const int rbegin = (int)ring.size();
        for (int e : cand)
            if (stamp[e] == t) {
                pos[e] = (int)ring.size();
                ring.push_back(e);
            }
        tiles.push_back(int4{(int)q0, (int)(q - q0), rbegin, (int)ring.size() - rbegin});
        for (long x = q0; x < q; x++) {
            const int nx[3] = {nbr[x].x, nbr[x].y, nbr[x].z};
            int r[3];
            for (int j = 0; j < 3; j++) {
                const int e = nx[j];
                r[j] = e < 0 ? -1 : (e >= q0 && e < q) ? e : (int)(vrow0 + pos[e]);
            }
            nb2[x] = int4{r[0], r[1], r[2], nbr[x].w};
        }
        q0 = q;
        t++;
    }
}

/*
 * Tiles of the blocked kernel: runs of consecutive elements (locality order), each grown by bisection on its
 * length to the longest run whose levels fit the workgroup: levels 0 .. 2K-1 (the run and every element within
 * 2K-1 face hops of it, the elements the kernel updates) at most `cap` lanes, one per thread (`cap` is the
 * workgroup's thread count), and level 2K (ghosts, which only supply their stresses to the first velocity update
 * of level 2K-1) at most cap/4 - 1 elements, held in LDS slots past the threads' (the last slot is the kernel's
 * zero slot).  tiles[t] = {first, own, first record, records}; levels[t*DGX_BLOCK_LEVELS + d] = end record of
 * level d, entries past level 2K equal to the record count.  Records: the run's elements in order, then each
 * level in increasing element id; gid = the element, lnb = {LDS slot of face j's neighbour, code}: the lane for
 * levels below 2K, cap + (record - lanes) for a ghost, -1 on the boundary (ghost records: -1 everywhere, never
 * read).  Every face neighbour of a level below 2K is in the tile.  Returns false if one element and its levels
 * do not fit.
 */
bool dgx_build_block_tiles(long tri, const std::vector<int4> &nbr, int K, int cap, std::vector<int4> &tiles,
                           std::vector<int> &levels, std::vector<int> &gid, std::vector<int4> &lnb)
{
    const int depth = 2 * K;       /* the ghost level */
    const long gcap = cap / 4 - 1; /* ghost slots (the last of the cap/4 slots is the zero slot) */
    std::vector<int> stamp(tri, 0), slot(tri, -1);
    int cur = 0;
    std::vector<int> front, next;
    std::vector<std::vector<int>> lv(depth + 1);
    /* whether the run [q0, q0 + n) and its levels fit; *total = its records (lanes + ghosts) */
    auto fits = [&](long q0, long n, bool keep, long *total) -> bool {
        cur++;
        long lanes = n;
        front.clear();
        for (long q = q0; q < q0 + n; q++) {
            stamp[q] = cur;
            front.push_back((int)q);
        }
        for (int d = 1; d <= depth; d++) {
            next.clear();
            for (int e : front) {
                const int nq[3] = {nbr[e].x, nbr[e].y, nbr[e].z};
                for (int x : nq)
                    if (x >= 0 && stamp[x] != cur) {
                        stamp[x] = cur;
                        next.push_back(x);
                    }
            }
            if (d < depth) {
                lanes += (long)next.size();
                if (lanes > cap) return false;
            } else if ((long)next.size() > gcap) {
                return false;
            }
            if (keep) {
                lv[d] = next;
                std::sort(lv[d].begin(), lv[d].end());
            }
            front.swap(next);
        }
        if (total) *total = lanes + (long)front.size();
        return true;
    };
    long q0 = 0, prev = 1;
    while (q0 < tri) {
        long lo = 1, hi = std::min<long>(cap, tri - q0);
        if (!fits(q0, 1, false, nullptr)) return false;
        if (prev <= hi) { /* tiles along the order have similar sizes: start from the previous one */
            if (fits(q0, prev, false, nullptr)) lo = prev;
            else hi = prev - 1;
        }
        while (lo < hi) {
            const long mid = (lo + hi + 1) / 2;
            if (fits(q0, mid, false, nullptr)) lo = mid;
            else hi = mid - 1;
        }
        const long n = lo;
        long total = 0;
        fits(q0, n, true, &total);
        const int base = (int)gid.size();
        for (long q = q0; q < q0 + n; q++) gid.push_back((int)q);
        int lev[DGX_BLOCK_LEVELS];
        lev[0] = (int)n;
        for (int d = 1; d <= depth; d++) {
            gid.insert(gid.end(), lv[d].begin(), lv[d].end());
            lev[d] = lev[d - 1] + (int)lv[d].size();
        }
        for (int d = depth + 1; d < DGX_BLOCK_LEVELS; d++) lev[d] = (int)total;
        levels.insert(levels.end(), lev, lev + DGX_BLOCK_LEVELS);
        const int lanes = lev[depth - 1];
        for (int l = 0; l < (int)total; l++) slot[gid[base + l]] = l < lanes ? l : cap + (l - lanes);
        for (int l = 0; l < (int)total; l++) {
            const int e = gid[base + l];
            int r[3] = {-1, -1, -1};
            if (l < lanes) {
                const int nq[3] = {nbr[e].x, nbr[e].y, nbr[e].z};
                for (int j = 0; j < 3; j++) r[j] = nq[j] < 0 ? -1 : slot[nq[j]];
            }
            lnb.push_back(int4{r[0], r[1], r[2], nbr[e].w});
        }
        tiles.push_back(int4{(int)q0, (int)n, base, (int)total});
        q0 += n;
        prev = n;
    }
    return true;
}

/* DGX_FP64_REFERENCE: one of its kernels (dgx_launch_reference's `which`).  The split source kernel (2, 3) reads
 * the caller-ordered source list, the fused step (6, 7) the aux-ordered one through the elements' aux records. */
static int launch_exact(dgx_solver *s, int which, int step_off, int t = -1, double w = 0)
{
    dgx::StepArgs<double> a = s->a64;
    a.step_off = step_off;
    a.src = (which == 2 || which == 3) ? s->src + s->nsrc : s->src;
    const hipError_t e = dgx_launch_reference(&a, s->refops, s->refmats, which, (int)s->nshots, (int)s->nsrc,
                                              (int)s->nrec, t, w, s->stream);
    if (e != hipSuccess) {
        dgx_set_error("kernel launch failed: %s", hipGetErrorString(e));
        return DGX_EHIP;
    }
    return DGX_OK;
}

static int launch(dgx_solver *s, int which, int step_off)
{
    hipError_t e;
    if (s->exact && which != K_ABSORB) {
        if (which == K_VEL || which == K_STRESS) return launch_exact(s, which == K_VEL ? 0 : 1, step_off);
        if (which == K_VEL_FUSED || which == K_STRESS_FUSED) return launch_exact(s, which == K_VEL_FUSED ? 6 : 7, step_off);
        dgx_set_error("DGX_FP64_REFERENCE runs the split step kernels only");
        return DGX_ESTATE;
    }
    if (s->layout == DGX_LAYOUT_SHOT_INNER) {
        if (which != K_VEL_FUSED && which != K_STRESS_FUSED) {
            dgx_set_error("the split entry points and the full-step kernel need the element layout (DGX_LAYOUT_ELEMENT)");
            return DGX_ESTATE;
        }
        s->a32.step_off = s->a64.step_off = step_off;
        e = s->fp32 ? dgx_launch_batched_f32(&s->a32, s->stream, which == K_VEL_FUSED ? 0 : 1)
                    : dgx_launch_batched_f64(&s->a64, s->stream, which == K_VEL_FUSED ? 0 : 1);
        if (e != hipSuccess) {
            dgx_set_error("kernel launch failed: %s", hipGetErrorString(e));
            return DGX_EHIP;
        }
        return DGX_OK;
    }
    if (which == K_FULL && !s->full) {
        dgx_set_error("the full-step kernel is not enabled for this solver (dgx_problem.step_kernel)");
        return DGX_ESTATE;
    }
    if (which == K_BLOCK && !s->blocked) {
        dgx_set_error("the blocked kernel is not enabled for this solver (dgx_problem.step_kernel)");
        return DGX_ESTATE;
    }
    if (s->fp32) {
        s->a32.step_off = step_off;
        e = dgx_launch_f32(&s->a32, (int)s->nshots, s->stream, which);
    } else {
        s->a64.step_off = step_off;
        e = dgx_launch_f64(&s->a64, (int)s->nshots, s->stream, which);
    }
    if (e != hipSuccess) {
        dgx_set_error("kernel launch failed: %s", hipGetErrorString(e));
        return DGX_EHIP;
    }
    return DGX_OK;
}

template <typename R>
static void fill_args(dgx_solver *s, dgx::StepArgs<R> &a, const dgx_operators *ops, double dt)
{
    /* Dr = Vr*inverV, Ds = Vs*inverV in FP64 from the file values (fun_DG.c:1176-1190 forms the same
     * products after scaling by the metric terms) */
    for (int i = 0; i < 6; i++)
        for (int n = 0; n < 6; n++) {
            double dr = 0, ds = 0;
            for (int k = 0; k < 6; k++) {
                dr += ops->Vr[i * 6 + k] * ops->inverV[k * 6 + n];
                ds += ops->Vs[i * 6 + k] * ops->inverV[k * 6 + n];
            }
            a.op.Dr[i * 6 + n] = (R)dr;
            a.op.Ds[i * 6 + n] = (R)ds;
        }
    const double *S[3] = {ops->S1, ops->S2, ops->S3};
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 18; k++) a.op.L[j * 18 + k] = (R)S[j][k];
    /* The velocity kernels have a variant with the P2 operators of the shipped operator files compiled
     * in; it serves this problem when its operators equal that table bit for bit.  The results agree
     * with the argument variant to rounding (with constants the compiler may fuse a different product
     * into an FMA).  DGX_GENERIC_OPERATORS=1 forces the variant that reads them from the argument. */
    static_assert(sizeof(a.op) == 126 * sizeof(R), "Ops<R> is Dr | Ds | L, packed");
    R table[126];
    for (int k = 0; k < 126; k++) table[k] = (R)dgx_p2_ops[k];
    const char *generic = getenv("DGX_GENERIC_OPERATORS");
    a.p2 = (memcmp(table, &a.op, sizeof table) == 0 && !(generic && strcmp(generic, "1") == 0)) ? 1 : 0;
    R *f = (R *)s->fields, *f2 = (R *)s->fields2;
    size_t per_field = (s->layout == DGX_LAYOUT_SHOT_INNER) ? (size_t)s->tri_pad * 6 * s->spad
                                                             : (size_t)s->nshots * s->vrows * 6;
    for (int c = 0; c < 5; c++) {
        a.f[c] = f + c * per_field;
        a.g[c] = f2 ? f2 + c * per_field : nullptr;
        /* compensated FP32: the low words follow the five fields in each copy */
        a.fl[c] = s->comp ? f + (5 + c) * per_field : nullptr;
        a.gl[c] = s->comp ? f2 + (5 + c) * per_field : nullptr;
    }
    a.fstride = s->vrows;
    a.nb2 = s->nb2;
    a.blevels = s->blevels;
    a.blrec = s->blrec;
    a.blnb = s->blnb;
    a.bkmax = s->bkmax;
    a.bthreads = s->bthreads;
    a.bsteps = s->bkmax;
    a.tiles = s->tiles;
    a.ring = s->ring;
    a.ntile = (int)s->ntile;
    a.tper = (int)((s->ntile + DGX_NXCD - 1) / DGX_NXCD);
    a.vrow0 = (int)s->tri_pad;
    a.nbr = s->nbr;
    a.vtx = s->vtx;
    a.pt = s->pt;
    R *m = (R *)s->mat;
    a.rho_inv = m;
    a.lam = m + s->tri_pad;
    a.mu = m + 2 * s->tri_pad;
    a.absorb = (const R *)s->absorb;
    a.aux = s->aux;
    a.src = s->src;
    a.rec = s->rec;
    a.wav = s->wav;
    a.trace = s->trace;
    a.step_base = s->step_base;
    a.step_off = 0;
    a.nsteps = (int)s->nsteps;
    a.nrec = (int)s->nrec;
    a.tri = (int)s->tri;
    a.nper = s->nper;
    a.dt = dt;
    a.spad = (int)s->spad;
    a.nshots = (int)s->nshots;
    a.shot_tile = s->shot_tile;
    a.wg_waves = s->wg_waves;
    a.il = s->il;
    a.shot_start = nullptr;
    a.shot_len = nullptr;
}

/* Blocks per XCD slice for the current layout and shot tile. */
static int blocks_per_xcd(const dgx_solver *s)
{
    long per_block = (s->layout == DGX_LAYOUT_SHOT_INNER) ? (long)s->wg_waves * (64 / s->shot_tile) : DGX_BLOCK;
    long nblk = (s->tri + per_block - 1) / per_block;
    return (int)((nblk + DGX_NXCD - 1) / DGX_NXCD);
}

static void free_solver(dgx_solver *s)
{
    if (!s) return;
    /* teardown: errors here cannot be acted on */
    for (auto &g : s->graph)
        if (g) (void)hipGraphExecDestroy(g);
    if (s->copy_stream) (void)hipStreamSynchronize(s->copy_stream);
    for (auto &e : s->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {s->snap_ready, s->snap_done})
        if (e) (void)hipEventDestroy(e);
    void *ptrs[] = {s->fields, s->fields2, s->nb2, s->tiles, s->ring, s->nbr, s->vtx, s->pt, s->mat, s->absorb,
                    s->aux, s->src, s->rec, s->wav, s->trace, s->step_base, s->d_perm, s->d_snap,
                    s->blevels, s->blrec, s->blnb, s->refops};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (s->h_snap) (void)hipHostFree(s->h_snap);
    if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

static void set_tile(dgx_solver *s, int tile, int waves)
{
    s->shot_tile = tile;
    s->wg_waves = waves;
    s->nper = blocks_per_xcd(s);
    s->a64.shot_tile = s->a32.shot_tile = tile;
    s->a64.wg_waves = s->a32.wg_waves = waves;
    s->a64.nper = s->a32.nper = s->nper;
    /* a tile that holds every (padded) shot of the row interleaves the 64/tile elements of a wavefront in
     * each node row, so that a wavefront's node loads are one contiguous span */
    s->il = (s->layout == DGX_LAYOUT_SHOT_INNER && tile == s->spad && tile < 64) ? __builtin_ctz(64 / tile) : 0;
    s->a64.il = s->a32.il = s->il;
}

/* Shot-innermost autotuner: time a few real steps with every tile shape the kernels are built for whose
 * shot tile divides the padded shot count (shots per lane group, and wavefronts per workgroup, which
 * with it fix the elements per workgroup), keep the fastest, then restore the initial state (zero
 * fields and traces, step 0).  `fixed_waves` > 0 tunes the shot tile alone. */
static int autotune(dgx_solver *s, int fixed_waves)
{
    const int cands[] = {4, 8, 16, 32, 64};
    const int nt = (int)std::min<long>(3, s->nsteps);
    int best = 0, best_waves = 4;
    double best_ms = 1e300;
    for (int tile : cands)
    for (int waves : {2, 4, 8}) {
        if (s->spad % tile || (fixed_waves && waves != fixed_waves) || !dgx_batched_shape_ok(tile, waves)) continue;
        set_tile(s, tile, waves);
        for (int rep = 0; rep < 2; rep++) { /* first pass warms up */
            HIP_TRY(hipEventRecord(s->ev[0], s->stream));
            for (int k = 0; k < nt; k++) {
                int rc = launch(s, K_VEL_FUSED, k);
                if (rc == DGX_OK) rc = launch(s, K_STRESS_FUSED, k);
                if (rc) return rc;
            }
            HIP_TRY(hipEventRecord(s->ev[1], s->stream));
            HIP_TRY(hipEventSynchronize(s->ev[1]));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
            if (rep == 1) {
                s->tune_tiles.push_back(tile);
                s->tune_waves.push_back(waves);
                s->tune_ms.push_back(ms / nt);
                if (ms < best_ms) { best_ms = ms; best = tile; best_waves = waves; }
            }
        }
    }
    if (!best) {
        dgx_set_error("dgx_create: no shot tile with %d wavefronts per workgroup divides the %ld padded shots",
                      fixed_waves, s->spad);
        return DGX_EARG;
    }
    set_tile(s, best, best_waves);
    size_t rs = s->real_size();
    HIP_TRY(hipMemsetAsync(s->fields, 0, (size_t)5 * s->tri_pad * 6 * s->spad * rs, s->stream));
    if (s->trace)
        HIP_TRY(hipMemsetAsync(s->trace, 0, (size_t)s->nshots * s->nsteps * 2 * s->nrec * sizeof(double), s->stream));
    HIP_TRY(hipMemsetAsync(s->step_base, 0, sizeof(int), s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return DGX_OK;
}

/*
 * Host-side mesh build shared by the solver and the server: validation (vertex ids, counter-clockwise
 * orientation, neighbour consistency, density), locality order (Hilbert index of element centroids),
 * face codes from the .neigh columns, vertices renumbered along that order, per-element material and
 * sponge factors.  `shot_inner` selects the element-major sponge layout of the shot-innermost kernels
 * (otherwise the 64-element blocked layout).  `explosive` (caller numbering, may be NULL) marks
 * elements whose neighbours must condition their face values.
 */
int dgx_build_mesh_host(const dgx_problem *p, int shot_inner, long tri_pad, const std::vector<char> *explosive,
                        dgx_mesh_host *out)
{
    const long T = p->tri;
    const double *P = p->point;
    const long *Tv = p->triangle;
    out->tri = T;
    out->tri_pad = tri_pad;
    double xmin = P[0], xmax = P[0], zmin = P[1], zmax = P[1];
    for (long i = 1; i < p->point_num; i++) {
        xmin = std::min(xmin, P[2 * i]); xmax = std::max(xmax, P[2 * i]);
        zmin = std::min(zmin, P[2 * i + 1]); zmax = std::max(zmax, P[2 * i + 1]);
    }
    std::vector<uint64_t> key(T);
    for (long e = 0; e < T; e++) {
        for (int k = 0; k < 3; k++)
            if (Tv[3 * e + k] < 1 || Tv[3 * e + k] > p->point_num) {
                dgx_set_error("element %ld: vertex id %ld out of range", e + 1, Tv[3 * e + k]);
                return DGX_EMESH;
            }
        const double *a = &P[2 * (Tv[3 * e] - 1)], *b = &P[2 * (Tv[3 * e + 1] - 1)], *c = &P[2 * (Tv[3 * e + 2] - 1)];
        double D = (b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1]);
        if (!(D > 0)) {
            dgx_set_error("element %ld is not counter-clockwise (2*area = %g)", e + 1, D);
            return DGX_EMESH;
        }
        double cx = (a[0] + b[0] + c[0]) / 3, cz = (a[1] + b[1] + c[1]) / 3;
        double ext = std::max(xmax - xmin, zmax - zmin);
        if (!(ext > 0)) ext = 1;
        key[e] = ((uint64_t)hilbert((cx - xmin) / ext, (cz - zmin) / ext) << 32) | (uint64_t)e;
    }
    std::sort(key.begin(), key.end());
    out->perm.resize(T);
    out->inv.resize(T);
    for (long q = 0; q < T; q++) {
        out->perm[q] = (long)(key[q] & 0xffffffffu);
        out->inv[out->perm[q]] = q;
    }
    std::vector<char> sponge(T, 0);
    if (p->absorb_fun)
        for (long e = 0; e < T; e++)
            for (int i = 0; i < 6; i++)
                if (p->absorb_fun[6 * e + i] != 1.0) sponge[e] = 1;

    /* face j of element e runs v_j -> v_{j+1} (j = 0, 1) or v0 -> v2 (j = 2); Triangle's .neigh column c
     * holds the element opposite vertex c (fun_DG.c:576-578) */
    static const int col_of_face[3] = {2, 0, 1};
    out->nbr.assign(tri_pad, make_int4(-1, -1, -1, 0));
    out->vtx.assign(tri_pad, make_int4(0, 0, 0, 0));
    for (long q = 0; q < T; q++) {
        long e = out->perm[q];
        const long *v = &Tv[3 * e];
        int nbq[3] = {-1, -1, -1};
        unsigned code = 0;
        for (int j = 0; j < 3; j++) {
            long nb = p->neighbor[3 * e + col_of_face[j]];
            if (nb == -1 || nb == 0) continue;
            if (nb < 1 || nb > T) {
                dgx_set_error("element %ld: neighbour %ld out of range", e + 1, nb);
                return DGX_EMESH;
            }
            const long *w = &Tv[3 * (nb - 1)];
            long fa = (j == 2) ? v[0] : v[j], fb = (j == 2) ? v[2] : v[j + 1];
            int m = -1;
            const long nb_face[3][2] = {{w[0], w[1]}, {w[1], w[2]}, {w[0], w[2]}};
            for (int t = 0; t < 3; t++)
                if ((nb_face[t][0] == fa && nb_face[t][1] == fb) || (nb_face[t][0] == fb && nb_face[t][1] == fa)) m = t;
            if (m < 0) {
                dgx_set_error("element %ld face %d: neighbour %ld does not share that edge", e + 1, j, nb);
                return DGX_EMESH;
            }
            unsigned rev = (nb_face[m][0] != fa) ? 1u : 0u; /* start vertices differ (fun_DG.c:1225,1251) */
            nbq[j] = (int)out->inv[nb - 1];
            code |= ((unsigned)m | (rev << 2)) << (3 * j);
            if (sponge[nb - 1] || (explosive && (*explosive)[nb - 1])) code |= DGX_CODE_NBSPECIAL(j);
        }
        if (sponge[e]) code |= DGX_CODE_SPONGE;
        out->nbr[q] = make_int4(nbq[0], nbq[1], nbq[2], (int)code);
        out->vtx[q] = make_int4((int)(v[0] - 1), (int)(v[1] - 1), (int)(v[2] - 1), 0);
    }
    /* vertices renumbered in order of first use along the element order, so the three vertex loads
     * of neighbouring elements hit the same cache lines */
    std::vector<int> vnew(p->point_num, -1);
    out->pt.clear();
    out->pt.reserve(p->point_num);
    for (long q = 0; q < T; q++)
        for (int k = 0; k < 3; k++) {
            int &slot = k == 0 ? out->vtx[q].x : k == 1 ? out->vtx[q].y : out->vtx[q].z;
            int v = slot;
            if (vnew[v] < 0) {
                vnew[v] = (int)out->pt.size();
                out->pt.push_back(make_double2(P[2 * v], P[2 * v + 1]));
            }
            slot = vnew[v];
        }
    /* materials (Elastic_A_B, fun_DG.c:1071-1088): 1/rho, lambda, mu per element */
    out->mat.assign(3 * tri_pad, 0.0);
    out->absorb.assign(6 * tri_pad, 1.0);
    for (long q = 0; q < T; q++) {
        long e = out->perm[q];
        if (!(p->den[e] > 0)) { dgx_set_error("element %ld: density must be positive", e + 1); return DGX_EARG; }
        out->mat[q] = 1 / p->den[e];
        out->mat[tri_pad + q] = p->lambda[e];
        out->mat[2 * tri_pad + q] = p->miu[e];
        if (p->absorb_fun)
            for (int i = 0; i < 6; i++) {
                long long o = shot_inner ? (long long)q * 6 + i : fidx(q, i);
                out->absorb[o] = p->absorb_fun[6 * e + i];
            }
    }
    return DGX_OK;
}

/* HIP_TRY for dgx_create: a failure frees the partly built solver */
#define CREATE_TRY(call)                                                                                       \
    do {                                                                                                       \
        hipError_t e_ = (call);                                                                                \
        if (e_ != hipSuccess) {                                                                                \
            dgx_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__);          \
            free_solver(s);                                                                                    \
            return e_ == hipErrorOutOfMemory ? DGX_ENOMEM : DGX_EHIP;                                          \
        }                                                                                                      \
    } while (0)

extern "C" int dgx_create(const dgx_problem *p, const dgx_operators *ops, dgx_solver **out)
{
    if (!p || !ops || !out) { dgx_set_error("dgx_create: null argument"); return DGX_EARG; }
    *out = nullptr;
    const long T = p->tri;
    if (T <= 0 || p->point_num <= 0 || !p->point || !p->triangle || !p->neighbor || !p->lambda || !p->miu ||
        !p->den) {
        dgx_set_error("dgx_create: empty mesh or missing arrays");
        return DGX_EARG;
    }
    if (T > (1L << 28) || p->point_num > (1L << 30)) { dgx_set_error("dgx_create: mesh too large"); return DGX_EARG; }
    if (p->nshots < 1 || p->nsteps < 0 || p->nrec < 0 || p->nsrc < 0 || (p->nsrc > 0 && !p->src) ||
        (p->nrec > 0 && !p->rec) || (p->nsrc > 0 && (!p->wavelet || p->nwavelets < 1)) ||
        (p->precision != DGX_FP64 && p->precision != DGX_FP32 && p->precision != DGX_FP32_COMPENSATED &&
         p->precision != DGX_FP64_REFERENCE) ||
        !(p->dt > 0)) {
        dgx_set_error("dgx_create: invalid shots/steps/sources/receivers/precision/dt");
        return DGX_EARG;
    }
    dgx_solver *s = new dgx_solver();
    s->fp32 = (p->precision == DGX_FP32 || p->precision == DGX_FP32_COMPENSATED);
    s->comp = (p->precision == DGX_FP32_COMPENSATED);
    s->exact = (p->precision == DGX_FP64_REFERENCE);
    s->tri = T;
    s->nshots = p->nshots;
    s->nsteps = p->nsteps;
    s->nrec = p->nrec;
    s->nsrc = p->nsrc;
    if (p->layout != DGX_LAYOUT_AUTO && p->layout != DGX_LAYOUT_ELEMENT && p->layout != DGX_LAYOUT_SHOT_INNER) {
        dgx_set_error("dgx_create: unknown layout %d", p->layout);
        free_solver(s);
        return DGX_EARG;
    }
    /* automatic choices: the blocked kernel in the element layout, one shot per workgroup row, at any shot count
     * and in both precisions (measured fastest: FP32 on one shot 2.2x the half-step pair, on 64 shots 1.28x the
     * shot-innermost kernels; FP64 on 20k elements 2.9x the pair, on 640k 1.0x, on 1.2M 1.43x, on 2M 1.38x, on
     * 8 and 64 shots of 1.2M elements 1.35x and 1.17x the shot-innermost kernels).  A shot tile or workgroup
     * shape given with the automatic layout asks for the shot-innermost kernels, from 8 shots */
    const bool wants_inner = p->shot_tile != 0 || p->wg_waves != 0;
    if (p->layout == DGX_LAYOUT_AUTO)
        s->layout = (wants_inner && p->nshots >= 8) ? DGX_LAYOUT_SHOT_INNER : DGX_LAYOUT_ELEMENT;
    else
        s->layout = p->layout;
    if (s->exact && (s->layout != DGX_LAYOUT_ELEMENT || (p->step_kernel != DGX_STEP_AUTO && p->step_kernel != DGX_STEP_HALF) ||
                     p->block_steps)) {
        dgx_set_error("dgx_create: DGX_FP64_REFERENCE runs in the element layout, one kernel per step function "
                      "(step_kernel AUTO or HALF, block_steps 0)");
        free_solver(s);
        return DGX_EARG;
    }
    const bool auto_blocked = p->step_kernel == DGX_STEP_AUTO && s->layout == DGX_LAYOUT_ELEMENT && !s->exact;
    if (p->step_kernel != DGX_STEP_AUTO && p->step_kernel != DGX_STEP_HALF && p->step_kernel != DGX_STEP_FULL &&
        p->step_kernel != DGX_STEP_BLOCKED) {
        dgx_set_error("dgx_create: unknown step_kernel %d", p->step_kernel);
        free_solver(s);
        return DGX_EARG;
    }
    if ((p->step_kernel == DGX_STEP_FULL || p->step_kernel == DGX_STEP_BLOCKED) && s->layout != DGX_LAYOUT_ELEMENT) {
        dgx_set_error("dgx_create: the full-step and blocked kernels need the element layout");
        free_solver(s);
        return DGX_EARG;
    }
    if (p->block_steps < 0 || p->block_steps > DGX_BLOCK_KMAX ||
        (p->block_steps && p->step_kernel != DGX_STEP_BLOCKED && !auto_blocked &&
         !(s->comp && p->step_kernel == DGX_STEP_AUTO))) {
        dgx_set_error("dgx_create: block_steps must be 0 or 1..%d, with the blocked kernel", DGX_BLOCK_KMAX);
        free_solver(s);
        return DGX_EARG;
    }
    if (s->comp && ((p->step_kernel != DGX_STEP_AUTO && p->step_kernel != DGX_STEP_BLOCKED) ||
                    s->layout != DGX_LAYOUT_ELEMENT)) {
        dgx_set_error("dgx_create: compensated FP32 runs on the blocked kernel (element layout)");
        free_solver(s);
        return DGX_EARG;
    }
    s->full = p->step_kernel == DGX_STEP_FULL;
    s->blocked = p->step_kernel == DGX_STEP_BLOCKED || s->comp || auto_blocked;
    /* steps per launch by default: FP32 3 up to 300,000 elements (compensated FP32: 50,000), where the launch and
     * the tile's loads are a large part of a step, and 3 where the two field copies exceed the 256 MB Infinity
     * Cache, so that fewer passes over them pay for the larger rings; 2 in between, where the rings'
     * recomputation dominates.  FP32 with the factored P2 derivatives, ms per step, K = 2 / 3: 120k 0.0097 /
     * 0.0079, 250k 0.0154 / 0.0142, 400k 0.0195 / 0.0196, 500k 0.0243 / 0.0226, 560k 0.0256 / 0.0254, 640k
     * 0.0284 / 0.0293, 640k x 2 shots 0.0656 / 0.0602, 2M 0.1059 / 0.0969; 64 shots of 1.2M elements 3.845 /
     * 3.706 (before the factored derivatives).  FP64 2 (its 256-thread tiles: 20k 0.0063 / 0.0073, 640k 0.0714 /
     * 0.0817, 2M 0.2059 / 0.2500) */
    const double copies_bytes = 2.0 * 30 * (s->fp32 ? 4 : 8) * (s->comp ? 2 : 1) * (double)T * (double)p->nshots;
    const bool beyond_cache = copies_bytes > 256.0 * 1024 * 1024;
    const long k3_upto = s->comp ? 50000 : 300000;
    s->bkmax = s->blocked ? (p->block_steps ? p->block_steps : (s->fp32 && (T <= k3_upto || beyond_cache) ? 3 : 2))
                          : 0;
    if (s->layout == DGX_LAYOUT_SHOT_INNER) {
        const int t = p->shot_tile;
        if (t != 0 && t != 4 && t != 8 && t != 16 && t != 32 && t != 64) {
            dgx_set_error("dgx_create: shot_tile must be 0 (autotune), 4, 8, 16, 32 or 64");
            free_solver(s);
            return DGX_EARG;
        }
        const int w = p->wg_waves;
        if (w != 0 && (t ? !dgx_batched_shape_ok(t, w) : (w != 2 && w != 4 && w != 8))) {
            dgx_set_error("dgx_create: wg_waves must be 0 (autotune), 2, 4 or 8, and 4 with shot tiles 4 and 8");
            free_solver(s);
            return DGX_EARG;
        }
        const long unit = t ? t : 8;
        s->spad = (p->nshots + unit - 1) / unit * unit;
        s->shot_tile = t ? t : 8;
        s->wg_waves = w ? w : 4;
        s->tri_pad = (T + 15) / 16 * 16; /* element-major; rows of interleaved elements need up to 16 */
        s->il = (s->shot_tile == s->spad && s->shot_tile < 64) ? __builtin_ctz(64 / s->shot_tile) : 0;
    } else if (p->shot_tile || p->wg_waves) {
        dgx_set_error("dgx_create: shot_tile and wg_waves apply to the shot-innermost layout only");
        free_solver(s);
        return DGX_EARG;
    }
    s->nper = blocks_per_xcd(s);
    if (s->layout == DGX_LAYOUT_ELEMENT) s->tri_pad = (long)s->nper * DGX_NXCD * DGX_BLOCK;

    /* ---- sources / receivers per element (caller numbering) ---- */
    std::vector<std::vector<int>> src_of(0), rec_of(0);
    std::vector<char> explosive(T, 0);
    std::vector<int> elem_aux(T, -1);
    std::vector<long> aux_elems;
    auto aux_slot = [&](long e) {
        if (elem_aux[e] < 0) {
            elem_aux[e] = (int)aux_elems.size();
            aux_elems.push_back(e);
            src_of.emplace_back();
            rec_of.emplace_back();
        }
        return elem_aux[e];
    };
    for (long k = 0; k < p->nsrc; k++) {
        const dgx_source &sr = p->src[k];
        if (sr.elem < 1 || sr.elem > T || sr.shot < 0 || sr.shot >= p->nshots || sr.wavelet < 0 ||
            sr.wavelet >= p->nwavelets || (sr.kind != DGX_SRC_FORCE_Z && sr.kind != DGX_SRC_EXPLOSIVE)) {
            dgx_set_error("source %ld: invalid element/shot/wavelet/kind", k);
            free_solver(s);
            return DGX_EARG;
        }
        src_of[aux_slot(sr.elem - 1)].push_back((int)k);
        if (sr.kind == DGX_SRC_EXPLOSIVE) explosive[sr.elem - 1] = 1;
    }
    for (long k = 0; k < p->nrec; k++) {
        const dgx_receiver &rc = p->rec[k];
        if (rc.elem < 1 || rc.elem > T) {
            dgx_set_error("receiver %ld: element %ld out of range", k, rc.elem);
            free_solver(s);
            return DGX_EARG;
        }
        rec_of[aux_slot(rc.elem - 1)].push_back((int)k);
    }
    if (aux_elems.size() > DGX_MAX_AUX) { dgx_set_error("too many source/receiver elements"); free_solver(s); return DGX_EARG; }

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        dgx_set_error("dgx_create: no HIP device available");
        free_solver(s);
        return DGX_EHIP;
    }
    CREATE_TRY(hipGetDevice(&s->device));

    /* ---- mesh: order, topology, vertices, materials, sponge, built on the device; the solver owns the
     * arrays from here (free_solver releases them) ---- */
    dgx_mesh_dev md;
    {
        int rc = dgx_build_mesh_device(p, s->layout == DGX_LAYOUT_SHOT_INNER, s->tri_pad, &explosive, &md);
        if (rc) { free_solver(s); return rc; }
    }
    s->nbr = md.d_nbr;
    s->vtx = md.d_vtx;
    s->pt = md.d_pt;
    if (s->fp32) {
        float *m = nullptr, *a = nullptr;
        const hipError_t e = dgx_mesh_f32(md.d_mat, 3 * s->tri_pad, md.d_absorb, 6 * s->tri_pad, &m, &a);
        (void)hipFree(md.d_mat);
        (void)hipFree(md.d_absorb);
        s->mat = m;
        s->absorb = a;
        CREATE_TRY(e);
    } else {
        s->mat = md.d_mat;
        s->absorb = md.d_absorb;
    }
    s->perm = std::move(md.perm);
    s->inv = std::move(md.inv);
    std::vector<int4> &nbr_host = md.nbr;

    /* ---- DGX_FP64_REFERENCE: the reference's per-element Q and material factors, formed on the host in chunks of
     * the caller's elements and placed in device order, entry k of element q at ((q/64)*34 + k)*64 + q%64 (Dx
     * and Dz are formed by the kernels from the vertices and Vr, Vs, inverV) ---- */
    std::vector<double> refops_host;
    if (s->exact) {
        memcpy(s->refmats, ops->Vr, sizeof ops->Vr);
        memcpy(s->refmats + 36, ops->Vs, sizeof ops->Vs);
        memcpy(s->refmats + 72, ops->inverV, sizeof ops->inverV);
        refops_host.assign((size_t)s->tri_pad * 34, 0.0);
        const long chunk = 65536;
        std::vector<double> Dx(chunk * 36), Dz(chunk * 36), Q(chunk * 30), A(chunk * 5), B(chunk * 5);
        for (long e0 = 0; e0 < T; e0 += chunk) {
            const long n = std::min(chunk, T - e0);
            const int rc = dgx_reference_operators(n, p->point, p->triangle + 3 * e0, p->lambda + e0, p->miu + e0,
                                                   p->den + e0, ops, Dx.data(), Dz.data(), Q.data(), A.data(),
                                                   B.data());
            if (rc) { dgx_set_error("dgx_create: reference operators failed"); free_solver(s); return rc; }
            for (long k = 0; k < n; k++) {
                const long q = s->inv[e0 + k];
                double *dst = &refops_host[(size_t)(q >> 6) * 34 * 64 + (q & 63)];
                for (int c = 0; c < 30; c++) dst[(size_t)c * 64] = Q[k * 30 + c];
                /* A = {1/rho, 1/rho, lambda + 2 mu, lambda, mu}, B = {1/rho, 1/rho, lambda, lambda + 2 mu, mu} */
                dst[30 * 64] = A[k * 5 + 0];
                dst[31 * 64] = A[k * 5 + 2];
                dst[32 * 64] = A[k * 5 + 3];
                dst[33 * 64] = A[k * 5 + 4];
            }
        }
    }

    std::vector<dgx_aux> aux_host(aux_elems.size());
    std::vector<dgx_src_dev> src_host;
    std::vector<dgx_rec_dev> rec_host;
    for (size_t x = 0; x < aux_elems.size(); x++) {
        long q = s->inv[aux_elems[x]];
        nbr_host[q].w = (int)((unsigned)nbr_host[q].w | ((unsigned)(x + 1) << DGX_CODE_AUX_SHIFT));
        aux_host[x].src_begin = (int)src_host.size();
        for (int k : src_of[x]) {
            const dgx_source &sr = p->src[k];
            dgx_src_dev d;
            d.elem = (int)q;
            d.kind = sr.kind;
            d.shot = sr.shot;
            d.wavelet = (int)sr.wavelet;
            for (int i = 0; i < 6; i++) d.fun[i] = sr.fun[i];
            src_host.push_back(d);
        }
        aux_host[x].src_count = (int)src_host.size() - aux_host[x].src_begin;
        aux_host[x].rec_begin = (int)rec_host.size();
        for (int k : rec_of[x]) {
            dgx_rec_dev d;
            d.elem = (int)q;
            d.index = k;
            d.pad0 = d.pad1 = 0;
            for (int i = 0; i < 6; i++) d.w[i] = p->rec[k].weight[i];
            rec_host.push_back(d);
        }
        aux_host[x].rec_count = (int)rec_host.size() - aux_host[x].rec_begin;
    }
    /* split-path source list in the caller's order (k_source applies them in list order) */
    std::vector<dgx_src_dev> src_list(p->nsrc);
    for (long k = 0; k < p->nsrc; k++) {
        const dgx_source &sr = p->src[k];
        src_list[k].elem = (int)s->inv[sr.elem - 1];
        src_list[k].kind = sr.kind;
        src_list[k].shot = sr.shot;
        src_list[k].wavelet = (int)sr.wavelet;
        for (int i = 0; i < 6; i++) src_list[k].fun[i] = sr.fun[i];
    }
    s->src_host = src_list;

    /* ---- full-step tiles: rows of the element range, then the tiles' ring rows ---- */
    s->vrows = s->tri_pad;
    std::vector<int4> tiles_host, nb2_host;
    std::vector<int> ring_host;
    if (s->full) {
        build_tiles(T, nbr_host, s->tri_pad, tiles_host, ring_host, nb2_host);
        s->ntile = (long)tiles_host.size();
        s->vrows = s->tri_pad + ((long)ring_host.size() + 63) / 64 * 64;
        if (s->vrows * 6 >= (1L << 31)) { dgx_set_error("mesh too large for 32-bit row indices"); free_solver(s); return DGX_EARG; }
    }

    /* ---- blocked-kernel tiles (code words carry their source/receiver bits by now) ---- */
    std::vector<int4> btiles_host, blnb_host;
    std::vector<int> blevels_host, bgid_host;
    if (s->blocked) {
        /* threads per workgroup of k_block: two workgroups per CU, so that one's loads and barriers overlap the
         * other's arithmetic.  FP32: 512 (128 VGPRs; measured 1.2-1.3x faster than one 1024-thread workgroup per
         * CU despite the larger rings).  FP64: 256 (about 235 VGPRs allow only two waves per SIMD; measured 1.07-
         * 1.16x faster than 512-thread tiles although their rings recompute 1.4-1.6 instead of 0.9 of the own
         * work).  DGX_BLOCK_THREADS selects 1024 (FP32) or 512 (FP64) instead; compensated FP32 runs only on
         * 512-thread tiles (its kernel is built for that count) */
        int cap = s->fp32 ? 512 : 256;
        if (const char *bt = getenv("DGX_BLOCK_THREADS")) {
            const int v = atoi(bt);
            if (s->fp32 && !s->comp && v == 1024) cap = 1024;
            if (!s->fp32 && v == 512) cap = 512;
        }
        s->bthreads = cap;
        if (dgx_build_block_tiles(T, nbr_host, s->bkmax, cap, btiles_host, blevels_host, bgid_host, blnb_host)) {
            s->ntile = (long)btiles_host.size();
            s->bown = T;
            s->blanes = (long)bgid_host.size();
        } else if (auto_blocked) { /* the automatic choice falls back to the half-step pair */
            s->blocked = 0;
            s->bkmax = 0;
            s->bthreads = 0;
            btiles_host.clear(); blevels_host.clear(); bgid_host.clear(); blnb_host.clear();
        } else {
            dgx_set_error("dgx_create: an element and its %d levels of neighbours exceed %d threads", 2 * s->bkmax,
                          cap);
            free_solver(s);
            return DGX_EMESH;
        }
    }

    /* ---- device allocations (all host-side validation is done above) ---- */
    {
        double *scr = nullptr; /* dgx_get_fields_noalloc's staging on this device: look it up now */
        CREATE_TRY(dgx_noalloc_scratch(&scr));
    }
    CREATE_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    for (auto &ev : s->ev) CREATE_TRY(hipEventCreate(&ev));
    size_t nreal_fields = (s->layout == DGX_LAYOUT_SHOT_INNER) ? (size_t)5 * s->tri_pad * 6 * s->spad
                                                                 : (size_t)5 * s->nshots * s->vrows * 6;
    if (s->comp) nreal_fields *= 2; /* high and low words */
    CREATE_TRY(hipMalloc(&s->fields, nreal_fields * s->real_size()));
    CREATE_TRY(hipMemset(s->fields, 0, nreal_fields * s->real_size()));
    if (s->full || s->blocked) {
        CREATE_TRY(hipMalloc(&s->fields2, nreal_fields * s->real_size()));
        CREATE_TRY(hipMemset(s->fields2, 0, nreal_fields * s->real_size()));
    }
    if (s->full) {
        CREATE_TRY(upload(&s->tiles, tiles_host));
        CREATE_TRY(upload(&s->ring, ring_host));
        CREATE_TRY(upload(&s->nb2, nb2_host));
    }
    if (s->blocked) {
        CREATE_TRY(upload(&s->tiles, btiles_host));
        CREATE_TRY(upload(&s->blevels, blevels_host));
        CREATE_TRY(dgx_upload_lane_records(bgid_host.data(), (long)bgid_host.size(), s->vtx, &s->blrec));
        CREATE_TRY(upload(&s->blnb, blnb_host));
    }
    if (s->exact) {
        CREATE_TRY(upload(&s->refops, refops_host));
        std::vector<double>().swap(refops_host);
    }
    if (!aux_elems.empty()) /* the code words now carry their source/receiver records */
        CREATE_TRY(hipMemcpy(s->nbr, nbr_host.data(), nbr_host.size() * sizeof(int4), hipMemcpyHostToDevice));
    CREATE_TRY(upload(&s->aux, aux_host));
    {
        /* fused kernels read sources through the aux ranges; the split source kernel reads the list in
         * caller order: keep both, aux-ordered first */
        std::vector<dgx_src_dev> all = src_host;
        all.insert(all.end(), src_list.begin(), src_list.end());
        CREATE_TRY(upload(&s->src, all));
    }
    CREATE_TRY(upload(&s->rec, rec_host));
    if (p->nsrc > 0) {
        std::vector<double> w(p->wavelet, p->wavelet + (size_t)p->nwavelets * p->nsteps);
        CREATE_TRY(upload(&s->wav, w));
    }
    size_t ntrace = (size_t)s->nshots * s->nsteps * 2 * s->nrec;
    if (ntrace) {
        CREATE_TRY(hipMalloc((void **)&s->trace, ntrace * sizeof(double)));
        CREATE_TRY(hipMemset(s->trace, 0, ntrace * sizeof(double)));
    }
    {
        /* device element -> caller element, for the reorders of dgx_get_fields and dgx_get_fields_noalloc */
        std::vector<int> perm(s->perm.begin(), s->perm.end());
        CREATE_TRY(upload(&s->d_perm, perm));
    }
    CREATE_TRY(hipMalloc((void **)&s->step_base, sizeof(int)));
    CREATE_TRY(hipMemset(s->step_base, 0, sizeof(int)));
    CREATE_TRY(hipDeviceSynchronize());

    if (s->fp32) fill_args(s, s->a32, ops, p->dt);
    else fill_args(s, s->a64, ops, p->dt);
    if (s->layout == DGX_LAYOUT_SHOT_INNER && p->shot_tile == 0 && s->nsteps > 0) {
        int rc = autotune(s, p->wg_waves);
        if (rc) { free_solver(s); return rc; }
    }
    *out = s;
    return DGX_OK;
}

extern "C" int dgx_full_step_tiling(const dgx_problem *p, long *ntile, long *nring, long *vrow0, long *vrows,
                                    int4 *tiles, int *ring, int4 *nb2, int4 *nbr)
{
    if (!p || p->tri <= 0) return DGX_EARG;
    const long nblk = (p->tri + DGX_BLOCK - 1) / DGX_BLOCK;
    const long tri_pad = (nblk + DGX_NXCD - 1) / DGX_NXCD * DGX_NXCD * DGX_BLOCK;
    dgx_mesh_host mh;
    if (int rc = dgx_build_mesh_host(p, 0, tri_pad, nullptr, &mh)) return rc;
    std::vector<int4> t, n2;
    std::vector<int> r;
    build_tiles(p->tri, mh.nbr, tri_pad, t, r, n2);
    if (ntile) *ntile = (long)t.size();
    if (nring) *nring = (long)r.size();
    if (vrow0) *vrow0 = tri_pad;
    if (vrows) *vrows = tri_pad + ((long)r.size() + 63) / 64 * 64;
    if (tiles) std::copy(t.begin(), t.end(), tiles);
    if (ring) std::copy(r.begin(), r.end(), ring);
    if (nb2) std::copy(n2.begin(), n2.end(), nb2);
    if (nbr) std::copy(mh.nbr.begin(), mh.nbr.begin() + p->tri, nbr); /* mh.nbr is padded */
    return DGX_OK;
}

extern "C" int dgx_block_tiling(const dgx_problem *p, int K, int cap, long *ntile, long *nlanes, int4 *tiles,
                                int *levels, int *gid, int4 *lnb, int4 *nbr)
{
    if (!p || p->tri <= 0 || K < 1 || K > DGX_BLOCK_KMAX || cap < 1) return DGX_EARG;
    const long nblk = (p->tri + DGX_BLOCK - 1) / DGX_BLOCK;
    const long tri_pad = (nblk + DGX_NXCD - 1) / DGX_NXCD * DGX_NXCD * DGX_BLOCK;
    dgx_mesh_host mh;
    if (int rc = dgx_build_mesh_host(p, 0, tri_pad, nullptr, &mh)) return rc;
    std::vector<int4> t, ln;
    std::vector<int> lv, g;
    if (!dgx_build_block_tiles(p->tri, mh.nbr, K, cap, t, lv, g, ln)) return DGX_EMESH;
    if (ntile) *ntile = (long)t.size();
    if (nlanes) *nlanes = (long)g.size();
    if (tiles) std::copy(t.begin(), t.end(), tiles);
    if (levels) std::copy(lv.begin(), lv.end(), levels);
    if (gid) std::copy(g.begin(), g.end(), gid);
    if (lnb) std::copy(ln.begin(), ln.end(), lnb);
    if (nbr) std::copy(mh.nbr.begin(), mh.nbr.begin() + p->tri, nbr);
    return DGX_OK;
}

extern "C" int dgx_destroy(dgx_solver *s)
{
    if (!s) return DGX_EARG;
    (void)hipStreamSynchronize(s->stream);
    free_solver(s);
    return DGX_OK;
}

/* DGX_FP64_REFERENCE: one step as the reference composes it (DG_leap_Frog_OpenMP.c:405-425), on the step at the
 * device counter + off, in two kernels: the velocity kernel applies the sponge and the sources to the values it
 * loads, the stress kernel applies them again to the stresses it updates and records the receivers. */
static int step_exact(dgx_solver *s, int off)
{
    int rc = launch_exact(s, 6, off);
    if (rc == DGX_OK) rc = launch_exact(s, 7, off);
    return rc;
}

static int step_eager(dgx_solver *s)
{
    int off = (int)(s->step - s->base_host);
    int rc;
    if (s->exact) {
        rc = step_exact(s, off);
        if (rc) return rc;
    } else if (s->full) {
        rc = launch(s, K_FULL, off);
        if (rc) return rc;
        swap_copies(s);
    } else {
        rc = launch(s, K_VEL_FUSED, off);
        if (rc) return rc;
        rc = launch(s, K_STRESS_FUSED, off);
        if (rc) return rc;
    }
    s->step++;
    return DGX_OK;
}

/* One launch of the blocked kernel advancing nst steps from the step after s->step. */
static int launch_block(dgx_solver *s, int step_off, int nst)
{
    s->a32.bsteps = s->a64.bsteps = nst;
    const int rc = launch(s, K_BLOCK, step_off);
    s->a32.bsteps = s->a64.bsteps = s->bkmax;
    return rc;
}

/* Steps per graph replay: 32 for the half-step and full-step kernels, kGraphLaunches launches of bkmax steps
 * for the blocked kernel. */
static long graph_steps(const dgx_solver *s) { return s->blocked ? (long)kGraphLaunches * s->bkmax : kGraphSteps; }

static int build_graph(dgx_solver *s)
{
    /* K steps with offsets 0..K-1 from the device counter, then advance the counter by K.  A full-step
     * graph alternates the two copies; K is even, so it ends on the copy it starts from. */
    static_assert(kGraphSteps % 2 == 0 && kGraphLaunches % 2 == 0,
                  "full-step and blocked graphs must end on the copy they start from");
    hipGraph_t g;
    const int c0 = s->cur;
    HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = DGX_OK;
    if (s->blocked) {
        for (int k = 0; k < kGraphLaunches && rc == DGX_OK; k++) {
            rc = launch_block(s, k * s->bkmax, s->bkmax);
            swap_copies(s);
        }
    } else {
        for (int k = 0; k < kGraphSteps && rc == DGX_OK; k++) {
            if (s->exact) {
                rc = step_exact(s, k);
            } else if (s->full) {
                rc = launch(s, K_FULL, k);
                swap_copies(s);
            } else {
                rc = launch(s, K_VEL_FUSED, k);
                if (rc == DGX_OK) rc = launch(s, K_STRESS_FUSED, k);
            }
        }
    }
    while (s->cur != c0) swap_copies(s); /* only after a failed launch */
    if (rc == DGX_OK && dgx_launch_advance(s->step_base, (int)graph_steps(s), s->stream) != hipSuccess) rc = DGX_EHIP;
    hipError_t e = hipStreamEndCapture(s->stream, &g);
    if (rc != DGX_OK) return rc;
    if (e != hipSuccess) { dgx_set_error("graph capture failed: %s", hipGetErrorString(e)); return DGX_EHIP; }
    e = hipGraphInstantiate(&s->graph[c0], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { dgx_set_error("graph instantiate failed: %s", hipGetErrorString(e)); return DGX_EHIP; }
    return DGX_OK;
}

extern "C" int dgx_step(dgx_solver *s, long n)
{
    if (!s || n < 0) return DGX_EARG;
    if (s->step + n > s->nsteps) {
        dgx_set_error("dgx_step: %ld + %ld steps exceed the capacity of %ld", s->step, n, s->nsteps);
        return DGX_ESTATE;
    }
    const long gs = graph_steps(s);
    while (n > 0) {
        if (n >= gs) {
            /* a replay works on steps *step_base + 0..K-1: bring the device counter up to date first */
            if (s->step != s->base_host) {
                if (dgx_launch_advance(s->step_base, (int)(s->step - s->base_host), s->stream) != hipSuccess) {
                    dgx_set_error("step counter update failed");
                    return DGX_EHIP;
                }
                s->base_host = s->step;
            }
            if (!s->graph[s->cur]) {
                int rc = build_graph(s);
                if (rc) return rc;
            }
            HIP_TRY(hipGraphLaunch(s->graph[s->cur], s->stream));
            s->step += gs;
            s->base_host += gs;
            n -= gs;
            continue;
        }
        if (s->blocked) {
            const int nst = (int)std::min<long>(n, s->bkmax);
            int rc = launch_block(s, (int)(s->step - s->base_host), nst);
            if (rc) return rc;
            swap_copies(s);
            s->step += nst;
            n -= nst;
            continue;
        }
        int rc = step_eager(s);
        if (rc) return rc;
        n--;
    }
    return DGX_OK;
}

/* The split entry points act on single FP64 or FP32 values: not on the compensated precision's word pairs. */
static int split_ok(const dgx_solver *s)
{
    if (!s->comp) return DGX_OK;
    dgx_set_error("the split entry points need DGX_FP64 or DGX_FP32 (compensated FP32 runs through dgx_step)");
    return DGX_ESTATE;
}

extern "C" int dgx_absorb_boundary(dgx_solver *s)
{
    if (!s) return DGX_EARG;
    if (int rc = split_ok(s)) return rc;
    return launch(s, K_ABSORB, 0);
}
extern "C" int dgx_lu_velocity(dgx_solver *s)
{
    if (!s) return DGX_EARG;
    if (int rc = split_ok(s)) return rc;
    return launch(s, K_VEL, 0);
}
extern "C" int dgx_lu_stress_elastic(dgx_solver *s)
{
    if (!s) return DGX_EARG;
    if (int rc = split_ok(s)) return rc;
    return launch(s, K_STRESS, 0);
}

static int needs_element_layout(const dgx_solver *s)
{
    if (int rc = split_ok(s)) return rc;
    if (s->layout == DGX_LAYOUT_ELEMENT) return DGX_OK;
    dgx_set_error("the split entry points need the element layout (DGX_LAYOUT_ELEMENT)");
    return DGX_ESTATE;
}

extern "C" int dgx_source_kernel(dgx_solver *s, double w)
{
    if (!s) return DGX_EARG;
    if (int rc = needs_element_layout(s)) return rc;
    if (s->nsrc == 0) return DGX_OK;
    if (s->exact) return launch_exact(s, 3, 0, -1, w);
    /* the caller-ordered list follows the aux-ordered copy in the same allocation */
    dgx::StepArgs<double> a64 = s->a64;
    dgx::StepArgs<float> a32 = s->a32;
    a64.src = s->src + s->nsrc;
    a32.src = s->src + s->nsrc;
    hipError_t e = dgx_launch_source(s->fp32 ? (const void *)&a32 : (const void *)&a64, s->fp32, (int)s->nsrc, w,
                                     s->stream);
    if (e != hipSuccess) { dgx_set_error("source launch failed: %s", hipGetErrorString(e)); return DGX_EHIP; }
    return DGX_OK;
}

extern "C" int dgx_record(dgx_solver *s, long t)
{
    if (!s || t < 0 || t >= s->nsteps) return DGX_EARG;
    if (int rc = needs_element_layout(s)) return rc;
    if (s->exact) return launch_exact(s, 5, 0, (int)t);
    hipError_t e = dgx_launch_record(s->args(), s->fp32, (int)t, (int)s->nshots, (int)s->nrec, s->stream);
    if (e != hipSuccess) { dgx_set_error("record launch failed: %s", hipGetErrorString(e)); return DGX_EHIP; }
    return DGX_OK;
}

/* Index of (q, i) of one shot inside the raw copy of a field: for the element layout the raw copy is
 * that shot's array; for the shot-innermost layout it is the whole field and the shot is added. */
static inline long long raw_index(const dgx_solver *s, long q, int i, long shot)
{
    if (s->layout == DGX_LAYOUT_SHOT_INNER) {
        const long long L = s->il;
        return ((((long long)(q >> L) * 6 + i) << L) + (q & ((1 << L) - 1))) * s->spad + shot;
    }
    return fidx(q, i);
}

template <typename R>
static void unpack(const dgx_solver *s, const std::vector<char> &raw, long shot, double *out)
{
    const R *r = (const R *)raw.data();
    for (long q = 0; q < s->tri; q++) {
        long e = s->perm[q];
        for (int i = 0; i < 6; i++) out[6 * e + i] = (double)r[raw_index(s, q, i, shot)];
    }
}

template <typename R>
static void pack(const dgx_solver *s, const double *in, long shot, std::vector<char> &raw)
{
    R *r = (R *)raw.data();
    for (long q = 0; q < s->tri; q++) {
        long e = s->perm[q];
        for (int i = 0; i < 6; i++) r[raw_index(s, q, i, shot)] = (R)in[6 * e + i];
    }
}

/* Bytes and byte offset of the raw copy of field c used for `shot`. */
static void raw_span(const dgx_solver *s, int c, long shot, size_t *off, size_t *len)
{
    size_t rs = s->real_size();
    if (s->layout == DGX_LAYOUT_SHOT_INNER) {
        *len = (size_t)s->tri_pad * 6 * s->spad * rs;
        *off = (size_t)c * *len;
    } else {
        *len = (size_t)s->tri_pad * 6 * rs;
        *off = ((size_t)c * s->nshots + shot) * (size_t)s->vrows * 6 * rs;
    }
}

/* Snapshot buffers, allocated at the first snapshot and kept: the device and pinned host copies hold all
 * five fields of one shot (tri*6 doubles each). */
static int snapshot_buffers(dgx_solver *s)
{
    /* each piece once, so that a call after a failed allocation neither leaks nor repeats the others */
    const size_t n = (size_t)s->tri * 6 * 5;
    if (!s->d_snap) HIP_TRY(hipMalloc((void **)&s->d_snap, n * sizeof(double)));
    if (!s->copy_stream) HIP_TRY(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
    if (!s->snap_ready) HIP_TRY(hipEventCreateWithFlags(&s->snap_ready, hipEventDisableTiming));
    if (!s->snap_done) HIP_TRY(hipEventCreateWithFlags(&s->snap_done, hipEventDisableTiming));
    if (!s->h_snap) HIP_TRY(hipHostMalloc((void **)&s->h_snap, n * sizeof(double), hipHostMallocDefault));
    return DGX_OK;
}

extern "C" int dgx_snapshot_begin(dgx_solver *s, long shot, int mask)
{
    if (!s || shot < 0 || shot >= s->nshots || mask <= 0 || mask > 31) {
        dgx_set_error("dgx_snapshot_begin: bad shot or field mask");
        return DGX_EARG;
    }
    if (s->snap_mask) {
        dgx_set_error("dgx_snapshot_begin: the previous snapshot has not been collected (dgx_snapshot_wait)");
        return DGX_ESTATE;
    }
    int rc = snapshot_buffers(s);
    if (rc) return rc;
    const size_t nf = (size_t)s->tri * 6;
    const long spad = s->layout == DGX_LAYOUT_SHOT_INNER ? s->spad : 0;
    /* on the solver's stream, after the steps enqueued so far: device-side reorder into the caller's order */
    for (int c = 0; c < 5; c++) {
        if (!((mask >> c) & 1)) continue;
        size_t off, len;
        raw_span(s, c, shot, &off, &len);
        const void *lo = nullptr;
        if (s->comp) {
            size_t loff, llen;
            raw_span(s, c + 5, shot, &loff, &llen);
            lo = (const char *)s->state() + loff;
        }
        HIP_TRY(dgx_launch_unpermute((const char *)s->state() + off, lo, s->fp32, s->d_perm, s->d_snap + c * nf, s->tri,
                                     spad, s->il, shot, s->stream));
    }
    HIP_TRY(hipEventRecord(s->snap_ready, s->stream));
    /* on the copy stream: the download, concurrent with the steps enqueued after this call */
    HIP_TRY(hipStreamWaitEvent(s->copy_stream, s->snap_ready, 0));
    for (int c = 0; c < 5; c++)
        if ((mask >> c) & 1)
            HIP_TRY(hipMemcpyAsync(s->h_snap + c * nf, s->d_snap + c * nf, nf * sizeof(double), hipMemcpyDeviceToHost,
                                   s->copy_stream));
    HIP_TRY(hipEventRecord(s->snap_done, s->copy_stream));
    s->snap_mask = mask;
    return DGX_OK;
}

extern "C" int dgx_snapshot_wait(dgx_solver *s, double *vx, double *vz, double *txx, double *tzz, double *txz)
{
    if (!s) return DGX_EARG;
    if (!s->snap_mask) {
        dgx_set_error("dgx_snapshot_wait: no snapshot in flight");
        return DGX_ESTATE;
    }
    double *outs[5] = {vx, vz, txx, tzz, txz};
    const size_t nf = (size_t)s->tri * 6;
    const int mask = s->snap_mask;
    s->snap_mask = 0;
    HIP_TRY(hipEventSynchronize(s->snap_done));
    for (int c = 0; c < 5; c++)
        if (((mask >> c) & 1) && outs[c]) std::memcpy(outs[c], s->h_snap + c * nf, nf * sizeof(double));
    return DGX_OK;
}

/* The caller's current device, restored when a call that switched to the solver's device returns. */
struct DeviceGuard {
    int prev = -1;
    hipError_t use(int dev)
    {
        hipError_t e = hipGetDevice(&prev);
        if (e != hipSuccess) { prev = -1; return e; }
        if (prev == dev) { prev = -1; return hipSuccess; }
        return hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

extern "C" int dgx_get_fields_noalloc(dgx_solver *s, long shot, double *vx, double *vz, double *txx, double *tzz,
                                      double *txz)
{
    if (!s || shot < 0 || shot >= s->nshots) {
        dgx_set_error("dgx_get_fields_noalloc: bad solver or shot");
        return DGX_EARG;
    }
    double *outs[5] = {vx, vz, txx, tzz, txz};
    int nreq = 0;
    for (int c = 0; c < 5; c++) nreq += outs[c] != nullptr;
    if (!nreq) return DGX_OK;
    const size_t nf = (size_t)s->tri * 6;
    if ((size_t)nreq * nf * sizeof(double) > (size_t)DGX_NOALLOC_SCRATCH_BYTES) {
        dgx_set_error("dgx_get_fields_noalloc: %d field(s) of %ld elements need %zu bytes of staging, more than the "
                      "%ld bytes of the static array; request fewer fields per call or use dgx_get_fields",
                      nreq, s->tri, (size_t)nreq * nf * sizeof(double), (long)DGX_NOALLOC_SCRATCH_BYTES);
        return DGX_EARG;
    }
    DeviceGuard dg;
    HIP_TRY(dg.use(s->device));
    double *scr = nullptr;
    HIP_TRY(dgx_noalloc_scratch(&scr));
    const long spad = s->layout == DGX_LAYOUT_SHOT_INNER ? s->spad : 0;
    /* on the solver's stream, after the steps enqueued so far: reorder each field into its slice of the
     * array, then copy the slices out */
    for (int c = 0, k = 0; c < 5; c++) {
        if (!outs[c]) continue;
        size_t off, len;
        raw_span(s, c, shot, &off, &len);
        const void *lo = nullptr;
        if (s->comp) {
            size_t loff, llen;
            raw_span(s, c + 5, shot, &loff, &llen);
            lo = (const char *)s->state() + loff;
        }
        HIP_TRY(dgx_launch_unpermute((const char *)s->state() + off, lo, s->fp32, s->d_perm, scr + k * nf, s->tri,
                                     spad, s->il, shot, s->stream));
        k++;
    }
    for (int c = 0, k = 0; c < 5; c++) {
        if (!outs[c]) continue;
        HIP_TRY(hipMemcpyAsync(outs[c], scr + k * nf, nf * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        k++;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return DGX_OK;
}

extern "C" int dgx_get_fields(dgx_solver *s, long shot, double *vx, double *vz, double *txx, double *tzz, double *txz)
{
    if (!s || shot < 0 || shot >= s->nshots) return DGX_EARG;
    double *outs[5] = {vx, vz, txx, tzz, txz};
    if (!s->snap_mask) {
        /* no snapshot in flight: the snapshot path (device-side reorder, pinned download) */
        int mask = 0;
        for (int c = 0; c < 5; c++)
            if (outs[c]) mask |= 1 << c;
        if (!mask) return DGX_OK;
        int rc = dgx_snapshot_begin(s, shot, mask);
        if (rc) return rc;
        return dgx_snapshot_wait(s, vx, vz, txx, tzz, txz);
    }
    /* a snapshot of the caller's is in flight and owns the snapshot buffers: copy and reorder on the host */
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<char> raw;
    for (int c = 0; c < 5; c++) {
        if (!outs[c]) continue;
        size_t off, len;
        raw_span(s, c, shot, &off, &len);
        raw.resize(len);
        HIP_TRY(hipMemcpy(raw.data(), (const char *)s->state() + off, len, hipMemcpyDeviceToHost));
        if (s->fp32) unpack<float>(s, raw, shot, outs[c]);
        else unpack<double>(s, raw, shot, outs[c]);
        if (s->comp) { /* add the low words */
            std::vector<double> lo((size_t)s->tri * 6);
            raw_span(s, c + 5, shot, &off, &len);
            HIP_TRY(hipMemcpy(raw.data(), (const char *)s->state() + off, len, hipMemcpyDeviceToHost));
            unpack<float>(s, raw, shot, lo.data());
            for (size_t k = 0; k < lo.size(); k++) outs[c][k] += lo[k];
        }
    }
    return DGX_OK;
}

extern "C" int dgx_set_fields(dgx_solver *s, long shot, const double *vx, const double *vz, const double *txx,
                              const double *tzz, const double *txz)
{
    if (!s || shot < 0 || shot >= s->nshots) return DGX_EARG;
    const double *ins[5] = {vx, vz, txx, tzz, txz};
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<char> raw;
    for (int c = 0; c < 5; c++) {
        if (!ins[c]) continue;
        size_t off, len;
        raw_span(s, c, shot, &off, &len);
        raw.assign(len, 0);
        /* read-modify-write so the other shots sharing the array keep their values */
        HIP_TRY(hipMemcpy(raw.data(), (const char *)s->state() + off, len, hipMemcpyDeviceToHost));
        if (s->fp32) pack<float>(s, ins[c], shot, raw);
        else pack<double>(s, ins[c], shot, raw);
        HIP_TRY(hipMemcpy((char *)s->state() + off, raw.data(), len, hipMemcpyHostToDevice));
        if (s->comp) { /* low word: what the FP32 high word leaves of the value */
            std::vector<double> rest((size_t)s->tri * 6);
            for (size_t k = 0; k < rest.size(); k++) rest[k] = ins[c][k] - (double)(float)ins[c][k];
            raw_span(s, c + 5, shot, &off, &len);
            HIP_TRY(hipMemcpy(raw.data(), (const char *)s->state() + off, len, hipMemcpyDeviceToHost));
            pack<float>(s, rest.data(), shot, raw);
            HIP_TRY(hipMemcpy((char *)s->state() + off, raw.data(), len, hipMemcpyHostToDevice));
        }
    }
    return DGX_OK;
}

extern "C" int dgx_get_traces(dgx_solver *s, long shot, double *rec_x, double *rec_z)
{
    if (!s || shot < 0 || shot >= s->nshots) return DGX_EARG;
    if (s->nrec == 0 || s->nsteps == 0) return DGX_OK;
    std::vector<double> t((size_t)s->nsteps * 2 * s->nrec);
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(t.data(), s->trace + (size_t)shot * s->nsteps * 2 * s->nrec, t.size() * sizeof(double),
                      hipMemcpyDeviceToHost));
    for (long st = 0; st < s->nsteps; st++)
        for (long k = 0; k < s->nrec; k++) {
            if (rec_x) rec_x[st * s->nrec + k] = t[(st * 2 + 0) * s->nrec + k];
            if (rec_z) rec_z[st * s->nrec + k] = t[(st * 2 + 1) * s->nrec + k];
        }
    return DGX_OK;
}

extern "C" int dgx_synchronize(dgx_solver *s)
{
    if (!s) return DGX_EARG;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return DGX_OK;
}

extern "C" long dgx_current_step(const dgx_solver *s) { return s ? s->step : -1; }

extern "C" int dgx_kernel_bench(dgx_solver *s, int which, long n, double *ms)
{
    /* n back-to-back launches of one fused kernel between two events: the kernel's own duration
     * without per-launch event overhead.  The launches repeat the current step, so the fields are
     * no longer a valid state afterwards. */
    if (!s || n <= 0 || which < 0 || which > 3) {
        dgx_set_error("dgx_kernel_bench: bad arguments");
        return DGX_EARG;
    }
    if (s->step + (which == 3 ? s->bkmax : 1) > s->nsteps) {
        dgx_set_error("dgx_kernel_bench: the launches repeat the next step, and step %ld + %d exceeds the %ld-step "
                      "capacity", s->step, which == 3 ? s->bkmax : 1, s->nsteps);
        return DGX_EARG;
    }
    int off = (int)(s->step - s->base_host);
    int k = which == 0 ? K_VEL_FUSED : which == 1 ? K_STRESS_FUSED : which == 2 ? K_FULL : K_BLOCK;
    int rc = launch(s, k, off); /* warm */
    if (rc) return rc;
    HIP_TRY(hipEventRecord(s->ev[0], s->stream));
    for (long i = 0; i < n; i++) {
        rc = launch(s, k, off);
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(s->ev[1], s->stream));
    HIP_TRY(hipEventSynchronize(s->ev[1]));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, s->ev[0], s->ev[1]));
    if (ms) *ms = t / n;
    return DGX_OK;
}

extern "C" int dgx_copy_probe(long bytes, long reps, int variant, double *ms)
{
    if (bytes < 1024 || reps < 1 || variant < 0 || variant > 2) return DGX_EARG;
    long n32 = bytes / 32;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0, e1;
    HIP_TRY(hipMalloc(&a, n32 * 32));
    HIP_TRY(hipMalloc(&b, n32 * 32));
    HIP_TRY(hipMemset(a, 0, n32 * 32));
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(dgx_launch_copy(a, b, n32, variant, 0));
    HIP_TRY(hipEventRecord(e0, 0));
    for (long r = 0; r < reps; r++) HIP_TRY(dgx_launch_copy(a, b, n32, variant, 0));
    HIP_TRY(hipEventRecord(e1, 0));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return DGX_OK;
}

extern "C" int dgx_layout_info(const dgx_solver *s, int *layout, int *shot_tile, int *tiles, double *ms_per_step,
                               int cap)
{
    if (!s) return DGX_EARG;
    if (layout) *layout = s->layout;
    if (shot_tile) *shot_tile = s->shot_tile;
    int n = (int)s->tune_tiles.size();
    for (int k = 0; k < n && k < cap; k++) {
        if (tiles) tiles[k] = s->tune_tiles[k];
        if (ms_per_step) ms_per_step[k] = s->tune_ms[k];
    }
    return n;
}

extern "C" int dgx_tile_info(const dgx_solver *s, int *shot_tile, int *wg_waves, int *tiles, int *waves,
                             double *ms_per_step, int cap)
{
    if (!s) return DGX_EARG;
    if (shot_tile) *shot_tile = s->layout == DGX_LAYOUT_SHOT_INNER ? s->shot_tile : 0;
    if (wg_waves) *wg_waves = s->layout == DGX_LAYOUT_SHOT_INNER ? s->wg_waves : 0;
    int n = (int)s->tune_tiles.size();
    for (int k = 0; k < n && k < cap; k++) {
        if (tiles) tiles[k] = s->tune_tiles[k];
        if (waves) waves[k] = s->tune_waves[k];
        if (ms_per_step) ms_per_step[k] = s->tune_ms[k];
    }
    return n;
}

extern "C" int dgx_block_info(const dgx_solver *s, int *steps_per_launch, long *ntile, long *own, long *lanes)
{
    if (!s) return DGX_EARG;
    if (steps_per_launch) *steps_per_launch = s->bkmax;
    if (ntile) *ntile = s->blocked ? s->ntile : 0;
    if (own) *own = s->bown;
    if (lanes) *lanes = s->blanes;
    return DGX_OK;
}

extern "C" int dgx_compiled_operators(const dgx_solver *s)
{
    if (!s) return DGX_EARG;
    if (s->layout != DGX_LAYOUT_ELEMENT) return 0;
    return s->fp32 ? s->a32.p2 : s->a64.p2;
}

extern "C" int dgx_kernel_times(dgx_solver *s, long n, double *ms_velocity, double *ms_stress)
{
    /* n fused steps, each kernel bracketed by events on the solver's stream */
    if (!s || n <= 0 || s->step + n > s->nsteps) return DGX_EARG;
    double tv = 0, ts = 0;
    for (long k = 0; k < n; k++) {
        int off = (int)(s->step - s->base_host);
        HIP_TRY(hipEventRecord(s->ev[0], s->stream));
        int rc = launch(s, K_VEL_FUSED, off);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(s->ev[1], s->stream));
        rc = launch(s, K_STRESS_FUSED, off);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(s->ev[2], s->stream));
        HIP_TRY(hipEventSynchronize(s->ev[2]));
        float a = 0, b = 0;
        HIP_TRY(hipEventElapsedTime(&a, s->ev[0], s->ev[1]));
        HIP_TRY(hipEventElapsedTime(&b, s->ev[1], s->ev[2]));
        tv += a;
        ts += b;
        s->step++;
    }
    if (ms_velocity) *ms_velocity = tv / n;
    if (ms_stress) *ms_stress = ts / n;
    return DGX_OK;
}
