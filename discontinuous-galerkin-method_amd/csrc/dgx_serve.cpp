/*
 * dgx_serve.cpp -- request loop (include/dgx_serve.h): resident meshes, slot groups of 4, 8, 16, 32 or 64 shots in the
 * shot-innermost layout, memory-bounded admission, one kernel per half-step over the tiles of all active
 * batches, hipGraph-captured blocks of steps, retirement with trace read-back.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "dgx.h"
#include "dgx_internal.h"
#include "dgx_kernels.h"
#include "dgx_p2ops.h"
#include "dgx_serve.h"

extern "C" {
hipError_t dgx_launch_serve(int fp32, const void *ops, const void *table, const int *sched,
                            const unsigned long long *runmask, const int *step_base, int step_off, int vel, int group,
                            int nb, int total, hipStream_t st);
hipError_t dgx_launch_clear_slots(int fp32, void *const *f, long long n_nodes, int group, unsigned long long mask,
                                  hipStream_t st);
hipError_t dgx_launch_advance(int *step_base, int n, hipStream_t st);
hipError_t dgx_launch_scatter(int *dst, const int2 *pairs, int n, hipStream_t st);
hipError_t dgx_launch_move_slots(int fp32, void *const *src, void *const *dst, long long n_nodes, int group, int n,
                                 const int *from, const int *to, hipStream_t st);
hipError_t dgx_launch_serve_block(int fp32, int p2, const void *table, const int *sched, const int *step_base,
                                  int step_off, int launch, int nb, int total, hipStream_t st);
}
/* the blocked kernel's tiling (dgx_api.cpp) */
bool dgx_build_block_tiles(long tri, const std::vector<int4> &nbr, int K, int cap, std::vector<int4> &tiles,
                           std::vector<int> &levels, std::vector<int> &gid, std::vector<int4> &lnb);

#define SV_TRY(call)                                                                                           \
    do {                                                                                                       \
        hipError_t e_ = (call);                                                                                \
        if (e_ != hipSuccess) {                                                                                \
            dgx_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__);          \
            return DGX_EHIP;                                                                                   \
        }                                                                                                      \
    } while (0)

namespace {

const int MAXG = 64;      /* largest slot group (shots per batch = shot tile of the serving kernel) */
const int MAX_BATCHES = 4096;


/* A material model on a resident mesh: 1/rho | lambda | mu per element in device order, and the time step
 * its requests run with.  Model 0 of every mesh is the one given at dgx_server_add_mesh and stays resident;
 * the others come with requests, are shared by every request with the same arrays and dt, and are on the
 * device while one of their requests runs. */
struct ModelDev {
    bool used = false;
    uint64_t hash = 0;
    std::vector<double> mat; /* host copy, 3 * tri (empty for model 0, whose copy lives on the device only) */
    double dt = 0;
    long refs = 0;           /* requests queued or running with this model */
    long live = 0;           /* running requests */
    void *d_mat = nullptr;
};

struct MeshDev {
    long tri = 0;
    long tri_pad = 0;      /* single-shot slots: element rows of a field (a multiple of 64) */
    /* single-shot slots: the blocked kernel's tiling of the mesh */
    int4 *btiles = nullptr, *blnb = nullptr;
    int *blevels = nullptr;
    int4 *blrec = nullptr;
    long ntile = 0;
    std::vector<long> inv; /* caller element -> device element */
    std::vector<int4> nbr_host; /* device-order neighbours (finding the faces next to a source) */
    int4 *nbr = nullptr, *vtx = nullptr;
    double2 *pt = nullptr;
    void *absorb = nullptr;
    std::vector<ModelDev> models;
    std::vector<int> batches;
};

struct Request {
    long id = 0;
    int mesh = 0, nshots = 0;
    int model = 0; /* index into the mesh's models */
    long nsteps = 0, nrec = 0;
    std::vector<dgx_source> src;
    std::vector<double> wav;
    std::vector<dgx_receiver> rec;
    int state = DGX_REQ_QUEUED;
    double arrival = 0, t_admit = 0, t_done = 0;
    long start = 0;
    std::vector<std::pair<int, int>> slots; /* (batch, slot) per shot */
    std::vector<double *> d_wav, d_trace;
    std::vector<double> res_x, res_z;
    long long bytes = 0;
};

struct Batch {
    bool live = true; /* false: freed (all slots were empty when memory was needed); the entry is reused */
    int mesh = 0;
    int model = 0;    /* every shot of a group runs on one material model of its mesh */
    void *fields = nullptr;
    long long field_bytes = 0;
    int G = 16;                     /* slots */
    long req[MAXG];
    int shot[MAXG];
    int h_start[MAXG], h_len[MAXG], h_nrec[MAXG];
    double *h_wav[MAXG], *h_trace[MAXG];
    int *d_slot_int = nullptr;      /* start[G] | len[G] | nrec[G] */
    double **d_slot_ptr = nullptr;  /* wav[G] | trace[G] */
    int *d_aux_of_elem = nullptr;
    std::vector<int> aux_elems;     /* device elements whose per-element word is currently nonzero */
    dgx_aux *d_aux = nullptr;
    dgx_src_dev *d_src = nullptr;
    dgx_rec_dev *d_rec = nullptr;
    int2 *d_pairs = nullptr;
    long aux_cap = 0, src_cap = 0, rec_cap = 0, pair_cap = 0;
    bool dirty = false;
    int used() const
    {
        int n = 0;
        for (int k = 0; k < G; k++) n += req[k] >= 0;
        return n;
    }
};

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename T>
hipError_t grow(T **p, long *cap, long need)
{
    if (need <= *cap) return hipSuccess;
    long c = std::max<long>(need, *cap * 2 + 16);
    if (*p) {
        hipError_t e = hipFree(*p);
        if (e != hipSuccess) return e;
    }
    *p = nullptr;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * c);
    if (e != hipSuccess) return e;
    *cap = c;
    return hipSuccess;
}

} // namespace

struct dgx_server {
    int fp32 = 1;
    int group = 16; /* shots per slot group; 1: single-shot slots stepped by the blocked kernel */
    int p2 = 0;     /* the operators equal the compiled-in P2 table (the blocked kernel's constants) */
    int nb = 0, tiles = 0; /* stepped batches and their tiles (sched); the step graph is built for these */
    std::vector<int> order; /* batch entry of each stepped batch, in table order */
    unsigned long long *d_runmask = nullptr; /* [step of block][stepped batch]: slots running at that step */
    long long budget = 0, in_use = 0;
    int K = 100;
    /* single-shot slots: steps per launch of the blocked kernel.  3 in FP32 when K allows it (a multiple of 6):
     * the slots' two field copies are far beyond the Infinity Cache, where one pass over them per three steps
     * pays for the larger rings (the solver's blocked kernel on 64 shots of 1.2M elements: 3.71 against 3.85
     * ms per step); else 2 */
    int bk = 2;
    dgx_operators ops;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int *d_step_base = nullptr;
    long global_step = 0;
    void *d_table = nullptr;
    int *d_sched = nullptr;
    std::vector<MeshDev> meshes;
    std::vector<Batch> batches;
    std::map<long, Request> reqs;
    std::deque<long> queue;
    long next_id = 1;
    long long dof_updates = 0;
    long long lane_updates = 0; /* slot lanes stepped, idle slots included */
    long long moved = 0;        /* shots moved between groups by compaction */
    double busy = 0;
    double gpu = 0;             /* graph execution time between events on the stream */
    hipEvent_t ev[2] = {nullptr, nullptr};
    long blocks = 0;
    bool table_dirty = true;
    dgx::Ops<double> op64;
    dgx::Ops<float> op32;
    /* memory pool of per-request device buffers, by size */
    std::multimap<size_t, double *> pool;
    long fail_alloc_at = 0, allocs = 0; /* test hook: DGX_SERVE_FAIL_ALLOC (pool_get) */

    size_t rsize() const { return fp32 ? 4 : 8; }
};

/* Threads per workgroup (lanes per tile) of the single-shot slots' blocked kernel: 512 in FP32, 256 in FP64 (two
 * workgroups per CU in both, as for the solver's blocked kernel; dgx_launch_serve_block launches that many). */
static int serve_block_threads(const dgx_server *sv) { return sv->fp32 ? 512 : 256; }

/* Slot groups that hold n shots. */
static int groups_for(const dgx_server *sv, int n) { return (n + sv->group - 1) / sv->group; }

/* Node rows of G lanes per field in a slot group of mesh m: the serving kernel interleaves 64/G elements
 * per row, so the element count is rounded up to a multiple of 64/G. */
static long long group_rows(const dgx_server *sv, const MeshDev &m)
{
    if (sv->group == 1) return m.tri_pad * 6; /* element layout, 64-element blocks */
    const long long per = 64 / sv->group;
    return (m.tri + per - 1) / per * per * 6;
}

/* Device bytes of one slot group of mesh m (fields and the per-element word). */
static long long group_bytes(const dgx_server *sv, const MeshDev &m)
{
    const int copies = sv->group == 1 ? 2 : 1; /* single-shot slots alternate between two copies */
    return (long long)copies * 5 * group_rows(sv, m) * sv->group * sv->rsize() + 4 * m.tri;
}

static void make_ops(dgx_server *sv)
{
    for (int i = 0; i < 6; i++)
        for (int n = 0; n < 6; n++) {
            double dr = 0, ds = 0;
            for (int k = 0; k < 6; k++) {
                dr += sv->ops.Vr[i * 6 + k] * sv->ops.inverV[k * 6 + n];
                ds += sv->ops.Vs[i * 6 + k] * sv->ops.inverV[k * 6 + n];
            }
            sv->op64.Dr[i * 6 + n] = dr;
            sv->op64.Ds[i * 6 + n] = ds;
            sv->op32.Dr[i * 6 + n] = (float)dr;
            sv->op32.Ds[i * 6 + n] = (float)ds;
        }
    const double *S[3] = {sv->ops.S1, sv->ops.S2, sv->ops.S3};
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 18; k++) {
            sv->op64.L[j * 18 + k] = S[j][k];
            sv->op32.L[j * 18 + k] = (float)S[j][k];
        }
}

/* Free every pooled request buffer (the pool lies outside the budget and bytes_in_use). */
static void pool_release(dgx_server *sv)
{
    for (auto &kv : sv->pool) (void)hipFree(kv.second);
    sv->pool.clear();
}

static double *pool_get(dgx_server *sv, size_t n)
{
    auto it = sv->pool.lower_bound(n);
    if (it != sv->pool.end() && it->first <= 2 * n) {
        double *p = it->second;
        sv->pool.erase(it);
        return p;
    }
    /* DGX_SERVE_FAIL_ALLOC=k (read at create): the k-th device allocation of a request buffer fails, as if
     * HBM were exhausted, so that tests can drive admission through its failure path */
    if (sv->fail_alloc_at > 0 && ++sv->allocs == sv->fail_alloc_at) return nullptr;
    double *p = nullptr;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(double);
    if (hipMalloc((void **)&p, bytes) == hipSuccess) return p;
    /* pooled buffers of other sizes may hold the memory: give them back and try once more */
    (void)hipGetLastError();
    pool_release(sv);
    p = nullptr;
    if (hipMalloc((void **)&p, bytes) == hipSuccess) return p;
    (void)hipGetLastError();
    return nullptr;
}

static void pool_put(dgx_server *sv, double *p, size_t n) { sv->pool.emplace(n, p); }

extern "C" double dgx_server_now(void) { return now_s(); }

extern "C" int dgx_server_create(const dgx_server_config *cfg, const dgx_operators *ops, dgx_server **out)
{
    if (!ops || !out) return DGX_EARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        dgx_set_error("dgx_server_create: no HIP device available");
        return DGX_EHIP;
    }
    if (cfg && cfg->precision != DGX_FP64 && cfg->precision != DGX_FP32) {
        dgx_set_error("dgx_server_create: precision must be DGX_FP32 or DGX_FP64");
        return DGX_EARG;
    }
    dgx_server *sv = new dgx_server();
    sv->fp32 = cfg ? (cfg->precision == DGX_FP64 ? 0 : 1) : 1;
    sv->K = (cfg && cfg->steps_per_block > 0) ? cfg->steps_per_block : 100;
    sv->group = (cfg && cfg->group_shots) ? cfg->group_shots : 16;
    if (sv->group != 1 && sv->group != 4 && sv->group != 8 && sv->group != 16 && sv->group != 32 && sv->group != 64) {
        dgx_set_error("dgx_server_create: group_shots must be 0, 1, 4, 8, 16, 32 or 64");
        delete sv;
        return DGX_EARG;
    }
    sv->bk = (sv->fp32 && sv->K % 6 == 0) ? 3 : 2;
    if (sv->group == 1 && sv->K % (2 * sv->bk)) {
        dgx_set_error("dgx_server_create: with single-shot slots steps_per_block must be a multiple of %d",
                      2 * sv->bk);
        delete sv;
        return DGX_EARG;
    }
    sv->ops = *ops;
    make_ops(sv);
    {
        /* the blocked kernel compiles in the shipped P2 operators: usable when these equal them bit for bit */
        bool eq = true;
        const double *o64 = &sv->op64.Dr[0];
        for (int k = 0; k < 126 && eq; k++) {
            const double v = o64[k];
            eq = sv->fp32 ? ((float)v == (float)dgx_p2_ops[k]) : (v == dgx_p2_ops[k]);
        }
        const char *generic = getenv("DGX_GENERIC_OPERATORS");
        sv->p2 = eq && !(generic && strcmp(generic, "1") == 0);
    }
    if (const char *f = getenv("DGX_SERVE_FAIL_ALLOC")) sv->fail_alloc_at = atol(f);
    /* a failure from here on frees what was already created */
#define CREATE_TRY(call)                                                                                       \
    do {                                                                                                       \
        hipError_t e_ = (call);                                                                                \
        if (e_ != hipSuccess) {                                                                                \
            dgx_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__);          \
            (void)dgx_server_destroy(sv);                                                                      \
            return DGX_EHIP;                                                                                   \
        }                                                                                                      \
    } while (0)
    size_t fr = 0, tot = 0;
    CREATE_TRY(hipMemGetInfo(&fr, &tot));
    sv->budget = (cfg && cfg->hbm_budget > 0) ? cfg->hbm_budget : (long long)(0.8 * fr);
    CREATE_TRY(hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking));
    CREATE_TRY(hipEventCreate(&sv->ev[0]));
    CREATE_TRY(hipEventCreate(&sv->ev[1]));
    CREATE_TRY(hipMalloc((void **)&sv->d_step_base, sizeof(int)));
    CREATE_TRY(hipMemset(sv->d_step_base, 0, sizeof(int)));
    size_t args = sv->fp32 ? sizeof(dgx::StepArgs<float>) : sizeof(dgx::StepArgs<double>);
    CREATE_TRY(hipMalloc(&sv->d_table, args * MAX_BATCHES));
    CREATE_TRY(hipMalloc((void **)&sv->d_sched, sizeof(int) * (MAX_BATCHES + 3)));
    CREATE_TRY(hipMalloc((void **)&sv->d_runmask, sizeof(unsigned long long) * MAX_BATCHES * sv->K));
    int zero[3] = {0, 0, 0};
    CREATE_TRY(hipMemcpy(sv->d_sched, zero, sizeof(zero), hipMemcpyHostToDevice));
#undef CREATE_TRY
    *out = sv;
    return DGX_OK;
}

extern "C" int dgx_server_destroy(dgx_server *sv)
{
    if (!sv) return DGX_EARG;
    if (sv->stream) (void)hipStreamSynchronize(sv->stream);
    if (sv->graph) (void)hipGraphExecDestroy(sv->graph);
    for (auto &m : sv->meshes) {
        void *p[] = {m.nbr, m.vtx, m.pt, m.absorb, m.btiles, m.blnb, m.blevels, m.blrec};
        for (void *x : p)
            if (x) (void)hipFree(x);
        for (auto &md : m.models)
            if (md.d_mat) (void)hipFree(md.d_mat);
    }
    for (auto &b : sv->batches) {
        void *p[] = {b.fields, b.d_slot_int, b.d_slot_ptr, b.d_aux_of_elem, b.d_aux, b.d_src, b.d_rec, b.d_pairs};
        for (void *x : p)
            if (x) (void)hipFree(x);
    }
    for (auto &r : sv->reqs) {
        for (double *p : r.second.d_wav) (void)hipFree(p);
        for (double *p : r.second.d_trace) (void)hipFree(p);
    }
    pool_release(sv);
    void *p[] = {sv->d_step_base, sv->d_table, sv->d_sched, sv->d_runmask};
    for (void *x : p)
        if (x) (void)hipFree(x);
    for (hipEvent_t e : sv->ev)
        if (e) (void)hipEventDestroy(e);
    if (sv->stream) (void)hipStreamDestroy(sv->stream);
    delete sv;
    return DGX_OK;
}

template <typename T>
static hipError_t upload_vec(T **dst, const std::vector<T> &v)
{
    *dst = nullptr;
    hipError_t e = hipMalloc((void **)dst, sizeof(T) * std::max<size_t>(v.size(), 1));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
}

/* doubles to the device in the server's precision */
static hipError_t upload_reals(const dgx_server *sv, void **dst, const std::vector<double> &v)
{
    if (sv->fp32) {
        std::vector<float> f(v.begin(), v.end());
        float *d = nullptr;
        hipError_t e = upload_vec(&d, f);
        *dst = d;
        return e;
    }
    double *d = nullptr;
    hipError_t e = upload_vec(&d, v);
    *dst = d;
    return e;
}

/* Bytes of one material model on the device. */
static long long model_bytes(const dgx_server *sv, const MeshDev &m) { return 3LL * m.tri * (long long)sv->rsize(); }

/* 64-bit FNV-1a over the model's arrays and time step. */
static uint64_t model_hash(const std::vector<double> &mat, double dt)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void *p, size_t n) {
        const unsigned char *c = (const unsigned char *)p;
        for (size_t i = 0; i < n; i++) h = (h ^ c[i]) * 1099511628211ull;
    };
    mix(mat.data(), mat.size() * sizeof(double));
    mix(&dt, sizeof dt);
    return h;
}

extern "C" int dgx_server_add_mesh(dgx_server *sv, const dgx_problem *p, int *mesh_id)
{
    if (!sv || !p || !mesh_id) return DGX_EARG;
    if (p->tri <= 0 || !p->point || !p->triangle || !p->neighbor || !p->lambda || !p->miu || !p->den ||
        !(p->dt > 0)) {
        dgx_set_error("dgx_server_add_mesh: incomplete mesh");
        return DGX_EARG;
    }
    /* the mesh built on the device (single-shot slots: the element layout, 64-element blocks, whose sponge
     * factors are blocked the same way) */
    const long tri_pad = sv->group == 1 ? (p->tri + 63) / 64 * 64 : p->tri;
    dgx_mesh_dev mh;
    int rc = dgx_build_mesh_device(p, sv->group == 1 ? 0 : 1, tri_pad, nullptr, &mh);
    if (rc) return rc;
    std::vector<int4> bt, bl;
    std::vector<int> blv, bg;
    if (sv->group == 1 &&
        !dgx_build_block_tiles(p->tri, mh.nbr, sv->bk, serve_block_threads(sv), bt, blv, bg, bl)) {
        dgx_mesh_dev_free(&mh);
        dgx_set_error("dgx_server_add_mesh: an element and its neighbour levels exceed %d threads",
                      serve_block_threads(sv));
        return DGX_EMESH;
    }
    MeshDev m;
    m.tri = p->tri;
    m.tri_pad = tri_pad;
    m.ntile = (long)bt.size();
    m.inv = mh.inv;
    m.nbr_host.assign(mh.nbr.begin(), mh.nbr.begin() + mh.tri);
    m.nbr = mh.d_nbr;
    m.vtx = mh.d_vtx;
    m.pt = mh.d_pt;
    ModelDev md;
    md.used = true;
    md.dt = p->dt;
    md.refs = 1; /* never released */
    /* a failure frees what this call allocated */
    auto fail = [&](hipError_t e) {
        void *q[] = {m.nbr, m.vtx, m.pt, m.absorb, md.d_mat, m.btiles, m.blnb, m.blevels, m.blrec, mh.d_mat, mh.d_absorb};
        for (void *x : q)
            if (x) (void)hipFree(x);
        dgx_set_error("dgx_server_add_mesh: %s", hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? DGX_ENOMEM : DGX_EHIP;
    };
    /* every model of the mesh is stored 1/rho | lambda | mu with stride tri (the build pads to tri_pad) */
    hipError_t e = dgx_mesh_mat_packed(mh.d_mat, tri_pad, p->tri, sv->fp32, &md.d_mat);
    if (e == hipSuccess) {
        if (sv->fp32) {
            float *m32 = nullptr, *a32 = nullptr;
            e = dgx_mesh_f32(nullptr, 0, mh.d_absorb, 6 * tri_pad, &m32, &a32);
            if (m32) (void)hipFree(m32);
            m.absorb = a32;
            (void)hipFree(mh.d_absorb);
        } else {
            m.absorb = mh.d_absorb;
        }
        mh.d_absorb = nullptr;
    }
    (void)hipFree(mh.d_mat);
    mh.d_mat = nullptr;
    if (e == hipSuccess && sv->group == 1) {
        e = upload_vec(&m.btiles, bt);
        if (e == hipSuccess) e = upload_vec(&m.blnb, bl);
        if (e == hipSuccess) e = upload_vec(&m.blevels, blv);
        if (e == hipSuccess) e = dgx_upload_lane_records(bg.data(), (long)bg.size(), m.vtx, &m.blrec);
    }
    if (e != hipSuccess) return fail(e);
    m.models.push_back(std::move(md));
    sv->meshes.push_back(std::move(m));
    *mesh_id = (int)sv->meshes.size() - 1;
    return DGX_OK;
}

extern "C" int dgx_server_submit(dgx_server *sv, const dgx_request *rq, long *id)
{
    if (!sv || !rq || !id) return DGX_EARG;
    if (rq->mesh < 0 || rq->mesh >= (int)sv->meshes.size() || rq->nshots < 1 || rq->nshots > DGX_SERVE_MAX_SHOTS ||
        rq->nsteps < 1 || rq->nrec < 0 || !rq->src || !rq->wavelet || (rq->nrec > 0 && !rq->rec)) {
        dgx_set_error("dgx_server_submit: invalid request");
        return DGX_EARG;
    }
    const long T = sv->meshes[rq->mesh].tri;
    for (int k = 0; k < rq->nshots; k++)
        if (rq->src[k].elem < 1 || rq->src[k].elem > T ||
            (rq->src[k].kind != DGX_SRC_FORCE_Z && rq->src[k].kind != DGX_SRC_EXPLOSIVE)) {
            dgx_set_error("dgx_server_submit: source %d invalid", k);
            return DGX_EARG;
        }
    for (long k = 0; k < rq->nrec; k++)
        if (rq->rec[k].elem < 1 || rq->rec[k].elem > T) {
            dgx_set_error("dgx_server_submit: receiver %ld invalid", k);
            return DGX_EARG;
        }
    MeshDev &md = sv->meshes[rq->mesh];
    int model = 0;
    if (rq->lambda || rq->miu || rq->den) {
        if (!rq->lambda || !rq->miu || !rq->den || !(rq->dt > 0)) {
            dgx_set_error("dgx_server_submit: a material model needs lambda, miu, den and dt > 0");
            return DGX_EARG;
        }
        /* 1/rho | lambda | mu in device order, as dgx_build_mesh_host lays out the mesh's own model */
        std::vector<double> mat(3 * T);
        for (long e = 0; e < T; e++) {
            if (!(rq->den[e] > 0)) {
                dgx_set_error("dgx_server_submit: element %ld: density must be positive", e + 1);
                return DGX_EARG;
            }
            const long q = md.inv[e];
            mat[q] = 1 / rq->den[e];
            mat[T + q] = rq->lambda[e];
            mat[2 * T + q] = rq->miu[e];
        }
        const uint64_t h = model_hash(mat, rq->dt);
        model = -1;
        for (int k = 1; k < (int)md.models.size() && model < 0; k++)
            if (md.models[k].used && md.models[k].hash == h && md.models[k].dt == rq->dt && md.models[k].mat == mat)
                model = k;
        if (model < 0) {
            for (int k = 1; k < (int)md.models.size() && model < 0; k++)
                if (!md.models[k].used) model = k;
            if (model < 0) {
                model = (int)md.models.size();
                md.models.emplace_back();
            }
            ModelDev &nm = md.models[model];
            nm = ModelDev();
            nm.used = true;
            nm.hash = h;
            nm.mat = std::move(mat);
            nm.dt = rq->dt;
        }
    }
    Request r;
    r.id = sv->next_id++;
    r.mesh = rq->mesh;
    r.model = model;
    r.nshots = rq->nshots;
    r.nsteps = rq->nsteps;
    r.nrec = rq->nrec;
    r.src.assign(rq->src, rq->src + rq->nshots);
    r.wav.assign(rq->wavelet, rq->wavelet + (size_t)rq->nshots * rq->nsteps);
    if (rq->nrec) r.rec.assign(rq->rec, rq->rec + rq->nrec);
    r.arrival = rq->arrival;
    r.bytes = (long long)rq->nshots * rq->nsteps * (1 + 2 * rq->nrec) * 8;
    const long long alone = r.bytes + groups_for(sv, rq->nshots) * group_bytes(sv, md) +
                            (model ? model_bytes(sv, md) : 0);
    if (alone > sv->budget) {
        dgx_set_error("dgx_server_submit: the request needs %lld bytes with its slot groups, more than the %lld-byte "
                      "budget", alone, sv->budget);
        sv->next_id--;
        if (model && md.models[model].refs == 0) md.models[model] = ModelDev(); /* created for it: drop */
        return DGX_ENOMEM;
    }
    md.models[model].refs++;
    *id = r.id;
    sv->queue.push_back(r.id);
    sv->reqs.emplace(r.id, std::move(r));
    return DGX_OK;
}

/* Free every slot group that has no request, so that its memory can go to a request for any mesh.
 * Called between blocks of steps (the stream is idle); returns the number of groups freed. */
static int reclaim_empty(dgx_server *sv)
{
    int n = 0;
    for (int bi = 0; bi < (int)sv->batches.size(); bi++) {
        Batch &b = sv->batches[bi];
        if (!b.live || b.used() != 0) continue;
        MeshDev &m = sv->meshes[b.mesh];
        void *p[] = {b.fields, b.d_slot_int, b.d_slot_ptr, b.d_aux_of_elem, b.d_aux, b.d_src, b.d_rec, b.d_pairs};
        for (void *x : p)
            if (x) (void)hipFree(x);
        sv->in_use -= group_bytes(sv, m);
        m.batches.erase(std::remove(m.batches.begin(), m.batches.end(), bi), m.batches.end());
        Batch dead;
        dead.live = false;
        sv->batches[bi] = std::move(dead);
        n++;
    }
    if (n) sv->table_dirty = true;
    return n;
}

static int new_batch(dgx_server *sv, int mesh, int model, int *bi)
{
    const MeshDev &m = sv->meshes[mesh];
    int slot = -1;
    for (int k = 0; k < (int)sv->batches.size() && slot < 0; k++)
        if (!sv->batches[k].live) slot = k;
    if (slot < 0 && (long)sv->batches.size() >= MAX_BATCHES) return DGX_ENOMEM;
    Batch b;
    b.mesh = mesh;
    b.model = model;
    b.G = sv->group;
    const int G = b.G;
    b.field_bytes = (long long)(G == 1 ? 2 : 1) * 5 * group_rows(sv, m) * G * sv->rsize();
    for (int k = 0; k < MAXG; k++) {
        b.req[k] = -1;
        b.shot[k] = 0;
        b.h_start[k] = 0;
        b.h_len[k] = 0;
        b.h_nrec[k] = 0;
        b.h_wav[k] = nullptr;
        b.h_trace[k] = nullptr;
    }
    hipError_t e = hipMalloc(&b.fields, b.field_bytes);
    if (e == hipSuccess) e = hipMemsetAsync(b.fields, 0, b.field_bytes, sv->stream);
    if (e == hipSuccess) e = hipMalloc((void **)&b.d_slot_int, sizeof(int) * 3 * G);
    if (e == hipSuccess) e = hipMalloc((void **)&b.d_slot_ptr, sizeof(double *) * 2 * G);
    if (e == hipSuccess) e = hipMalloc((void **)&b.d_aux_of_elem, sizeof(int) * m.tri);
    if (e == hipSuccess) e = hipMemsetAsync(b.d_aux_of_elem, 0, sizeof(int) * m.tri, sv->stream); /* no aux, no source */
    if (e != hipSuccess) {
        /* nothing of the half-built group is kept */
        (void)hipStreamSynchronize(sv->stream);
        void *p[] = {b.fields, b.d_slot_int, b.d_slot_ptr, b.d_aux_of_elem};
        for (void *x : p)
            if (x) (void)hipFree(x);
        dgx_set_error("slot group allocation failed: %s", hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? DGX_ENOMEM : DGX_EHIP;
    }
    sv->in_use += group_bytes(sv, m);
    if (slot >= 0) {
        sv->batches[slot] = std::move(b);
        *bi = slot;
    } else {
        sv->batches.push_back(std::move(b));
        *bi = (int)sv->batches.size() - 1;
    }
    sv->meshes[mesh].batches.push_back(*bi);
    sv->table_dirty = true;
    return DGX_OK;
}

/* Rebuild a batch's source/receiver tables and per-slot parameters after admissions/retirements. */
static int refresh_batch(dgx_server *sv, Batch &b)
{
    const MeshDev &m = sv->meshes[b.mesh];
    std::map<int, std::pair<std::vector<dgx_src_dev>, std::vector<dgx_rec_dev>>> by_elem;
    const int G = b.G;
    for (int k = 0; k < G; k++) {
        if (b.req[k] < 0) continue;
        const Request &r = sv->reqs.at(b.req[k]);
        const dgx_source &sr = r.src[b.shot[k]];
        dgx_src_dev d;
        d.elem = (int)m.inv[sr.elem - 1];
        d.kind = sr.kind;
        d.shot = k;
        d.wavelet = k;
        for (int i = 0; i < 6; i++) d.fun[i] = sr.fun[i];
        by_elem[d.elem].first.push_back(d);
        for (long j = 0; j < r.nrec; j++) {
            dgx_rec_dev rr;
            rr.elem = (int)m.inv[r.rec[j].elem - 1];
            rr.index = (int)j;
            rr.pad0 = k;
            rr.pad1 = 0;
            for (int i = 0; i < 6; i++) rr.w[i] = r.rec[j].weight[i];
            by_elem[rr.elem].second.push_back(rr);
        }
    }
    std::vector<dgx_aux> aux;
    std::vector<dgx_src_dev> src;
    std::vector<dgx_rec_dev> rec;
    std::vector<int> elems;
    for (auto &kv : by_elem) {
        dgx_aux x;
        x.src_begin = (int)src.size();
        x.src_count = (int)kv.second.first.size();
        x.rec_begin = (int)rec.size();
        x.rec_count = (int)kv.second.second.size();
        src.insert(src.end(), kv.second.first.begin(), kv.second.first.end());
        rec.insert(rec.end(), kv.second.second.begin(), kv.second.second.end());
        aux.push_back(x);
        elems.push_back(kv.first);
    }
    /* per-element words (aux record + 1, neighbour-source bits); every element whose word changes gets
     * exactly one pair, so the scatter has no two writes to one address */
    std::map<int, unsigned> word;
    for (size_t x = 0; x < elems.size(); x++) word[elems[x]] |= (unsigned)(x + 1);
    for (const dgx_src_dev &d : src) {
        if (d.kind != DGX_SRC_EXPLOSIVE) continue;
        const int4 ne = m.nbr_host[d.elem];
        for (int n : {ne.x, ne.y, ne.z}) {
            if (n < 0) continue;
            const int4 nn = m.nbr_host[n];
            const int back[3] = {nn.x, nn.y, nn.z};
            for (int j = 0; j < 3; j++)
                if (back[j] == d.elem) word[n] |= 1u << (DGX_DYN_NB_SHIFT + j);
        }
    }
    std::map<int, int> fin;
    for (int q : b.aux_elems) fin[q] = 0;
    for (auto &kv : word) fin[kv.first] = (int)kv.second;
    std::vector<int2> pairs;
    for (auto &kv : fin) pairs.push_back(make_int2(kv.first, kv.second));
    long old_a = b.aux_cap, old_s = b.src_cap, old_r = b.rec_cap;
    SV_TRY(grow(&b.d_aux, &b.aux_cap, (long)aux.size()));
    SV_TRY(grow(&b.d_src, &b.src_cap, (long)src.size()));
    SV_TRY(grow(&b.d_rec, &b.rec_cap, (long)rec.size()));
    SV_TRY(grow(&b.d_pairs, &b.pair_cap, (long)pairs.size()));
    if (old_a != b.aux_cap || old_s != b.src_cap || old_r != b.rec_cap) sv->table_dirty = true;
    if (!aux.empty()) SV_TRY(hipMemcpyAsync(b.d_aux, aux.data(), sizeof(dgx_aux) * aux.size(), hipMemcpyHostToDevice, sv->stream));
    if (!src.empty()) SV_TRY(hipMemcpyAsync(b.d_src, src.data(), sizeof(dgx_src_dev) * src.size(), hipMemcpyHostToDevice, sv->stream));
    if (!rec.empty()) SV_TRY(hipMemcpyAsync(b.d_rec, rec.data(), sizeof(dgx_rec_dev) * rec.size(), hipMemcpyHostToDevice, sv->stream));
    if (!pairs.empty()) {
        SV_TRY(hipMemcpyAsync(b.d_pairs, pairs.data(), sizeof(int2) * pairs.size(), hipMemcpyHostToDevice, sv->stream));
        SV_TRY(dgx_launch_scatter(b.d_aux_of_elem, b.d_pairs, (int)pairs.size(), sv->stream));
    }
    b.aux_elems.clear();
    for (auto &kv : word) b.aux_elems.push_back(kv.first);
    int si[3 * MAXG];
    double *sp[2 * MAXG];
    for (int k = 0; k < G; k++) {
        si[k] = b.h_start[k];
        si[G + k] = b.h_len[k];
        si[2 * G + k] = b.h_nrec[k];
        sp[k] = b.h_wav[k];
        sp[G + k] = b.h_trace[k];
    }
    SV_TRY(hipMemcpyAsync(b.d_slot_int, si, sizeof(int) * 3 * G, hipMemcpyHostToDevice, sv->stream));
    SV_TRY(hipMemcpyAsync(b.d_slot_ptr, sp, sizeof(double *) * 2 * G, hipMemcpyHostToDevice, sv->stream));
    /* the staging copies above read host memory asynchronously: wait before it goes out of scope */
    SV_TRY(hipStreamSynchronize(sv->stream));
    b.dirty = false;
    return DGX_OK;
}

template <typename R>
static void fill_batch_args(const dgx_server *sv, const Batch &b, const dgx::Ops<R> &op, dgx::StepArgs<R> &a)
{
    const MeshDev &m = sv->meshes[b.mesh];
    std::memset(&a, 0, sizeof(a));
    a.op = op;
    R *f = (R *)b.fields;
    const int G = b.G;
    size_t per = (size_t)group_rows(sv, m) * G;
    for (int c = 0; c < 5; c++) a.f[c] = f + c * per;
    a.fstride = m.tri;
    if (G == 1) { /* blocked kernel: copy A, copy B (the kernel alternates them by launch), the mesh's tiling */
        for (int c = 0; c < 5; c++) a.g[c] = f + (5 + c) * per;
        a.fstride = 0;
        a.tiles = m.btiles;
        a.blevels = m.blevels;
        a.blrec = m.blrec;
        a.blnb = m.blnb;
        a.ntile = (int)m.ntile;
        a.bkmax = sv->bk;
        a.bsteps = sv->bk;
        a.bthreads = serve_block_threads(sv);
    }
    a.nbr = m.nbr;
    a.vtx = m.vtx;
    a.pt = m.pt;
    const ModelDev &md = m.models[b.model];
    const R *mat = (const R *)md.d_mat;
    a.rho_inv = mat;
    a.lam = mat + m.tri;
    a.mu = mat + 2 * m.tri;
    a.absorb = (const R *)m.absorb;
    a.aux = b.d_aux;
    a.src = b.d_src;
    a.rec = b.d_rec;
    a.nsteps = 0x7fffffff;
    a.tri = (int)m.tri;
    a.dt = md.dt;
    a.spad = G;
    a.shot_tile = G;
    a.nshots = G;
    a.shot_start = b.d_slot_int;
    a.shot_len = b.d_slot_int + G;
    a.slot_nrec = b.d_slot_int + 2 * G;
    a.slot_wav = (const double *const *)b.d_slot_ptr;
    a.slot_trace = b.d_slot_ptr + G;
    a.aux_of_elem = b.d_aux_of_elem;
}

static int upload_table(dgx_server *sv)
{
    std::vector<int> sched(2);
    std::vector<char> table;
    int nb = 0, tiles = 0;
    size_t asz = sv->fp32 ? sizeof(dgx::StepArgs<float>) : sizeof(dgx::StepArgs<double>);
    sv->order.clear();
    for (size_t bi = 0; bi < sv->batches.size(); bi++) {
        const Batch &b = sv->batches[bi];
        if (!b.live || b.used() == 0) continue;
        sv->order.push_back((int)bi);
        table.resize((nb + 1) * asz);
        if (sv->fp32) fill_batch_args(sv, b, sv->op32, *(dgx::StepArgs<float> *)(table.data() + nb * asz));
        else fill_batch_args(sv, b, sv->op64, *(dgx::StepArgs<double> *)(table.data() + nb * asz));
        sched.push_back(tiles);
        if (sv->group == 1) {
            tiles += (int)sv->meshes[b.mesh].ntile; /* the blocked kernel's tiles of the mesh */
        } else {
            const int per_tile = 256 / sv->group; /* a workgroup covers 256/G elements x G shots */
            tiles += (int)((sv->meshes[b.mesh].tri + per_tile - 1) / per_tile);
        }
        nb++;
    }
    sched.push_back(tiles);
    sched[0] = nb;
    sched[1] = tiles;
    if ((nb != sv->nb || tiles != sv->tiles) && sv->graph) {
        /* the kernels' grid and batch count are part of the captured graph */
        (void)hipGraphExecDestroy(sv->graph);
        sv->graph = nullptr;
    }
    sv->nb = nb;
    sv->tiles = tiles;
    if (nb) SV_TRY(hipMemcpyAsync(sv->d_table, table.data(), table.size(), hipMemcpyHostToDevice, sv->stream));
    SV_TRY(hipMemcpyAsync(sv->d_sched, sched.data(), sizeof(int) * sched.size(), hipMemcpyHostToDevice, sv->stream));
    SV_TRY(hipStreamSynchronize(sv->stream));
    sv->table_dirty = false;
    return DGX_OK;
}

/* Batch entries free for new groups (reused entries, and the table's room). */
static long batch_entries_free(const dgx_server *sv)
{
    long n = MAX_BATCHES - (long)sv->batches.size();
    for (const auto &b : sv->batches)
        if (!b.live) n++;
    return n;
}

/*
 * Admit queued requests in arrival order (later ones may pass one that does not fit yet).  A request's
 * shots are independent, so they may spread over several groups of its mesh: free slots are taken from
 * the fullest groups first, and new groups are opened only for the shots left over.  Groups stay full,
 * instead of one partly used group per request.
 */
static int admit(dgx_server *sv)
{
    std::deque<long> still;
    const int G = sv->group;
    while (!sv->queue.empty()) {
        long id = sv->queue.front();
        sv->queue.pop_front();
        Request &r = sv->reqs.at(id);
        MeshDev &m = sv->meshes[r.mesh];
        ModelDev &md = m.models[r.model];
        /* this mesh's groups of the request's material model with free slots, fullest first */
        std::vector<std::pair<int, int>> open; /* (free slots, batch) */
        int free_total = 0;
        for (int bi : m.batches) {
            /* an empty group of another model of the mesh is memory for any of them */
            if (sv->batches[bi].model != r.model && sv->batches[bi].used() > 0) continue;
            const int fr = G - sv->batches[bi].used();
            if (fr > 0) {
                open.push_back({fr, bi});
                free_total += fr;
            }
        }
        std::stable_sort(open.begin(), open.end(),
                         [](const std::pair<int, int> &x, const std::pair<int, int> &y) { return x.first < y.first; });
        const int grow = groups_for(sv, std::max(0, r.nshots - free_total)); /* new groups needed */
        const long long mbytes = md.d_mat ? 0 : model_bytes(sv, m); /* the model comes to the device with it */
        long long need = r.bytes + grow * group_bytes(sv, m) + mbytes;
        bool fits = sv->in_use + need <= sv->budget && batch_entries_free(sv) >= grow;
        if (!fits && grow > 0 && reclaim_empty(sv) > 0) {
            /* memory and entries held by idle groups (of any mesh) went back: look again */
            sv->queue.push_front(id);
            continue;
        }
        if (!fits) { still.push_back(id); continue; }
        /* nothing of this request is placed until every step that can fail has succeeded; on a failure it and
         * the requests deferred or not yet looked at stay queued, in order, and its buffers go back to the pool
         * (a group opened for it stays empty until reclaimed) */
        auto requeue = [&](int rc) {
            still.push_back(id);
            for (long rest : sv->queue) still.push_back(rest);
            sv->queue = still;
            return rc;
        };
        if (!md.d_mat) {
            void *d = nullptr;
            if (upload_reals(sv, &d, md.mat) != hipSuccess) {
                if (d) (void)hipFree(d);
                (void)hipGetLastError();
                dgx_set_error("upload of the material model of request %ld failed", id);
                return requeue(DGX_ENOMEM);
            }
            md.d_mat = d;
            sv->in_use += mbytes;
        }
        for (int g = 0; g < grow; g++) {
            int bi = -1;
            int rc = new_batch(sv, r.mesh, r.model, &bi);
            if (rc) return requeue(rc);
            open.push_back({G, bi});
        }
        /* every shot's wavelet and trace buffer, and the wavelet uploads */
        const size_t nw = (size_t)r.nsteps, nt = (size_t)r.nsteps * 2 * std::max<long>(r.nrec, 1);
        std::vector<double *> ws, ts;
        int rc = DGX_OK;
        for (int k = 0; k < r.nshots && rc == DGX_OK; k++) {
            double *w = pool_get(sv, nw);
            if (w) ws.push_back(w);
            double *t = w ? pool_get(sv, nt) : nullptr;
            if (t) ts.push_back(t);
            if (!w || !t) {
                dgx_set_error("out of device memory for the buffers of request %ld", id);
                rc = DGX_ENOMEM;
            } else if (hipMemcpyAsync(w, &r.wav[(size_t)k * nw], sizeof(double) * nw, hipMemcpyHostToDevice,
                                      sv->stream) != hipSuccess) {
                dgx_set_error("wavelet upload of request %ld failed", id);
                rc = DGX_EHIP;
            }
        }
        if (rc != DGX_OK) {
            (void)hipStreamSynchronize(sv->stream);
            for (double *w : ws) pool_put(sv, w, nw);
            for (double *t : ts) pool_put(sv, t, nt);
            return requeue(rc);
        }
        int k = 0;
        for (const auto &ob : open) {
            if (k == r.nshots) break;
            Batch &b = sv->batches[ob.second];
            b.model = r.model;
            unsigned long long mask = 0;
            for (int slot = 0; slot < G && k < r.nshots; slot++) {
                if (b.req[slot] >= 0) continue;
                double *w = ws[k], *t = ts[k];
                b.req[slot] = id;
                b.shot[slot] = k;
                b.h_start[slot] = (int)sv->global_step;
                b.h_len[slot] = (int)r.nsteps;
                b.h_nrec[slot] = (int)r.nrec;
                b.h_wav[slot] = w;
                b.h_trace[slot] = t;
                r.slots.push_back({ob.second, slot});
                r.d_wav.push_back(w);
                r.d_trace.push_back(t);
                mask |= 1ull << slot;
                k++;
            }
            if (!mask) continue;
            void *f[5];
            size_t per = (size_t)group_rows(sv, m) * G * sv->rsize();
            for (int c = 0; c < 5; c++) f[c] = (char *)b.fields + c * per;
            SV_TRY(dgx_launch_clear_slots(sv->fp32, f, group_rows(sv, m), G, mask, sv->stream));
            b.dirty = true;
        }
        sv->table_dirty = true;
        md.live++;
        r.state = DGX_REQ_RUNNING;
        r.start = sv->global_step;
        r.t_admit = now_s();
        sv->in_use += r.bytes;
    }
    sv->queue = still;
    return DGX_OK;
}

static int retire(dgx_server *sv, int *completed)
{
    for (auto &kv : sv->reqs) {
        Request &r = kv.second;
        if (r.state != DGX_REQ_RUNNING || sv->global_step < r.start + r.nsteps) continue;
        const long n = r.nsteps, nr = r.nrec;
        r.res_x.assign((size_t)r.nshots * n * nr, 0.0);
        r.res_z.assign((size_t)r.nshots * n * nr, 0.0);
        std::vector<double> buf((size_t)n * 2 * std::max<long>(nr, 1));
        for (int k = 0; k < r.nshots; k++) {
            if (nr > 0) {
                SV_TRY(hipMemcpy(buf.data(), r.d_trace[k], sizeof(double) * n * 2 * nr, hipMemcpyDeviceToHost));
                for (long t = 0; t < n; t++)
                    for (long j = 0; j < nr; j++) {
                        r.res_x[((size_t)k * n + t) * nr + j] = buf[(t * 2 + 0) * nr + j];
                        r.res_z[((size_t)k * n + t) * nr + j] = buf[(t * 2 + 1) * nr + j];
                    }
            }
            Batch &b = sv->batches[r.slots[k].first];
            int slot = r.slots[k].second;
            b.req[slot] = -1;
            b.h_len[slot] = 0;
            b.h_wav[slot] = nullptr;
            b.h_trace[slot] = nullptr;
            b.dirty = true;
            pool_put(sv, r.d_wav[k], n);
            pool_put(sv, r.d_trace[k], (size_t)n * 2 * std::max<long>(nr, 1));
        }
        r.d_wav.clear();
        r.d_trace.clear();
        MeshDev &m = sv->meshes[r.mesh];
        ModelDev &md = m.models[r.model];
        md.live--;
        if (r.model && md.live == 0 && md.d_mat) {
            /* its groups hold no shot now (reclaimed when memory is needed); none is stepped again before a
             * request of this model re-uploads it */
            SV_TRY(hipStreamSynchronize(sv->stream));
            SV_TRY(hipFree(md.d_mat));
            md.d_mat = nullptr;
            sv->in_use -= model_bytes(sv, m);
        }
        if (--md.refs == 0 && r.model) md = ModelDev();
        r.state = DGX_REQ_DONE;
        r.t_done = now_s();
        sv->in_use -= r.bytes;
        sv->table_dirty = true;
        (*completed)++;
    }
    return DGX_OK;
}

/*
 * Compaction.  When the running shots of a mesh fit in fewer of its groups than now hold them (requests
 * finished and nothing queued filled their slots), the shots of the emptiest groups move into the free
 * slots of the fullest.  A move copies the shot's field lanes bit for bit and carries its step range,
 * wavelet and trace buffer along, so the shot continues exactly as it would have in place; the emptied
 * groups are no longer stepped.  Runs between blocks, after admission.
 */
static int compact(dgx_server *sv)
{
    const int G = sv->group;
    for (int mi = 0; mi < (int)sv->meshes.size(); mi++)
    for (int mo = 0; mo < (int)sv->meshes[mi].models.size(); mo++) {
        const MeshDev &m = sv->meshes[mi];
        std::vector<std::pair<int, int>> act; /* (used slots, batch) of this model's groups that are stepped */
        int used = 0;
        for (int bi : m.batches) {
            if (sv->batches[bi].model != mo) continue;
            const int u = sv->batches[bi].used();
            if (u > 0) {
                act.push_back({u, bi});
                used += u;
            }
        }
        const int keep = (used + G - 1) / G;
        if ((int)act.size() <= keep) continue;
        std::stable_sort(act.begin(), act.end(),
                         [](const std::pair<int, int> &x, const std::pair<int, int> &y) { return x.first > y.first; });
        size_t t = 0; /* target group (act[0 .. keep-1]) and its next slot to look at */
        int ts = 0;
        const size_t per = (size_t)group_rows(sv, m) * G * sv->rsize();
        for (size_t d = keep; d < act.size(); d++) {
            const int dbi = act[d].second;
            int s = 0;
            while (s < G) {
                /* one launch per (donor, target) pair */
                while (t < (size_t)keep && ts >= G) { t++; ts = 0; }
                if (t >= (size_t)keep) break; /* cannot happen: the targets have room for every donor shot */
                const int tbi = act[t].second;
                Batch &D = sv->batches[dbi];
                Batch &T = sv->batches[tbi];
                int from[MAXG], to[MAXG], n = 0;
                for (; s < G; s++) {
                    if (D.req[s] < 0) continue;
                    while (ts < G && T.req[ts] >= 0) ts++;
                    if (ts >= G) break;
                    from[n] = s;
                    to[n] = ts;
                    n++;
                    T.req[ts] = D.req[s];
                    T.shot[ts] = D.shot[s];
                    T.h_start[ts] = D.h_start[s];
                    T.h_len[ts] = D.h_len[s];
                    T.h_nrec[ts] = D.h_nrec[s];
                    T.h_wav[ts] = D.h_wav[s];
                    T.h_trace[ts] = D.h_trace[s];
                    for (auto &sl : sv->reqs.at(D.req[s]).slots)
                        if (sl.first == dbi && sl.second == s) sl = {tbi, ts};
                    D.req[s] = -1;
                    D.h_len[s] = 0;
                    D.h_wav[s] = nullptr;
                    D.h_trace[s] = nullptr;
                }
                if (!n) continue;
                void *src[5], *dst[5];
                for (int c = 0; c < 5; c++) {
                    src[c] = (char *)D.fields + c * per;
                    dst[c] = (char *)T.fields + c * per;
                }
                SV_TRY(dgx_launch_move_slots(sv->fp32, src, dst, group_rows(sv, m), G, n, from, to, sv->stream));
                D.dirty = true;
                T.dirty = true;
                sv->table_dirty = true;
                sv->moved += n;
            }
        }
    }
    return DGX_OK;
}

/* The slots of every stepped batch that run at each step of the next block (bit s: slot s), read by the
 * serving kernels through the scalar cache instead of a per-lane load of each slot's step window. */
static int upload_runmask(dgx_server *sv)
{
    const int nb = (int)sv->order.size();
    if (!nb) return DGX_OK;
    std::vector<unsigned long long> m((size_t)sv->K * nb, 0ull);
    for (int i = 0; i < nb; i++) {
        const Batch &b = sv->batches[sv->order[i]];
        for (int s = 0; s < b.G; s++) {
            if (b.req[s] < 0 || b.h_len[s] <= 0) continue;
            const long lo = std::max<long>(b.h_start[s], sv->global_step);
            const long hi = std::min<long>((long)b.h_start[s] + b.h_len[s], sv->global_step + sv->K);
            for (long g = lo; g < hi; g++) m[(size_t)(g - sv->global_step) * nb + i] |= 1ull << s;
        }
    }
    SV_TRY(hipMemcpyAsync(sv->d_runmask, m.data(), sizeof(unsigned long long) * m.size(), hipMemcpyHostToDevice,
                          sv->stream));
    /* the copy reads host memory asynchronously: wait before it goes out of scope */
    SV_TRY(hipStreamSynchronize(sv->stream));
    return DGX_OK;
}

static int build_graph(dgx_server *sv)
{
    hipGraph_t g;
    SV_TRY(hipStreamBeginCapture(sv->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    const void *ops = sv->fp32 ? (const void *)&sv->op32 : (const void *)&sv->op64;
    /* single-shot slots: K / bk launches of the blocked kernel (an even number: every slot's
     * state is back in copy A at the end of the block) */
    for (int k = 0; sv->group == 1 && k < sv->K / sv->bk && e == hipSuccess; k++)
        e = dgx_launch_serve_block(sv->fp32, sv->p2, sv->d_table, sv->d_sched, sv->d_step_base, k * sv->bk, k,
                                   sv->nb, sv->tiles, sv->stream);
    for (int k = 0; sv->group != 1 && k < sv->K && e == hipSuccess; k++) {
        e = dgx_launch_serve(sv->fp32, ops, sv->d_table, sv->d_sched, sv->d_runmask, sv->d_step_base, k, 1,
                             sv->group, sv->nb, sv->tiles, sv->stream);
        if (e == hipSuccess)
            e = dgx_launch_serve(sv->fp32, ops, sv->d_table, sv->d_sched, sv->d_runmask, sv->d_step_base, k, 0,
                                 sv->group, sv->nb, sv->tiles, sv->stream);
    }
    if (e == hipSuccess) e = dgx_launch_advance(sv->d_step_base, sv->K, sv->stream);
    hipError_t e2 = hipStreamEndCapture(sv->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        dgx_set_error("serving graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
        return DGX_EHIP;
    }
    e = hipGraphInstantiate(&sv->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { dgx_set_error("graph instantiate failed: %s", hipGetErrorString(e)); return DGX_EHIP; }
    return DGX_OK;
}

extern "C" int dgx_server_advance(dgx_server *sv, int nblocks)
{
    if (!sv || nblocks < 0) return DGX_EARG;
    double t0 = now_s();
    int completed = 0;
    for (int it = 0; it < nblocks; it++) {
        int rc = admit(sv);
        if (rc) return rc;
        rc = compact(sv);
        if (rc) return rc;
        for (auto &b : sv->batches)
            if (b.live && b.dirty) {
                rc = refresh_batch(sv, b);
                if (rc) return rc;
            }
        if (sv->table_dirty) {
            rc = upload_table(sv);
            if (rc) return rc;
        }
        long running = 0;
        long long dof = 0;
        for (auto &kv : sv->reqs)
            if (kv.second.state == DGX_REQ_RUNNING) {
                const Request &r = kv.second;
                running++;
                long left = r.start + r.nsteps - sv->global_step;
                dof += (long long)r.nshots * sv->meshes[r.mesh].tri * 30 * std::min<long>(left, sv->K);
            }
        if (!running) break;
        long long lanes = 0; /* groups with a running slot are stepped (upload_table) */
        for (const auto &b : sv->batches)
            if (b.live && b.used() > 0) lanes += (long long)b.G * sv->meshes[b.mesh].tri * 30 * sv->K;
        rc = upload_runmask(sv);
        if (rc) return rc;
        if (!sv->graph) {
            rc = build_graph(sv);
            if (rc) return rc;
        }
        SV_TRY(hipEventRecord(sv->ev[0], sv->stream));
        SV_TRY(hipGraphLaunch(sv->graph, sv->stream));
        SV_TRY(hipEventRecord(sv->ev[1], sv->stream));
        SV_TRY(hipStreamSynchronize(sv->stream));
        float ms = 0;
        SV_TRY(hipEventElapsedTime(&ms, sv->ev[0], sv->ev[1]));
        sv->gpu += ms * 1e-3;
        sv->global_step += sv->K;
        sv->dof_updates += dof;
        sv->lane_updates += lanes;
        sv->blocks++;
        rc = retire(sv, &completed);
        if (rc) return rc;
    }
    sv->busy += now_s() - t0;
    return completed;
}

extern "C" int dgx_server_status(const dgx_server *sv, long id, int *state, double *t_admit, double *t_done)
{
    if (!sv) return DGX_EARG;
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return DGX_EARG;
    if (state) *state = it->second.state;
    if (t_admit) *t_admit = it->second.t_admit;
    if (t_done) *t_done = it->second.t_done;
    return DGX_OK;
}

extern "C" int dgx_server_result(dgx_server *sv, long id, double *rec_x, double *rec_z)
{
    if (!sv) return DGX_EARG;
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end() || it->second.state != DGX_REQ_DONE) {
        dgx_set_error("request %ld is not finished", id);
        return DGX_ESTATE;
    }
    const Request &r = it->second;
    if (rec_x && !r.res_x.empty()) std::memcpy(rec_x, r.res_x.data(), sizeof(double) * r.res_x.size());
    if (rec_z && !r.res_z.empty()) std::memcpy(rec_z, r.res_z.data(), sizeof(double) * r.res_z.size());
    return DGX_OK;
}

extern "C" int dgx_server_release(dgx_server *sv, long id)
{
    if (!sv) return DGX_EARG;
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end() || it->second.state != DGX_REQ_DONE) return DGX_ESTATE;
    sv->reqs.erase(it);
    return DGX_OK;
}

extern "C" int dgx_server_get_stats(const dgx_server *sv, dgx_server_stats *st)
{
    if (!sv || !st) return DGX_EARG;
    std::memset(st, 0, sizeof(*st));
    for (auto &kv : sv->reqs) {
        if (kv.second.state == DGX_REQ_QUEUED) st->queued++;
        else if (kv.second.state == DGX_REQ_RUNNING) st->running++;
        else st->done++;
    }
    for (const auto &b : sv->batches) st->batches += b.live ? 1 : 0;
    st->bytes_in_use = sv->in_use;
    st->dof_updates = sv->dof_updates;
    st->busy_seconds = sv->busy;
    st->blocks = sv->blocks;
    st->global_step = sv->global_step;
    st->lane_updates = sv->lane_updates;
    st->moved = sv->moved;
    st->gpu_seconds = sv->gpu;
    return DGX_OK;
}
