/*
 * dgx_meshdev.hip -- mesh set-up on the device: element order, face codes, vertex renumbering, materials and
 * sponge factors (what dgx_build_mesh_host computes, for dgx_create and dgx_server_add_mesh), and the point
 * locator of sources and receivers (what dgx_locate_source / dgx_locate_receivers compute).
 *
 * Both produce the host versions' results bit for bit: the same keys, comparisons and floating-point
 * expressions, evaluated without contraction into fused multiply-adds (the host code is compiled without
 * them), and reductions whose result does not depend on the order (min, max, first / last index).  The
 * reference builds these arrays on the host in serial loops: Neighbor_Edge (fun_DG.c:557-739), NeighborNodes
 * (fun_DG.c:1195-1276), Elastic_A_B (fun_DG.c:1071-1088), Get_Source / Get_Receiver (fun_DG.c:284-418).
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "dgx.h"
#include "dgx_internal.h"
#include "dgx_kernels.h"
#include "dgx_setup.h"

#pragma clang fp contract(off)

extern "C" int dgx_source_at(const double *point, const long *triangle, const dgx_operators *ops, long e,
                             long vertex, double x, double z, int kind, int shot, long wavelet, dgx_source *out);
extern "C" int dgx_receiver_at(const double *point, const long *triangle, long e, long vertex, double x, double z,
                               dgx_receiver *out);

namespace {

constexpr int kThreads = 256;
constexpr unsigned long long kNone = ULLONG_MAX;

unsigned blocks_for(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

__device__ __forceinline__ long long fidx_d(long long q, int i) { return ((q >> 6) * 6 + i) * 64 + (q & 63); }

/* Hilbert index on a 2^16 grid: the host's hilbert() in dgx_api.cpp, step for step. */
__device__ uint32_t hilbert_d(double u, double v)
{
    const uint32_t n = 1u << 16;
    uint32_t x = (uint32_t)fmin(65535.0, fmax(0.0, u * 65536.0));
    uint32_t y = (uint32_t)fmin(65535.0, fmax(0.0, v * 65536.0));
    uint32_t d = 0;
    for (uint32_t s = n / 2; s > 0; s /= 2) {
        const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - (x & (s - 1)) + (x & ~(s - 1));
                y = s - 1 - (y & (s - 1)) + (y & ~(s - 1));
            }
            x &= (s - 1);
            y &= (s - 1);
            const uint32_t t = x;
            x = y;
            y = t;
        } else {
            x &= (s - 1);
            y &= (s - 1);
        }
    }
    return d;
}

/* Bounding box of the points: per-block min/max, then one block over the partials. bb = {xmin, xmax, zmin, zmax}. */
__global__ void k_bbox_part(const double *__restrict__ P, long long n, double *__restrict__ part)
{
    __shared__ double sh[4][kThreads];
    double a = INFINITY, b = -INFINITY, c = INFINITY, d = -INFINITY;
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
        const double x = P[2 * i], z = P[2 * i + 1];
        a = fmin(a, x); b = fmax(b, x); c = fmin(c, z); d = fmax(d, z);
    }
    sh[0][threadIdx.x] = a; sh[1][threadIdx.x] = b; sh[2][threadIdx.x] = c; sh[3][threadIdx.x] = d;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w /= 2) {
        if ((int)threadIdx.x < w) {
            sh[0][threadIdx.x] = fmin(sh[0][threadIdx.x], sh[0][threadIdx.x + w]);
            sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + w]);
            sh[2][threadIdx.x] = fmin(sh[2][threadIdx.x], sh[2][threadIdx.x + w]);
            sh[3][threadIdx.x] = fmax(sh[3][threadIdx.x], sh[3][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) part[blockIdx.x * 4 + k] = sh[k][0];
}

__global__ void k_bbox_final(const double *__restrict__ part, int nparts, double *__restrict__ bb)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = part[0], b = part[1], c = part[2], d = part[3];
    for (int k = 1; k < nparts; k++) {
        a = fmin(a, part[4 * k]); b = fmax(b, part[4 * k + 1]);
        c = fmin(c, part[4 * k + 2]); d = fmax(d, part[4 * k + 3]);
    }
    bb[0] = a; bb[1] = b; bb[2] = c; bb[3] = d;
}

/* Sort keys (Hilbert index of the centroid << 32 | element) and the checks of dgx_build_mesh_host's first loop.
 * err: min over failing elements of 2 e + (0: vertex id out of range, 1: not counter-clockwise). */
__global__ void k_keys(const double *__restrict__ P, const long *__restrict__ Tv, long long T, long long np,
                       const double *__restrict__ bb, uint64_t *__restrict__ key, unsigned long long *err)
{
    const long long e = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (e >= T) return;
    for (int k = 0; k < 3; k++) {
        const long v = Tv[3 * e + k];
        if (v < 1 || v > np) { atomicMin(err, (unsigned long long)(2 * e)); return; }
    }
    const double *a = &P[2 * (Tv[3 * e] - 1)], *b = &P[2 * (Tv[3 * e + 1] - 1)], *c = &P[2 * (Tv[3 * e + 2] - 1)];
    const double D = (b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1]);
    if (!(D > 0)) { atomicMin(err, (unsigned long long)(2 * e + 1)); return; }
    const double cx = (a[0] + b[0] + c[0]) / 3, cz = (a[1] + b[1] + c[1]) / 3;
    double ext = fmax(bb[1] - bb[0], bb[3] - bb[2]);
    if (!(ext > 0)) ext = 1;
    key[e] = ((uint64_t)hilbert_d((cx - bb[0]) / ext, (cz - bb[2]) / ext) << 32) | (uint64_t)e;
}

__global__ void k_perm(const uint64_t *__restrict__ sorted, long long T, int *__restrict__ perm, int *__restrict__ inv)
{
    const long long q = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (q >= T) return;
    const int e = (int)(sorted[q] & 0xffffffffu);
    perm[q] = e;
    inv[e] = (int)q;
}

__global__ void k_sponge_flag(const double *__restrict__ ab, long long T, unsigned char *__restrict__ flag)
{
    const long long e = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (e >= T) return;
    unsigned char f = 0;
    for (int i = 0; i < 6; i++)
        if (ab[6 * e + i] != 1.0) f = 1;
    flag[e] = f;
}

/* Neighbour/code and vertex records in device order (dgx_build_mesh_host's second loop).  err: min over failing
 * faces of 3 q + j (neighbour out of range, or not sharing the face's edge). */
__global__ void k_records(const long *__restrict__ Tv, const long *__restrict__ nbraw, const int *__restrict__ perm,
                          const int *__restrict__ inv, const unsigned char *__restrict__ sponge,
                          const unsigned char *__restrict__ explosive, long long T, long long tri_pad,
                          int4 *__restrict__ nbr, int4 *__restrict__ vtx, unsigned long long *err)
{
    const long long q = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (q >= tri_pad) return;
    if (q >= T) {
        nbr[q] = make_int4(-1, -1, -1, 0);
        vtx[q] = make_int4(0, 0, 0, 0);
        return;
    }
    const long long e = perm[q];
    const long *v = &Tv[3 * e];
    int nbq[3] = {-1, -1, -1};
    unsigned code = 0;
    const int col_of_face[3] = {2, 0, 1}; /* .neigh column c holds the element opposite vertex c */
    for (int j = 0; j < 3; j++) {
        const long nb = nbraw[3 * e + col_of_face[j]];
        if (nb == -1 || nb == 0) continue;
        if (nb < 1 || nb > T) { atomicMin(err, (unsigned long long)(3 * q + j)); return; }
        const long *w = &Tv[3 * (nb - 1)];
        const long fa = (j == 2) ? v[0] : v[j], fb = (j == 2) ? v[2] : v[j + 1];
        const long nf[3][2] = {{w[0], w[1]}, {w[1], w[2]}, {w[0], w[2]}};
        int m = -1;
        for (int t = 0; t < 3; t++)
            if ((nf[t][0] == fa && nf[t][1] == fb) || (nf[t][0] == fb && nf[t][1] == fa)) m = t;
        if (m < 0) { atomicMin(err, (unsigned long long)(3 * q + j)); return; }
        const unsigned rev = (nf[m][0] != fa) ? 1u : 0u;
        nbq[j] = inv[nb - 1];
        code |= ((unsigned)m | (rev << 2)) << (3 * j);
        if ((sponge && sponge[nb - 1]) || (explosive && explosive[nb - 1])) code |= DGX_CODE_NBSPECIAL(j);
    }
    if (sponge && sponge[e]) code |= DGX_CODE_SPONGE;
    nbr[q] = make_int4(nbq[0], nbq[1], nbq[2], (int)code);
    vtx[q] = make_int4((int)(v[0] - 1), (int)(v[1] - 1), (int)(v[2] - 1), 0);
}

__global__ void k_fill_int(int *__restrict__ a, long long n, int value)
{
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i < n) a[i] = value;
}

/* First use of every vertex along the element order: the smallest 3 q + k that names it; used counts the
 * vertices named at all (the atomic that replaces the initial value does it, once per vertex). */
__global__ void k_first_use(const int4 *__restrict__ vtx, long long T, int *__restrict__ first, int *__restrict__ used)
{
    const long long q = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (q >= T) return;
    const int4 v = vtx[q];
    const int vs[3] = {v.x, v.y, v.z};
    for (int k = 0; k < 3; k++)
        if (atomicMin(&first[vs[k]], (int)(3 * q + k)) == INT_MAX) atomicAdd(used, 1);
}

__global__ void k_vkeys(const int *__restrict__ first, long long np, uint64_t *__restrict__ key)
{
    const long long v = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (v < np) key[v] = ((uint64_t)(uint32_t)first[v] << 32) | (uint64_t)v;
}

/* Vertex of rank i in the order of first use becomes vertex i. */
__global__ void k_vrank(const uint64_t *__restrict__ sorted, long long nused, const double *__restrict__ P,
                        int *__restrict__ vnew, double2 *__restrict__ pt)
{
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i >= nused) return;
    const long long v = (long long)(sorted[i] & 0xffffffffu);
    vnew[v] = (int)i;
    pt[i] = make_double2(P[2 * v], P[2 * v + 1]);
}

__global__ void k_vremap(int4 *__restrict__ vtx, long long T, const int *__restrict__ vnew)
{
    const long long q = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (q >= T) return;
    int4 v = vtx[q];
    v.x = vnew[v.x]; v.y = vnew[v.y]; v.z = vnew[v.z];
    vtx[q] = v;
}

/* Materials (1/rho | lambda | mu, tri_pad each) and sponge factors in device order (zeros and ones past T).
 * err: min failing q (density not positive). */
__global__ void k_mat_absorb(const int *__restrict__ perm, const double *__restrict__ lam, const double *__restrict__ mu,
                             const double *__restrict__ den, const double *__restrict__ ab, long long T,
                             long long tri_pad, int shot_inner, double *__restrict__ mat, double *__restrict__ absorb,
                             unsigned long long *err)
{
    const long long q = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (q >= tri_pad) return;
    double f[6] = {1, 1, 1, 1, 1, 1};
    if (q < T) {
        const long long e = perm[q];
        if (!(den[e] > 0)) atomicMin(err, (unsigned long long)q);
        mat[q] = 1 / den[e];
        mat[tri_pad + q] = lam[e];
        mat[2 * tri_pad + q] = mu[e];
        if (ab)
            for (int i = 0; i < 6; i++) f[i] = ab[6 * e + i];
    } else {
        mat[q] = 0; mat[tri_pad + q] = 0; mat[2 * tri_pad + q] = 0;
    }
    for (int i = 0; i < 6; i++) absorb[shot_inner ? q * 6 + i : fidx_d(q, i)] = f[i];
}

/* Device buffers of one build, freed together. */
struct Scratch {
    std::vector<void *> bufs;
    ~Scratch()
    {
        for (void *b : bufs)
            if (b) (void)hipFree(b);
    }
    template <typename T>
    hipError_t alloc(T **p, size_t n)
    {
        *p = nullptr;
        if (n == 0) n = 1;
        hipError_t e = hipMalloc((void **)p, n * sizeof(T));
        if (e == hipSuccess) bufs.push_back(*p);
        return e;
    }
    template <typename T>
    hipError_t upload(T **p, const T *src, size_t n)
    {
        hipError_t e = alloc(p, n);
        if (e == hipSuccess && n) e = hipMemcpy(*p, src, n * sizeof(T), hipMemcpyHostToDevice);
        return e;
    }
};

hipError_t bbox_device(const double *dP, long long np, double *dbb, Scratch &sc)
{
    const int nparts = (int)std::min<long long>(1024, blocks_for(np));
    double *part = nullptr;
    hipError_t e = sc.alloc(&part, (size_t)nparts * 4);
    if (e != hipSuccess) return e;
    k_bbox_part<<<nparts, kThreads>>>(dP, np, part);
    k_bbox_final<<<1, 64>>>(part, nparts, dbb);
    return hipGetLastError();
}

hipError_t sort_keys(uint64_t *in, uint64_t *out, long long n, Scratch &sc)
{
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, in, out, (int)n, 0, 64);
    if (e != hipSuccess) return e;
    void *tmp = nullptr;
    e = sc.alloc((char **)&tmp, bytes);
    if (e != hipSuccess) return e;
    return hipcub::DeviceRadixSort::SortKeys(tmp, bytes, in, out, (int)n, 0, 64);
}

unsigned long long read_err(const unsigned long long *d)
{
    unsigned long long h = kNone;
    if (hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return kNone - 1;
    return h;
}

#define MD_TRY(call)                                                                                           \
    do {                                                                                                       \
        hipError_t e_ = (call);                                                                                \
        if (e_ != hipSuccess) {                                                                                \
            dgx_set_error("device mesh set-up: %s failed: %s", #call, hipGetErrorString(e_));                   \
            dgx_mesh_dev_free(out);                                                                            \
            return e_ == hipErrorOutOfMemory ? DGX_ENOMEM : DGX_EHIP;                                          \
        }                                                                                                      \
    } while (0)

} // namespace

void dgx_mesh_dev_free(dgx_mesh_dev *m)
{
    if (!m) return;
    void *p[] = {m->d_nbr, m->d_vtx, m->d_pt, m->d_mat, m->d_absorb};
    for (void *x : p)
        if (x) (void)hipFree(x);
    m->d_nbr = m->d_vtx = nullptr;
    m->d_pt = nullptr;
    m->d_mat = m->d_absorb = nullptr;
}

int dgx_build_mesh_device(const dgx_problem *p, int shot_inner, long tri_pad, const std::vector<char> *explosive,
                          dgx_mesh_dev *out)
{
    const long long T = p->tri, np = p->point_num;
    *out = dgx_mesh_dev();
    if (T <= 0 || np <= 0 || tri_pad < T || (!shot_inner && tri_pad % 64) || T > INT_MAX / 3 || np > INT_MAX) {
        dgx_set_error("device mesh set-up: bad sizes (tri %lld, points %lld, padded %ld)", T, np, tri_pad);
        return DGX_EARG;
    }
    out->tri = (long)T;
    out->tri_pad = tri_pad;
    Scratch sc;
    /* the caller's arrays, as they are */
    double *dP = nullptr, *dlam = nullptr, *dmu = nullptr, *dden = nullptr, *dab = nullptr, *dbb = nullptr;
    long *dTv = nullptr, *dnb = nullptr;
    unsigned char *dexp = nullptr, *dsponge = nullptr;
    MD_TRY(sc.upload(&dP, p->point, (size_t)np * 2));
    MD_TRY(sc.upload(&dTv, p->triangle, (size_t)T * 3));
    MD_TRY(sc.upload(&dnb, p->neighbor, (size_t)T * 3));
    MD_TRY(sc.upload(&dlam, p->lambda, (size_t)T));
    MD_TRY(sc.upload(&dmu, p->miu, (size_t)T));
    MD_TRY(sc.upload(&dden, p->den, (size_t)T));
    if (p->absorb_fun) MD_TRY(sc.upload(&dab, p->absorb_fun, (size_t)T * 6));
    if (explosive) MD_TRY(sc.upload(&dexp, (const unsigned char *)explosive->data(), (size_t)T));
    unsigned long long *derr = nullptr;
    MD_TRY(sc.alloc(&derr, 3));
    {
        const unsigned long long none[3] = {kNone, kNone, kNone};
        MD_TRY(hipMemcpy(derr, none, sizeof none, hipMemcpyHostToDevice));
    }
    MD_TRY(sc.alloc(&dbb, 4));
    MD_TRY(bbox_device(dP, np, dbb, sc));

    /* element order: Hilbert index of the centroid, ties by element id */
    uint64_t *dkey = nullptr, *dkeys = nullptr;
    MD_TRY(sc.alloc(&dkey, (size_t)T));
    MD_TRY(sc.alloc(&dkeys, (size_t)T));
    k_keys<<<blocks_for(T), kThreads>>>(dP, dTv, T, np, dbb, dkey, derr);
    MD_TRY(hipGetLastError());
    if (unsigned long long er = read_err(derr); er != kNone) {
        const long e = (long)(er >> 1);
        if (er == kNone - 1) { dgx_set_error("device mesh set-up: reading the error word failed"); return DGX_EHIP; }
        if ((er & 1) == 0) {
            for (int k = 0; k < 3; k++) {
                const long v = p->triangle[3 * e + k];
                if (v < 1 || v > p->point_num) { dgx_set_error("element %ld: vertex id %ld out of range", e + 1, v); break; }
            }
        } else {
            const double *a = &p->point[2 * (p->triangle[3 * e] - 1)], *b = &p->point[2 * (p->triangle[3 * e + 1] - 1)],
                         *c = &p->point[2 * (p->triangle[3 * e + 2] - 1)];
            const double D = (b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1]);
            dgx_set_error("element %ld is not counter-clockwise (2*area = %g)", e + 1, D);
        }
        return DGX_EMESH;
    }
    MD_TRY(sort_keys(dkey, dkeys, T, sc));
    int *dperm = nullptr, *dinv = nullptr;
    MD_TRY(sc.alloc(&dperm, (size_t)T));
    MD_TRY(sc.alloc(&dinv, (size_t)T));
    k_perm<<<blocks_for(T), kThreads>>>(dkeys, T, dperm, dinv);
    MD_TRY(hipGetLastError());
    if (dab) {
        MD_TRY(sc.alloc(&dsponge, (size_t)T));
        k_sponge_flag<<<blocks_for(T), kThreads>>>(dab, T, dsponge);
        MD_TRY(hipGetLastError());
    }

    /* neighbour/code and vertex records (kept on the device: out owns them from here) */
    MD_TRY(hipMalloc((void **)&out->d_nbr, (size_t)tri_pad * sizeof(int4)));
    MD_TRY(hipMalloc((void **)&out->d_vtx, (size_t)tri_pad * sizeof(int4)));
    k_records<<<blocks_for(tri_pad), kThreads>>>(dTv, dnb, dperm, dinv, dsponge, dexp, T, tri_pad, out->d_nbr,
                                                 out->d_vtx, derr + 1);
    MD_TRY(hipGetLastError());
    if (unsigned long long er = read_err(derr + 1); er != kNone) {
        dgx_mesh_dev_free(out);
        if (er == kNone - 1) { dgx_set_error("device mesh set-up: reading the error word failed"); return DGX_EHIP; }
        const long q = (long)(er / 3);
        const int j = (int)(er % 3);
        std::vector<int> perm1(1);
        if (hipMemcpy(perm1.data(), dperm + q, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
            dgx_set_error("device mesh set-up: reading the failing element failed");
            return DGX_EHIP;
        }
        const long e = perm1[0];
        static const int col_of_face[3] = {2, 0, 1};
        const long nb = p->neighbor[3 * e + col_of_face[j]];
        if (nb < 1 || nb > T) dgx_set_error("element %ld: neighbour %ld out of range", e + 1, nb);
        else dgx_set_error("element %ld face %d: neighbour %ld does not share that edge", e + 1, j, nb);
        return DGX_EMESH;
    }

    /* vertices renumbered in order of first use along the element order */
    int *dfirst = nullptr, *dused = nullptr, *dvnew = nullptr;
    uint64_t *dvkey = nullptr, *dvkeys = nullptr;
    MD_TRY(sc.alloc(&dfirst, (size_t)np));
    MD_TRY(sc.alloc(&dused, 1));
    MD_TRY(sc.alloc(&dvnew, (size_t)np));
    MD_TRY(sc.alloc(&dvkey, (size_t)np));
    MD_TRY(sc.alloc(&dvkeys, (size_t)np));
    k_fill_int<<<blocks_for(np), kThreads>>>(dfirst, np, INT_MAX);
    k_fill_int<<<1, kThreads>>>(dused, 1, 0);
    k_first_use<<<blocks_for(T), kThreads>>>(out->d_vtx, T, dfirst, dused);
    k_vkeys<<<blocks_for(np), kThreads>>>(dfirst, np, dvkey);
    MD_TRY(hipGetLastError());
    MD_TRY(sort_keys(dvkey, dvkeys, np, sc));
    int nused = 0;
    MD_TRY(hipMemcpy(&nused, dused, sizeof(int), hipMemcpyDeviceToHost));
    out->npt = nused;
    MD_TRY(hipMalloc((void **)&out->d_pt, (size_t)(nused > 0 ? nused : 1) * sizeof(double2)));
    k_vrank<<<blocks_for(nused), kThreads>>>(dvkeys, nused, dP, dvnew, out->d_pt);
    k_vremap<<<blocks_for(T), kThreads>>>(out->d_vtx, T, dvnew);
    MD_TRY(hipGetLastError());

    /* materials and sponge factors */
    MD_TRY(hipMalloc((void **)&out->d_mat, (size_t)3 * tri_pad * sizeof(double)));
    MD_TRY(hipMalloc((void **)&out->d_absorb, (size_t)6 * tri_pad * sizeof(double)));
    k_mat_absorb<<<blocks_for(tri_pad), kThreads>>>(dperm, dlam, dmu, dden, dab, T, tri_pad, shot_inner, out->d_mat,
                                                    out->d_absorb, derr + 2);
    MD_TRY(hipGetLastError());
    if (unsigned long long er = read_err(derr + 2); er != kNone) {
        std::vector<int> perm1(1);
        const hipError_t e2 = er == kNone - 1 ? hipErrorUnknown
                                              : hipMemcpy(perm1.data(), dperm + er, sizeof(int), hipMemcpyDeviceToHost);
        dgx_mesh_dev_free(out);
        if (e2 != hipSuccess) { dgx_set_error("device mesh set-up: reading the failing element failed"); return DGX_EHIP; }
        dgx_set_error("element %ld: density must be positive", (long)perm1[0] + 1);
        return DGX_EARG;
    }

    /* host copies: the order and the neighbour records */
    std::vector<int> perm32((size_t)T);
    MD_TRY(hipMemcpy(perm32.data(), dperm, (size_t)T * sizeof(int), hipMemcpyDeviceToHost));
    out->perm.resize((size_t)T);
    out->inv.resize((size_t)T);
    for (long long q = 0; q < T; q++) {
        out->perm[q] = perm32[q];
        out->inv[perm32[q]] = (long)q;
    }
    out->nbr.resize((size_t)tri_pad);
    MD_TRY(hipMemcpy(out->nbr.data(), out->d_nbr, (size_t)tri_pad * sizeof(int4), hipMemcpyDeviceToHost));
    return DGX_OK;
}

/* ---- FP32 copies and the server's model layout ---- */

namespace {

__global__ void k_to_f32(const double *__restrict__ a, long long n, float *__restrict__ b, int deficit)
{
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i < n) b[i] = deficit ? (float)(1.0 - a[i]) : (float)a[i];
}

template <typename R>
__global__ void k_pack_mat(const double *__restrict__ mat, long long tri_pad, long long tri, R *__restrict__ out)
{
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i >= 3 * tri) return;
    const long long c = i / tri, q = i - c * tri;
    out[i] = (R)mat[c * tri_pad + q];
}

} // namespace

hipError_t dgx_mesh_f32(const double *mat, long nmat, const double *absorb, long nab, float **mat32, float **absorb32)
{
    *mat32 = *absorb32 = nullptr;
    hipError_t e = hipMalloc((void **)mat32, (size_t)(nmat > 0 ? nmat : 1) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)absorb32, (size_t)(nab > 0 ? nab : 1) * sizeof(float));
    if (e == hipSuccess && nmat > 0) {
        k_to_f32<<<blocks_for(nmat), kThreads>>>(mat, nmat, *mat32, 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess && nab > 0) {
        k_to_f32<<<blocks_for(nab), kThreads>>>(absorb, nab, *absorb32, 1); /* the FP32 kernels read 1 - f */
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        if (*mat32) (void)hipFree(*mat32);
        if (*absorb32) (void)hipFree(*absorb32);
        *mat32 = *absorb32 = nullptr;
    }
    return e;
}

hipError_t dgx_mesh_mat_packed(const double *mat, long tri_pad, long tri, int fp32, void **out)
{
    *out = nullptr;
    hipError_t e = hipMalloc(out, (size_t)3 * tri * (fp32 ? sizeof(float) : sizeof(double)));
    if (e != hipSuccess) { *out = nullptr; return e; }
    if (fp32) k_pack_mat<float><<<blocks_for(3LL * tri), kThreads>>>(mat, tri_pad, tri, (float *)*out);
    else k_pack_mat<double><<<blocks_for(3LL * tri), kThreads>>>(mat, tri_pad, tri, (double *)*out);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) { (void)hipFree(*out); *out = nullptr; }
    return e;
}

/* ---- the equality check of the two builds (tests) ---- */

extern "C" int dgx_mesh_build_compare(const dgx_problem *p, int shot_inner, long tri_pad, const char *explosive,
                                      long *mismatch)
{
    if (!p || !mismatch || p->tri <= 0 || tri_pad < p->tri || (!shot_inner && tri_pad % 64)) return DGX_EARG;
    std::vector<char> ex;
    if (explosive) ex.assign(explosive, explosive + p->tri);
    dgx_mesh_host h;
    int rc = dgx_build_mesh_host(p, shot_inner, tri_pad, explosive ? &ex : nullptr, &h);
    if (rc) return rc;
    dgx_mesh_dev d;
    rc = dgx_build_mesh_device(p, shot_inner, tri_pad, explosive ? &ex : nullptr, &d);
    if (rc) return rc;
    std::vector<int4> vtx((size_t)tri_pad);
    std::vector<double2> pt((size_t)d.npt);
    std::vector<double> mat((size_t)3 * tri_pad), absorb((size_t)6 * tri_pad);
    hipError_t e = hipMemcpy(vtx.data(), d.d_vtx, vtx.size() * sizeof(int4), hipMemcpyDeviceToHost);
    if (e == hipSuccess && d.npt) e = hipMemcpy(pt.data(), d.d_pt, pt.size() * sizeof(double2), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(mat.data(), d.d_mat, mat.size() * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(absorb.data(), d.d_absorb, absorb.size() * sizeof(double), hipMemcpyDeviceToHost);
    dgx_mesh_dev_free(&d);
    if (e != hipSuccess) { dgx_set_error("dgx_mesh_build_compare: %s", hipGetErrorString(e)); return DGX_EHIP; }
    long bad = 0;
    const char *first = nullptr;
    auto count = [&](const char *what, long n) {
        if (n && !first) first = what;
        bad += n;
    };
    long n = 0;
    for (size_t q = 0; q < h.perm.size(); q++) n += h.perm[q] != d.perm[q] || h.inv[q] != d.inv[q];
    count("perm", n);
    n = 0;
    for (long q = 0; q < tri_pad; q++) n += std::memcmp(&h.nbr[q], &d.nbr[q], sizeof(int4)) != 0;
    count("nbr", n);
    n = 0;
    for (long q = 0; q < tri_pad; q++) n += std::memcmp(&h.vtx[q], &vtx[q], sizeof(int4)) != 0;
    count("vtx", n);
    n = (long)h.pt.size() != d.npt;
    for (size_t i = 0; i < std::min(h.pt.size(), pt.size()); i++) n += std::memcmp(&h.pt[i], &pt[i], sizeof(double2)) != 0;
    count("pt", n);
    n = 0;
    for (size_t i = 0; i < mat.size(); i++) n += std::memcmp(&h.mat[i], &mat[i], sizeof(double)) != 0;
    count("mat", n);
    n = 0;
    for (size_t i = 0; i < absorb.size(); i++) n += std::memcmp(&h.absorb[i], &absorb[i], sizeof(double)) != 0;
    count("absorb", n);
    *mismatch = bad;
    if (first) dgx_set_error("dgx_mesh_build_compare: first mismatching array: %s", first);
    return DGX_OK;
}

/* ---- point location on the device ---- */

namespace {

struct Grid {
    double x0, z0, h;
    long nx, nz;
};

__host__ __device__ __forceinline__ long clampl_d(long v, long lo, long hi) { return v < lo ? lo : (v > hi ? hi : v); }

__host__ __device__ __forceinline__ void cell_of_d(const Grid &g, double x, double z, long *ix, long *iz)
{
    *ix = clampl_d((long)floor((x - g.x0) / g.h), 0, g.nx - 1);
    *iz = clampl_d((long)floor((z - g.z0) / g.h), 0, g.nz - 1);
}

/* Vertex matches: for query k the first (smallest) or last (largest) vertex within tol in both coordinates whose
 * grid cell is within one cell of the query's (the host locator's 3 x 3 cell search). */
__global__ void k_vmatch(const double *__restrict__ P, long long np, Grid g, const double *__restrict__ qxz,
                         const long *__restrict__ qcell, int nq, double tol, int last, unsigned long long *best)
{
    const long long v = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (v >= np) return;
    const double px = P[2 * v], pz = P[2 * v + 1];
    long vx, vz;
    cell_of_d(g, px, pz, &vx, &vz);
    for (int k = 0; k < nq; k++) {
        const long dx = vx - qcell[2 * k], dz = vz - qcell[2 * k + 1];
        if (dx < -1 || dx > 1 || dz < -1 || dz > 1) continue;
        if (fabs(px - qxz[2 * k]) < tol && fabs(pz - qxz[2 * k + 1]) < tol) {
            if (last) atomicMax((long long *)&best[k], (long long)v);
            else atomicMin(&best[k], (unsigned long long)v);
        }
    }
}

/* Smallest and largest element holding each vertex.  err: min failing element (vertex id out of range). */
__global__ void k_velem(const long *__restrict__ Tv, long long T, long long np, int *__restrict__ emin,
                        int *__restrict__ emax, unsigned long long *err)
{
    const long long e = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (e >= T) return;
    for (int k = 0; k < 3; k++) {
        const long v = Tv[3 * e + k] - 1;
        if (v < 0 || v >= np) { atomicMin(err, (unsigned long long)e); return; }
        atomicMin(&emin[v], (int)e);
        atomicMax(&emax[v], (int)e);
    }
}

/* Element of each vertex match: the smallest (first) or largest (last) element holding the vertex, -1 for no
 * match or a vertex no element holds. */
__global__ void k_match_elem(const unsigned long long *__restrict__ best, int nq, int last, const int *__restrict__ emin,
                             const int *__restrict__ emax, long long *__restrict__ out)
{
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= nq) return;
    const long long b = (long long)best[k];
    long long e = -1;
    if (last ? b >= 0 : best[k] != kNone) {
        const int x = last ? emax[b] : emin[b];
        if (last ? x >= 0 : x != INT_MAX) e = x;
    }
    out[k] = e;
}

/* The host locator's barycentric test (dgx_setup.c, inside_elem), same expressions. */
__device__ __forceinline__ bool inside_d(const double *P, const long *Tv, long long e, double x, double z)
{
    double v[3][2];
    for (int k = 0; k < 3; k++) {
        const long p = Tv[e * 3 + k] - 1;
        v[k][0] = P[p * 2];
        v[k][1] = P[p * 2 + 1];
    }
    const double xr = v[1][0] - v[0][0], zr = v[1][1] - v[0][1];
    const double xs = v[2][0] - v[0][0], zs = v[2][1] - v[0][1];
    const double det = xr * zs - xs * zr;
    const double dx = x - v[0][0], dz = z - v[0][1];
    const double a = (dx * zs - xs * dz) / det, b = (xr * dz - dx * zr) / det;
    const double tol = 1e-10;
    return a >= -tol && b >= -tol && a + b <= 1 + tol;
}

/* Point location: for each query without a vertex match, the smallest element whose grid-cell range (its
 * bounding box padded by 1e-9 h, as the host locator files it) covers the query's cell and which contains the
 * point. */
__global__ void k_locate(const double *__restrict__ P, const long *__restrict__ Tv, long long T, Grid g,
                         const double *__restrict__ qxz, const long *__restrict__ qcell, const int *__restrict__ todo,
                         int nt, unsigned long long *loc)
{
    const long long e = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (e >= T) return;
    double bx0 = 1e300, bx1 = -1e300, bz0 = 1e300, bz1 = -1e300;
    for (int k = 0; k < 3; k++) {
        const long p = Tv[e * 3 + k] - 1;
        bx0 = fmin(bx0, P[p * 2]); bx1 = fmax(bx1, P[p * 2]);
        bz0 = fmin(bz0, P[p * 2 + 1]); bz1 = fmax(bz1, P[p * 2 + 1]);
    }
    long ix0, iz0, ix1, iz1;
    cell_of_d(g, bx0 - 1e-9 * g.h, bz0 - 1e-9 * g.h, &ix0, &iz0);
    cell_of_d(g, bx1 + 1e-9 * g.h, bz1 + 1e-9 * g.h, &ix1, &iz1);
    for (int t = 0; t < nt; t++) {
        const int k = todo[t];
        const long cx = qcell[2 * k], cz = qcell[2 * k + 1];
        if (cx < ix0 || cx > ix1 || cz < iz0 || cz > iz1) continue;
        if (inside_d(P, Tv, e, qxz[2 * k], qxz[2 * k + 1])) atomicMin(&loc[k], (unsigned long long)e);
    }
}

/* Locate n points: element (0-based) and matched vertex (or -1) per point; -1 element: outside the mesh. */
int locate_device(long tri, long point_num, const double *point, const long *triangle, long n, const double *xz,
                  double tol, int last, std::vector<long> &elem, std::vector<long> &vert)
{
    if (tri <= 0 || point_num <= 0 || !point || !triangle || n < 0 || (n > 0 && !xz)) return DGX_EARG;
    if (n > INT_MAX / 2) return DGX_EARG;
    elem.assign((size_t)n, -1);
    vert.assign((size_t)n, -1);
    if (n == 0) return DGX_OK;
    Scratch sc;
    dgx_mesh_dev *out = nullptr; /* MD_TRY frees nothing here */
    double *dP = nullptr, *dbb = nullptr, *dq = nullptr;
    long *dTv = nullptr, *dqc = nullptr;
    MD_TRY(sc.upload(&dP, point, (size_t)point_num * 2));
    MD_TRY(sc.upload(&dTv, triangle, (size_t)tri * 3));
    MD_TRY(sc.alloc(&dbb, 4));
    MD_TRY(bbox_device(dP, point_num, dbb, sc));
    double bb[4];
    MD_TRY(hipMemcpy(bb, dbb, sizeof bb, hipMemcpyDeviceToHost));
    /* the host locator's grid (dgx_locator_create) */
    Grid g;
    const double area = (bb[1] - bb[0]) * (bb[3] - bb[2]);
    double h = sqrt((area > 0 ? area : 1.0) / (double)tri) * 2.0;
    if (!(h > 0)) h = 1.0;
    g.h = h;
    g.x0 = bb[0] - 1e-6 * h;
    g.z0 = bb[2] - 1e-6 * h;
    g.nx = (long)((bb[1] - g.x0) / h) + 1;
    g.nz = (long)((bb[3] - g.z0) / h) + 1;
    std::vector<long> qc((size_t)n * 2);
    for (long k = 0; k < n; k++) cell_of_d(g, xz[2 * k], xz[2 * k + 1], &qc[2 * k], &qc[2 * k + 1]);
    MD_TRY(sc.upload(&dq, xz, (size_t)n * 2));
    MD_TRY(sc.upload(&dqc, qc.data(), (size_t)n * 2));
    /* vertex matches and the elements holding each vertex */
    unsigned long long *dbest = nullptr, *derr = nullptr;
    int *demin = nullptr, *demax = nullptr;
    MD_TRY(sc.alloc(&dbest, (size_t)n));
    MD_TRY(sc.alloc(&derr, 1));
    MD_TRY(sc.alloc(&demin, (size_t)point_num));
    MD_TRY(sc.alloc(&demax, (size_t)point_num));
    {
        std::vector<unsigned long long> init((size_t)n, last ? (unsigned long long)-1LL : kNone);
        MD_TRY(hipMemcpy(dbest, init.data(), init.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
        const unsigned long long none = kNone;
        MD_TRY(hipMemcpy(derr, &none, sizeof none, hipMemcpyHostToDevice));
    }
    k_fill_int<<<blocks_for(point_num), kThreads>>>(demin, point_num, INT_MAX);
    k_fill_int<<<blocks_for(point_num), kThreads>>>(demax, point_num, -1);
    k_velem<<<blocks_for(tri), kThreads>>>(dTv, tri, point_num, demin, demax, derr);
    k_vmatch<<<blocks_for(point_num), kThreads>>>(dP, point_num, g, dq, dqc, (int)n, tol, last, dbest);
    MD_TRY(hipGetLastError());
    if (unsigned long long er = read_err(derr); er != kNone) {
        dgx_set_error("element %ld: vertex id out of range", (long)er + 1);
        return DGX_EMESH;
    }
    long long *dme = nullptr;
    MD_TRY(sc.alloc(&dme, (size_t)n));
    k_match_elem<<<blocks_for(n), kThreads>>>(dbest, (int)n, last, demin, demax, dme);
    MD_TRY(hipGetLastError());
    std::vector<unsigned long long> best((size_t)n);
    std::vector<long long> me((size_t)n);
    MD_TRY(hipMemcpy(best.data(), dbest, best.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    MD_TRY(hipMemcpy(me.data(), dme, me.size() * sizeof(long long), hipMemcpyDeviceToHost));
    std::vector<int> todo;
    for (long k = 0; k < n; k++) {
        if (me[k] >= 0) {
            elem[k] = (long)me[k];
            vert[k] = (long)(long long)best[k];
        } else {
            todo.push_back((int)k);
        }
    }
    if (!todo.empty()) {
        int *dtodo = nullptr;
        unsigned long long *dloc = nullptr;
        MD_TRY(sc.upload(&dtodo, todo.data(), todo.size()));
        MD_TRY(sc.alloc(&dloc, (size_t)n));
        std::vector<unsigned long long> init((size_t)n, kNone);
        MD_TRY(hipMemcpy(dloc, init.data(), init.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
        k_locate<<<blocks_for(tri), kThreads>>>(dP, dTv, tri, g, dq, dqc, dtodo, (int)todo.size(), dloc);
        MD_TRY(hipGetLastError());
        std::vector<unsigned long long> loc((size_t)n);
        MD_TRY(hipMemcpy(loc.data(), dloc, loc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        for (int k : todo)
            if (loc[k] != kNone) elem[k] = (long)loc[k];
    }
    return DGX_OK;
}

} // namespace

extern "C" int dgx_locate_receivers_device(long tri, long point_num, const double *point, const long *triangle, long n,
                                           const double *xz, dgx_receiver *out)
{
    if (n > 0 && !out) return DGX_EARG;
    std::vector<long> elem, vert;
    int rc = locate_device(tri, point_num, point, triangle, n, xz, 1e-2, 1, elem, vert);
    if (rc) return rc;
    for (long k = 0; k < n; k++) {
        if (elem[k] < 0) {
            dgx_set_error("receiver %ld at (%g, %g) lies outside the mesh", k, xz[2 * k], xz[2 * k + 1]);
            return DGX_EARG;
        }
        rc = dgx_receiver_at(point, triangle, elem[k], vert[k], xz[2 * k], xz[2 * k + 1], &out[k]);
        if (rc) return rc;
    }
    return DGX_OK;
}

extern "C" int dgx_locate_sources_device(long tri, long point_num, const double *point, const long *triangle,
                                         const dgx_operators *ops, long n, const double *xz, const int *kind,
                                         const int *shot, const long *wavelet, dgx_source *out)
{
    if (n > 0 && (!ops || !kind || !shot || !wavelet || !out)) return DGX_EARG;
    std::vector<long> elem, vert;
    int rc = locate_device(tri, point_num, point, triangle, n, xz, 1e-3, 0, elem, vert);
    if (rc) return rc;
    for (long k = 0; k < n; k++) {
        if (elem[k] < 0) {
            dgx_set_error("source (%g, %g) lies outside the mesh", xz[2 * k], xz[2 * k + 1]);
            return DGX_EARG;
        }
        rc = dgx_source_at(point, triangle, ops, elem[k], vert[k], xz[2 * k], xz[2 * k + 1], kind[k], shot[k],
                           wavelet[k], &out[k]);
        if (rc) return rc;
    }
    return DGX_OK;
}
