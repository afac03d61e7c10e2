/*
 * dgx_batched.hip -- shot-innermost kernels: many shots (independent wavefields) on one mesh.
 *
 * Layout: value (element q, node i, shot s) of a field lives at (q*6 + i)*S + s, S the padded shot
 * count.  A wavefront covers EW = 64/ST elements x ST consecutive shots (lane = e_sub*ST + shot), and a
 * workgroup NW wavefronts (NW*EW elements; shot tile and NW are picked by the autotuner at create), so
 * every load of one node of one field for one element is ST consecutive values, and so is every face
 * gather from a neighbour: the irregular part of the access pattern is amortised over ST shots.  The
 * per-element data (neighbours, code word, vertices, material) is identical across a lane group and is
 * read once per ST shots; with ST = 64 it is wave-uniform.
 *
 * The arithmetic is the same as the element-per-thread kernels (dgx_plain.hip), term for term, so
 * the two layouts agree to rounding; the tests compare them.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dgx_device.h"
#include "dgx_kernels.h"

namespace dgxb {

using dgx::as_global;
using dgx::face_node;
using dgx::Geom;
using dgx::geometry;
using dgx::nb_node;
using dgx::sponge;
using dgx::Ops;
using dgx::StepArgs;


/* Element and shot of this lane (workgroups of NW wavefronts); false when outside the mesh. */
template <int ST, int NW, typename R>
__device__ __forceinline__ bool my_work(const StepArgs<R> &a, int &q, int &s)
{
    constexpr int EW = 64 / ST;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = blockIdx.x;
    const int tile = (b & (DGX_NXCD - 1)) * a.nper + (b >> 3); /* XCD-contiguous element tiles */
    q = (tile * NW + wave) * EW + lane / ST;
    s = blockIdx.y * ST + (lane % ST);
    return q < a.tri;
}

/* Offset of value (element q, node i, shot s).  Rows of spad shots; with L > 0 each node row interleaves 2^L
 * consecutive elements, ((q >> L)*6 + i)*2^L + (q mod 2^L), so that the 64 lanes of a wavefront that holds
 * 2^L elements x spad shots (spad = 64 >> L) read and write one contiguous 256-byte (FP32) span per node.
 * L = 0 (one element per row) for the batched solver; the serving groups use L = log2(64/G). */
template <int L, typename R>
__device__ __forceinline__ long long bidx(const StepArgs<R> &a, int q, int i, int s)
{
    if constexpr (L == 0) return ((long long)q * 6 + i) * a.spad + s;
    else return ((((long long)(q >> L) * 6 + i) << L) + (q & ((1 << L) - 1))) * a.spad + s;
}

/* local step of shot s at the launch's global step; -1 when the shot is not running */
template <typename R>
__device__ __forceinline__ int shot_step(const StepArgs<R> &a, int s, int gstep)
{
    if (!a.shot_start) return gstep;
    const int t = gstep - a.shot_start[s];
    return (t >= 0 && t < a.shot_len[s]) ? t : -1;
}

/* aux record index + 1 of element q (0: none) */
template <typename R>
__device__ __forceinline__ unsigned aux_of(const StepArgs<R> &a, int q, unsigned code)
{
    return a.aux_of_elem ? ((unsigned)a.aux_of_elem[q] & DGX_DYN_AUX_MASK) : (code >> DGX_CODE_AUX_SHIFT);
}

template <typename R>
__device__ __forceinline__ double wavelet_at(const StepArgs<R> &a, const dgx_src_dev &src, int s, int lstep)
{
    return a.slot_wav ? as_global(a.slot_wav[s])[lstep] : a.wav[(long long)src.wavelet * a.nsteps + lstep];
}

template <typename R>
__device__ __forceinline__ void condition_stress(const StepArgs<R> &a, unsigned code, int q, int i, int s, int lstep,
                                                 R &txx, R &tzz, R &txz)
{
    if (code & DGX_CODE_SPONGE) {
        const R f = a.absorb[q * 6 + i];
        txx = sponge(txx, f); tzz = sponge(tzz, f); txz = sponge(txz, f);
    }
    const unsigned ax = aux_of(a, q, code);
    if (ax && lstep >= 0) {
        const dgx_aux x = a.aux[ax - 1];
        for (int k = x.src_begin; k < x.src_begin + x.src_count; k++) {
            const dgx_src_dev &src = a.src[k];
            if (src.shot != s || src.kind != 1) continue;
            const R inc = (R)(src.fun[i] * wavelet_at(a, src, s, lstep));
            txx += inc;
            tzz += inc;
        }
    }
}

/* condition_stress for all 6 nodes: one sponge branch and one source branch per element */
template <typename R>
__device__ __forceinline__ void condition_stress6(const StepArgs<R> &a, unsigned code, int q, int s, int lstep,
                                                  R *txx, R *tzz, R *txz)
{
    if (code & DGX_CODE_SPONGE) {
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const R f = a.absorb[q * 6 + i];
            txx[i] = sponge(txx[i], f); tzz[i] = sponge(tzz[i], f); txz[i] = sponge(txz[i], f);
        }
    }
    const unsigned ax = aux_of(a, q, code);
    if (ax && lstep >= 0) {
        const dgx_aux x = a.aux[ax - 1];
        for (int k = x.src_begin; k < x.src_begin + x.src_count; k++) {
            const dgx_src_dev &src = a.src[k];
            if (src.shot != s || src.kind != 1) continue;
            const double w = wavelet_at(a, src, s, lstep);
#pragma unroll
            for (int i = 0; i < 6; i++) {
                const R inc = (R)(src.fun[i] * w);
                txx[i] += inc;
                tzz[i] += inc;
            }
        }
    }
}

template <typename R, int L = 0>
__device__ __forceinline__ void velocity_body(const Ops<R> &op, const StepArgs<R> &a, int q, int s, int lstep)
{
    R *__restrict__ Vx = a.f[0];
    R *__restrict__ Vz = a.f[1];
    const R *__restrict__ Txx = a.f[2];
    const R *__restrict__ Tzz = a.f[3];
    const R *__restrict__ Txz = a.f[4];

    const int4 nb = a.nbr[q];
    const unsigned code = (unsigned)nb.w;
    const Geom g = geometry(a.pt, a.vtx[q]);

    /* stresses conditioned (sponge, explosive source) before use; velocities are conditioned at the
     * update.  FP32 issues the velocity loads up front; FP64 loads them at the update, which keeps
     * the register count (and so the number of resident wavefronts) where it is without them. */
    R txx[6], tzz[6], txz[6], vxe[6], vze[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const long long o = bidx<L>(a, q, i, s);
        txx[i] = Txx[o]; tzz[i] = Tzz[o]; txz[i] = Txz[o];
        if (sizeof(R) == 4) { vxe[i] = Vx[o]; vze[i] = Vz[o]; }
    }
    if constexpr (sizeof(R) == 4) {
        condition_stress6(a, code, q, s, lstep, txx, tzz, txz); /* FP32: per element */
    } else {
#pragma unroll
        for (int i = 0; i < 6; i++) condition_stress(a, code, q, i, s, lstep, txx[i], tzz[i], txz[i]);
    }

    const R rx = (R)g.rx, rz = (R)g.rz, sx = (R)g.sx, sz = (R)g.sz;
    R ux[6], wx[6], uz[6], wz[6];
#pragma unroll
    for (int n = 0; n < 6; n++) {
        ux[n] = rx * txx[n] + rz * txz[n];
        wx[n] = sx * txx[n] + sz * txz[n];
        uz[n] = rx * txz[n] + rz * tzz[n];
        wz[n] = sx * txz[n] + sz * tzz[n];
    }
    R rhx[6], rhz[6], flx[6], flz[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        R sx_ = 0, sz_ = 0;
#pragma unroll
        for (int n = 0; n < 6; n++) {
            sx_ += op.Dr[i * 6 + n] * ux[n] + op.Ds[i * 6 + n] * wx[n];
            sz_ += op.Dr[i * 6 + n] * uz[n] + op.Ds[i * 6 + n] * wz[n];
        }
        rhx[i] = sx_;
        rhz[i] = sz_;
        flx[i] = 0;
        flz[i] = 0;
    }
    {
        const int nbs[3] = {nb.x, nb.y, nb.z};
        R gxx[3][3], gzz[3][3], gxz[3][3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int ee = nbs[j] >= 0 ? nbs[j] : q;
            const unsigned m = (code >> (3 * j)) & 3u, rev = (code >> (3 * j + 2)) & 1u;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const long long o = bidx<L>(a, ee, nb_node(m, rev, k), s);
                gxx[j][k] = Txx[o]; gzz[j][k] = Tzz[o]; gxz[j][k] = Txz[o];
            }
        }
        /* faces whose neighbour needs its stresses conditioned: sponge (static, in the code word) or,
         * when serving, an explosive source of this batch (the element's own word) */
        const unsigned dyn = a.aux_of_elem ? ((unsigned)a.aux_of_elem[q] >> DGX_DYN_NB_SHIFT) : 0u;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int e = nbs[j];
            if (e < 0) continue;
            const bool special = (code & DGX_CODE_NBSPECIAL(j)) != 0 || ((dyn >> j) & 1u);
            if (!special) continue;
            const unsigned m = (code >> (3 * j)) & 3u, rev = (code >> (3 * j + 2)) & 1u;
            const unsigned ncode = (unsigned)a.nbr[e].w;
            for (int k = 0; k < 3; k++)
                condition_stress(a, ncode, e, nb_node(m, rev, k), s, lstep, gxx[j][k], gzz[j][k], gxz[j][k]);
        }
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const bool has = nbs[j] >= 0;
            const R ex = (R)g.ex[j], ez = (R)g.ez[j];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int on = face_node(j, k);
                R dxx = -txx[on], dzz = -tzz[on], dxz = -txz[on];
                if (has) { dxx += gxx[j][k]; dzz += gzz[j][k]; dxz += gxz[j][k]; }
                const R gx = ez * dxx - ex * dxz;
                const R gz = ez * dxz - ex * dzz;
#pragma unroll
                for (int i = 0; i < 6; i++) {
                    flx[i] += op.L[j * 18 + i * 3 + k] * gx;
                    flz[i] += op.L[j * 18 + i * 3 + k] * gz;
                }
            }
        }
    }
    const R invD = (R)g.invD;
    const R c = a.rho_inv[q] * (R)a.dt;
    const unsigned ax = aux_of(a, q, code);
    if constexpr (sizeof(R) == 4) {
        /* FP32: velocities already loaded; sponge and sources applied per element (one branch each),
         * which measures faster than per node; FP64 keeps the per-node form below, which needs fewer
         * registers */
        if (code & DGX_CODE_SPONGE) {
#pragma unroll
            for (int i = 0; i < 6; i++) {
                const R f = a.absorb[q * 6 + i];
                vxe[i] = sponge(vxe[i], f);
                vze[i] = sponge(vze[i], f);
            }
        }
        if (ax && lstep >= 0) {
            const dgx_aux x = a.aux[ax - 1];
            for (int k = x.src_begin; k < x.src_begin + x.src_count; k++) {
                const dgx_src_dev &src = a.src[k];
                if (src.shot != s || src.kind != 0) continue;
                const double w = wavelet_at(a, src, s, lstep);
#pragma unroll
                for (int i = 0; i < 6; i++) vze[i] += (R)(src.fun[i] * w);
            }
        }
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const long long o = bidx<L>(a, q, i, s);
            Vx[o] = (rhx[i] + flx[i] * invD) * c + vxe[i];
            Vz[o] = (rhz[i] + flz[i] * invD) * c + vze[i];
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const long long o = bidx<L>(a, q, i, s);
        R vx, vz;
        if (sizeof(R) == 4) { vx = vxe[i]; vz = vze[i]; }
        else { vx = Vx[o]; vz = Vz[o]; }
        if (code & DGX_CODE_SPONGE) {
            const R f = a.absorb[q * 6 + i];
            vx = sponge(vx, f);
            vz = sponge(vz, f);
        }
        if (ax && lstep >= 0) {
            const dgx_aux x = a.aux[ax - 1];
            for (int k = x.src_begin; k < x.src_begin + x.src_count; k++) {
                const dgx_src_dev &src = a.src[k];
                if (src.shot == s && src.kind == 0) vz += (R)(src.fun[i] * wavelet_at(a, src, s, lstep));
            }
        }
        Vx[o] = (rhx[i] + flx[i] * invD) * c + vx;
        Vz[o] = (rhz[i] + flz[i] * invD) * c + vz;
    }
}

template <typename R, int L = 0>
__device__ __forceinline__ void stress_body(const Ops<R> &op, const StepArgs<R> &a, int q, int s, int lstep)
{
    const R *__restrict__ Vx = a.f[0];
    const R *__restrict__ Vz = a.f[1];
    R *__restrict__ Txx = a.f[2];
    R *__restrict__ Tzz = a.f[3];
    R *__restrict__ Txz = a.f[4];

    const int4 nb = a.nbr[q];
    const unsigned code = (unsigned)nb.w;
    const Geom g = geometry(a.pt, a.vtx[q]);

    R vx[6], vz[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const long long o = bidx<L>(a, q, i, s);
        vx[i] = Vx[o]; vz[i] = Vz[o];
    }
    const R rx = (R)g.rx, rz = (R)g.rz, sx = (R)g.sx, sz = (R)g.sz;
    R exx[6], ezz[6], exz[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        R rvx = 0, svx = 0, rvz = 0, svz = 0;
#pragma unroll
        for (int n = 0; n < 6; n++) {
            rvx += op.Dr[i * 6 + n] * vx[n];
            svx += op.Ds[i * 6 + n] * vx[n];
            rvz += op.Dr[i * 6 + n] * vz[n];
            svz += op.Ds[i * 6 + n] * vz[n];
        }
        exx[i] = rx * rvx + sx * svx;
        ezz[i] = rz * rvz + sz * svz;
        exz[i] = rx * rvz + sx * svz + rz * rvx + sz * svx;
    }
    R fx[6], fz[6], fxz[6];
#pragma unroll
    for (int i = 0; i < 6; i++) { fx[i] = 0; fz[i] = 0; fxz[i] = 0; }
    {
        const int nbs[3] = {nb.x, nb.y, nb.z};
        R gvx[3][3], gvz[3][3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int ee = nbs[j] >= 0 ? nbs[j] : q;
            const unsigned m = (code >> (3 * j)) & 3u, rev = (code >> (3 * j + 2)) & 1u;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const long long o = bidx<L>(a, ee, nb_node(m, rev, k), s);
                gvx[j][k] = Vx[o]; gvz[j][k] = Vz[o];
            }
        }
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const bool has = nbs[j] >= 0;
            const R ex = (R)g.ex[j], ez = (R)g.ez[j];
            R px[3], pz[3], pxz[3];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int on = face_node(j, k);
                R dvx = -vx[on], dvz = -vz[on];
                if (has) { dvx += gvx[j][k]; dvz += gvz[j][k]; }
                px[k] = ez * dvx;
                pz[k] = -ex * dvz;
                pxz[k] = ez * dvz - ex * dvx;
            }
#pragma unroll
            for (int i = 0; i < 6; i++) {
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const R l = op.L[j * 18 + i * 3 + k];
                    fx[i] += l * px[k];
                    fz[i] += l * pz[k];
                    fxz[i] += l * pxz[k];
                }
            }
        }
    }
    const R invD = (R)g.invD;
    const R lam = a.lam[q], mu = a.mu[q], l2m = lam + 2 * mu, dt = (R)a.dt;
    if constexpr (sizeof(R) == 4) {
        /* FP32: own stresses loaded together and conditioned per element */
        R txx[6], tzz[6], txz[6];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const long long o = bidx<L>(a, q, i, s);
            txx[i] = Txx[o]; tzz[i] = Tzz[o]; txz[i] = Txz[o];
        }
        condition_stress6(a, code, q, s, lstep, txx, tzz, txz);
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const long long o = bidx<L>(a, q, i, s);
            const R axx = exx[i] + fx[i] * invD, azz = ezz[i] + fz[i] * invD, axz = exz[i] + fxz[i] * invD;
            Txx[o] = (l2m * axx + lam * azz) * dt + txx[i];
            Tzz[o] = (lam * axx + l2m * azz) * dt + tzz[i];
            Txz[o] = (mu * axz) * dt + txz[i];
        }
    } else
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const long long o = bidx<L>(a, q, i, s);
        R txx = Txx[o], tzz = Tzz[o], txz = Txz[o];
        condition_stress(a, code, q, i, s, lstep, txx, tzz, txz);
        const R axx = exx[i] + fx[i] * invD, azz = ezz[i] + fz[i] * invD, axz = exz[i] + fxz[i] * invD;
        Txx[o] = (l2m * axx + lam * azz) * dt + txx;
        Tzz[o] = (lam * axx + l2m * azz) * dt + tzz;
        Txz[o] = (mu * axz) * dt + txz;
    }
    const unsigned ax = aux_of(a, q, code);
    if (ax && lstep >= 0 && s < a.nshots && (a.slot_trace || lstep < a.nsteps)) {
        const dgx_aux x = a.aux[ax - 1];
        for (int r = x.rec_begin; r < x.rec_begin + x.rec_count; r++) {
            const dgx_rec_dev &rc = a.rec[r];
            if (a.slot_trace && rc.pad0 != s) continue;
            double sx_ = 0, sz_ = 0;
#pragma unroll
            for (int i = 0; i < 6; i++) {
                sx_ += rc.w[i] * (double)vx[i];
                sz_ += rc.w[i] * (double)vz[i];
            }
            const int nrec = a.slot_trace ? a.slot_nrec[s] : a.nrec;
            double *t = a.slot_trace ? as_global(a.slot_trace[s]) + (long long)lstep * 2 * nrec
                                     : a.trace + ((long long)s * a.nsteps + lstep) * 2 * nrec;
            t[rc.index] = sx_;
            t[nrec + rc.index] = sz_;
        }
    }
}

/* IL: the tile holds every shot of the row (ST == spad < 64), whose node rows interleave the 64/ST elements
 * of a wavefront (bidx with L = log2(64/ST)) */
template <typename R, int ST, int NW, bool IL>
__global__ void __launch_bounds__(64 * NW) kb_velocity(StepArgs<R> a)
{
    constexpr int L = IL ? __builtin_ctz(64 / ST) : 0;
    int q, s;
    if (!my_work<ST, NW>(a, q, s)) return;
    /* with 64 shots per tile a wavefront holds one element: saying so turns the element's records, its
     * neighbour ids and the row offsets into scalar values (measured: faster for the velocity kernel,
     * slightly slower for the stress kernel, which keeps the per-lane form) */
    if (ST == 64) q = __builtin_amdgcn_readfirstlane(q);
    velocity_body<R, L>(a.op, a, q, s, shot_step(a, s, *a.step_base + a.step_off));
}

template <typename R, int ST, int NW, bool IL>
__global__ void __launch_bounds__(64 * NW) kb_stress(StepArgs<R> a)
{
    constexpr int L = IL ? __builtin_ctz(64 / ST) : 0;
    int q, s;
    if (!my_work<ST, NW>(a, q, s)) return;
    stress_body<R, L>(a.op, a, q, s, shot_step(a, s, *a.step_base + a.step_off));
}

template <typename R>
__device__ __forceinline__ StepArgs<R> global_args(const StepArgs<R> &t)
{
    StepArgs<R> a = t;
    for (int c = 0; c < 5; c++) a.f[c] = as_global(a.f[c]);
    a.nbr = as_global(a.nbr);
    a.vtx = as_global(a.vtx);
    a.pt = as_global(a.pt);
    a.rho_inv = as_global(a.rho_inv);
    a.lam = as_global(a.lam);
    a.mu = as_global(a.mu);
    a.absorb = as_global(a.absorb);
    a.aux = as_global(a.aux);
    a.src = as_global(a.src);
    a.rec = as_global(a.rec);
    a.shot_start = as_global(a.shot_start);
    a.shot_len = as_global(a.shot_len);
    a.aux_of_elem = as_global(a.aux_of_elem);
    a.slot_wav = as_global(a.slot_wav);
    a.slot_trace = as_global(a.slot_trace);
    a.slot_nrec = as_global(a.slot_nrec);
    return a;
}

/*
 * Serving: one launch per half-step over every active (mesh, slot group) batch.  sched = {nbatch,
 * total tiles, tile offset of batch 0, ..., tile offset of batch nbatch}; a tile is 256/G elements x G
 * shots (G = slots per group, 4, 8, 16, 32 or 64: 64/G elements per wavefront, stored interleaved in the
 * group's node rows so that each node load and store of a wavefront is one contiguous span; bidx).  One workgroup per tile, and the
 * tiles are dealt to the 8 XCDs in contiguous ranges swept in order, as in the batched kernels: face
 * neighbours are stepped close together in time on one XCD and their rows are still in its L2.  (A
 * persistent grid-stride loop over the same tiles let the workgroups drift apart over their ~200 tiles
 * each; measured on 640k elements x 64 shots, it fetched 1.28-1.44x the field bytes from HBM against
 * 1.02-1.04x here, and ran 20 % slower.)  The grid depends on the tile count, so the host recaptures the
 * step graph when the set of stepped batches changes; the table itself is read from device memory.
 *
 * Everything a wavefront needs before its first element load comes through the scalar cache: the batch
 * (binary search over the tile offsets), its table entry, and which of its slots run at this step
 * (runmask[step_off * nb + batch], bit s = slot s, written by the host before each block).  The slot's
 * local step, needed only where a source or receiver sits, is read per lane alongside the element loads.
 */
template <typename R, bool VEL, int G>
__global__ void __launch_bounds__(DGX_BLOCK) ks_step(const Ops<R> op, const StepArgs<R> *__restrict__ table,
                                                        const int *__restrict__ sched,
                                                        const unsigned long long *__restrict__ runmask,
                                                        const int *step_base, int step_off, int nb, int total,
                                                        int nper)
{
    constexpr int EW = 64 / G; /* elements per wavefront, interleaved in one node row of the group */
    constexpr int LG = G == 64 ? 0 : G == 32 ? 1 : G == 16 ? 2 : G == 8 ? 3 : 4;
    static_assert((1 << LG) == EW, "slot groups of 4, 8, 16, 32 or 64 shots");
    const int tile = (int)(blockIdx.x & (DGX_NXCD - 1)) * nper + (int)(blockIdx.x >> 3);
    if (tile >= total) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    /* batch of the tile: the last batch whose first tile sched[2 + b] is at or below it (the offsets
     * increase strictly; every batch has at least one tile) */
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sched[2 + mid] <= tile) lo = mid;
        else hi = mid - 1;
    }
    const int b = lo;
    const unsigned long long run = runmask[(long long)step_off * nb + b];
    const int s = lane % G;
    /* a slot that is not running (free, or its request has not started / has finished) is skipped: its
     * fields are never read again before the slot is cleared for a new request */
    if (!((run >> s) & 1ull)) return;
    const StepArgs<R> a = global_args(table[b]);
    int q = (tile - sched[2 + b]) * (4 * EW) + wave * EW + lane / G;
    if (G == 64 && VEL) q = __builtin_amdgcn_readfirstlane(q); /* one element per wavefront */
    if (q >= a.tri) return;
    const int t = *step_base + step_off - a.shot_start[s];
    if (VEL) velocity_body<R, LG>(op, a, q, s, t);
    else stress_body<R, LG>(op, a, q, s, t);
}

/* dst[pairs[k].x] = pairs[k].y */
__global__ void k_scatter_int(int *dst, const int2 *pairs, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[pairs[k].x] = pairs[k].y;
}

/* zero the field columns of the slots in `mask` (of a G-slot group) before they are reused */
template <typename R>
__global__ void k_clear_slots(R *f0, R *f1, R *f2, R *f3, R *f4, long long n_nodes, int G, unsigned long long mask)
{
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n_nodes * G;
         k += (long long)gridDim.x * blockDim.x) {
        if (!((mask >> (k % G)) & 1ull)) continue;
        f0[k] = 0; f1[k] = 0; f2[k] = 0; f3[k] = 0; f4[k] = 0;
    }
}

/* Slot lanes moved between two G-slot groups of one mesh (serving compaction). */
template <typename R>
struct MoveArgs {
    const R *src[5];
    R *dst[5];
    int from[64], to[64];
};

/* dst[c][node*G + to[k]] = src[c][node*G + from[k]] for every field c, node and move k < n: the moved
 * shots continue bit for bit in their new slots */
template <typename R>
__global__ void k_move_slots(MoveArgs<R> m, long long n_nodes, int G, int n)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n_nodes * n;
         t += (long long)gridDim.x * blockDim.x) {
        const long long node = t / n;
        const int k = (int)(t - node * n);
        const long long s = node * G + m.from[k], d = node * G + m.to[k];
#pragma unroll
        for (int c = 0; c < 5; c++) m.dst[c][d] = m.src[c][s];
    }
}

} // namespace dgxb

template <typename R>
static hipError_t launch_move(void *const *src, void *const *dst, long long n_nodes, int group, int n, const int *from,
                              const int *to, hipStream_t st)
{
    dgxb::MoveArgs<R> m;
    for (int c = 0; c < 5; c++) {
        m.src[c] = (const R *)src[c];
        m.dst[c] = (R *)dst[c];
    }
    for (int k = 0; k < 64; k++) {
        m.from[k] = k < n ? from[k] : 0;
        m.to[k] = k < n ? to[k] : 0;
    }
    dgxb::k_move_slots<R><<<2048, 256, 0, st>>>(m, n_nodes, group, n);
    return hipGetLastError();
}

template <typename R, int ST, int NW>
static hipError_t launch_b(const dgx::StepArgs<R> &a, hipStream_t st, int which)
{
    dim3 grid(DGX_NXCD * a.nper, a.spad / ST), block(64 * NW);
    if constexpr (ST < 64) {
        if (a.il) {
            if (a.spad != ST || (1 << a.il) != 64 / ST) return hipErrorInvalidValue;
            if (which == 0) dgxb::kb_velocity<R, ST, NW, true><<<grid, block, 0, st>>>(a);
            else dgxb::kb_stress<R, ST, NW, true><<<grid, block, 0, st>>>(a);
            return hipGetLastError();
        }
    }
    if (a.il) return hipErrorInvalidValue;
    if (which == 0) dgxb::kb_velocity<R, ST, NW, false><<<grid, block, 0, st>>>(a);
    else dgxb::kb_stress<R, ST, NW, false><<<grid, block, 0, st>>>(a);
    return hipGetLastError();
}

/* The tile shapes the shot-innermost kernels are built for: shot tiles 4 and 8 with 4 wavefronts per
 * workgroup, and 16, 32, 64 with 2, 4 or 8 (dgx_batched_shape_ok lists the same set for the host). */
template <typename R, int ST>
static hipError_t launch_bw(const dgx::StepArgs<R> &a, hipStream_t st, int which)
{
    switch (a.wg_waves) {
    case 2: return launch_b<R, ST, 2>(a, st, which);
    case 4: return launch_b<R, ST, 4>(a, st, which);
    case 8: return launch_b<R, ST, 8>(a, st, which);
    }
    return hipErrorInvalidValue;
}

template <typename R>
static hipError_t dispatch_b(const dgx::StepArgs<R> &a, hipStream_t st, int which)
{
    switch (a.shot_tile) {
    case 4: return a.wg_waves == 4 ? launch_b<R, 4, 4>(a, st, which) : hipErrorInvalidValue;
    case 8: return a.wg_waves == 4 ? launch_b<R, 8, 4>(a, st, which) : hipErrorInvalidValue;
    case 16: return launch_bw<R, 16>(a, st, which);
    case 32: return launch_bw<R, 32>(a, st, which);
    case 64: return launch_bw<R, 64>(a, st, which);
    }
    return hipErrorInvalidValue;
}

template <typename R>
static hipError_t launch_serve(const void *ops, const void *table, const int *sched,
                               const unsigned long long *runmask, const int *step_base, int step_off, int vel,
                               int group, int nb, int total, hipStream_t st)
{
    const dgx::Ops<R> &op = *(const dgx::Ops<R> *)ops;
    const dgx::StepArgs<R> *t = (const dgx::StepArgs<R> *)table;
    if (total <= 0) return hipSuccess;
    const int nper = (total + DGX_NXCD - 1) / DGX_NXCD; /* tiles per XCD */
    const int grid = DGX_NXCD * nper;
    if (group == 4) {
        if (vel) dgxb::ks_step<R, true, 4><<<grid, DGX_BLOCK, 0, st>>>(op, t, sched, runmask, step_base, step_off, nb, total, nper);
        else dgxb::ks_step<R, false, 4><<<grid, DGX_BLOCK, 0, st>>>(op, t, sched, runmask, step_base, step_off, nb, total, nper);
    } else if (group == 8) {
        if (vel) dgxb::ks_step<R, true, 8><<<grid, DGX_BLOCK, 0, st>>>(op, t, sched, runmask, step_base, step_off, nb, total, nper);
        else dgxb::ks_step<R, false, 8><<<grid, DGX_BLOCK, 0, st>>>(op, t, sched, runmask, step_base, step_off, nb, total, nper);
    } else if (group == 16) {
        if (vel) dgxb::ks_step<R, true, 16><<<grid, DGX_BLOCK, 0, st>>>(op, t, sched, runmask, step_base, step_off, nb, total, nper);
        else dgxb::ks_step<R, false, 16><<<grid, DGX_BLOCK, 0, st>>>(op, t, sched, runmask, step_base, step_off, nb, total, nper);
    } else if (group == 32) {
        if (vel) dgxb::ks_step<R, true, 32><<<grid, DGX_BLOCK, 0, st>>>(op, t, sched, runmask, step_base, step_off, nb, total, nper);
        else dgxb::ks_step<R, false, 32><<<grid, DGX_BLOCK, 0, st>>>(op, t, sched, runmask, step_base, step_off, nb, total, nper);
    } else if (group == 64) {
        if (vel) dgxb::ks_step<R, true, 64><<<grid, DGX_BLOCK, 0, st>>>(op, t, sched, runmask, step_base, step_off, nb, total, nper);
        else dgxb::ks_step<R, false, 64><<<grid, DGX_BLOCK, 0, st>>>(op, t, sched, runmask, step_base, step_off, nb, total, nper);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

extern "C" {
hipError_t dgx_launch_serve(int fp32, const void *ops, const void *table, const int *sched,
                            const unsigned long long *runmask, const int *step_base, int step_off, int vel, int group,
                            int nb, int total, hipStream_t st)
{
    /* the operators are the same for every mesh: they travel as a kernel argument (scalar registers),
     * not through the batch table */
    if (fp32) return launch_serve<float>(ops, table, sched, runmask, step_base, step_off, vel, group, nb, total, st);
    return launch_serve<double>(ops, table, sched, runmask, step_base, step_off, vel, group, nb, total, st);
}
hipError_t dgx_launch_scatter(int *dst, const int2 *pairs, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    dgxb::k_scatter_int<<<(n + 255) / 256, 256, 0, st>>>(dst, pairs, n);
    return hipGetLastError();
}
hipError_t dgx_launch_clear_slots(int fp32, void *const *f, long long n_nodes, int group, unsigned long long mask,
                                  hipStream_t st)
{
    if (fp32) dgxb::k_clear_slots<float><<<1024, 256, 0, st>>>((float *)f[0], (float *)f[1], (float *)f[2], (float *)f[3], (float *)f[4], n_nodes, group, mask);
    else dgxb::k_clear_slots<double><<<1024, 256, 0, st>>>((double *)f[0], (double *)f[1], (double *)f[2], (double *)f[3], (double *)f[4], n_nodes, group, mask);
    return hipGetLastError();
}
hipError_t dgx_launch_move_slots(int fp32, void *const *src, void *const *dst, long long n_nodes, int group, int n,
                                 const int *from, const int *to, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (n > 64 || n > group) return hipErrorInvalidValue;
    for (int k = 0; k < n; k++)
        if (from[k] < 0 || from[k] >= group || to[k] < 0 || to[k] >= group) return hipErrorInvalidValue;
    return fp32 ? launch_move<float>(src, dst, n_nodes, group, n, from, to, st)
                : launch_move<double>(src, dst, n_nodes, group, n, from, to, st);
}
int dgx_batched_shape_ok(int shot_tile, int wg_waves)
{
    if (shot_tile == 4 || shot_tile == 8) return wg_waves == 4;
    if (shot_tile == 16 || shot_tile == 32 || shot_tile == 64) return wg_waves == 2 || wg_waves == 4 || wg_waves == 8;
    return 0;
}
hipError_t dgx_launch_batched_f64(const void *args, hipStream_t st, int which)
{
    return dispatch_b(*(const dgx::StepArgs<double> *)args, st, which);
}
hipError_t dgx_launch_batched_f32(const void *args, hipStream_t st, int which)
{
    return dispatch_b(*(const dgx::StepArgs<float> *)args, st, which);
}
}
