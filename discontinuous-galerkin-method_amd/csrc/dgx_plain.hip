/*
 * dgx_plain.hip -- element-per-thread kernels of the velocity and stress half-steps (gfx950).
 *
 * One thread advances one element of one shot.  The reference-element operators Dr = Vr*inverV,
 * Ds = Vs*inverV and the face lifts S1..S3 arrive as kernel arguments (uniform, so they sit in
 * scalar registers / the scalar cache).  Geometric factors are recomputed in FP64 from the three
 * vertex coordinates (dgx_device.h), so the reference's per-element Dx/Dz/Q arrays (fun_DG.c:1091-1192)
 * are never materialised.
 *
 * Velocity (OpenMP_Function.c:75-185), with the material factored out and the Dx/Dz products
 * regrouped as Dr(rx u + rz w) + Ds(sx u + sz w):
 *   Vx += dt/rho [ Dr(rx Txx + rz Txz) + Ds(sx Txx + sz Txz) + 1/D sum_j S_j (ez_j dTxx_j - ex_j dTxz_j) ]
 *   Vz += dt/rho [ Dr(rx Txz + rz Tzz) + Ds(sx Txz + sz Tzz) + 1/D sum_j S_j (ez_j dTxz_j - ex_j dTzz_j) ]
 * Stress (OpenMP_Function.c:188-306) in strain-rate form:
 *   exx = Dx Vx + 1/D sum_j S_j(ez_j dVx_j),  ezz = Dz Vz + 1/D sum_j S_j(-ex_j dVz_j),
 *   exz = Dx Vz + Dz Vx + 1/D sum_j S_j(ez_j dVz_j - ex_j dVx_j)
 *   Txx += dt[(l+2m) exx + l ezz],  Tzz += dt[l exx + (l+2m) ezz],  Txz += dt m exz
 * The face jump is dU = U_nb - U_own (U_nb = 0 on the outer boundary, fun_DG.c:724-736).
 *
 * Fused variants fold the rest of the reference step into these two kernels: the sponge multiply
 * (OpenMP_Function.c:25-46) is applied to every value as it is loaded, the source increment
 * (OpenMP_Function.c:49-58) right after it, and the receiver pick (OpenMP_Function.c:61-72) is done by
 * the stress kernel from the velocities it already holds.  Each field is read once per half-step.
 * The half-step kernels stage the values their workgroup's 256 elements hold (conditioned stresses for the
 * velocity update, new velocities for the stress update) in LDS, and read the face neighbours that lie in
 * the workgroup from there; only the others are gathered from global memory.
 * k_fullstep runs both half-steps in one launch (opt-in; see its comment).
 */
#include <hip/hip_runtime.h>

#include <mutex>

#include "dgx.h"
#include "dgx_device.h"
#include "dgx_kernels.h"
#include "dgx_p2ops.h"

namespace dgx {

__device__ __forceinline__ int fidx(int q, int i) { return ((q >> 6) * 6 + i) * 64 + (q & 63); }

/* Reference-element operators.  P2: the P2 Fekete operators of the shipped operator files, compiled in
 * as immediates (the host selects these kernels only when the caller's operators equal them bit for
 * bit; results agree with the argument variant to rounding, since the compiler may fuse a different
 * product into an FMA); otherwise the caller's operators from the kernel argument.  As arguments, the
 * 126 FP64 values do not fit the scalar registers and the FP64 velocity kernel spills about 200 of them
 * to vector-register lanes; compiled in, it spills none.  The stress update compiles in only the face
 * lifts: with all operators compiled in its FP64 kernel needs more than the 168 VGPRs of 3 waves per
 * SIMD, with the lifts alone it does not and runs faster in both precisions. */
template <bool P2, typename R>
__device__ __forceinline__ R op_dr(const StepArgs<R> &a, int k)
{
    if constexpr (P2) return (R)dgx_p2_ops[k];
    else return a.op.Dr[k];
}
template <bool P2, typename R>
__device__ __forceinline__ R op_ds(const StepArgs<R> &a, int k)
{
    if constexpr (P2) return (R)dgx_p2_ops[36 + k];
    else return a.op.Ds[k];
}
template <bool P2, typename R>
__device__ __forceinline__ R op_l(const StepArgs<R> &a, int k)
{
    if constexpr (P2) return (R)dgx_p2_ops[72 + k];
    else return a.op.L[k];
}

/* Whether the blocked kernel (SPARSE) leaves out the product with Dr (k < 36) or Ds (k >= 36) entry k of the
 * compiled-in P2 table: exact zeros in both precisions (their products add nothing), and in FP32 also the
 * entries below 1e-6 in magnitude.  Those are noise of the operator files on structural zeros (6e-10 to 2e-8);
 * each is smaller than half an FP32 ulp of the O(1) entries beside it (6e-8), so FP32 cannot tell them from
 * zero.  Of the 72 entries, FP64 skips 8 and FP32 28. */
template <typename R, bool SPARSE>
__device__ __forceinline__ constexpr bool p2_skip(int k)
{
    if constexpr (!SPARSE) return false;
    const double v = dgx_p2_ops[k];
    return v == 0.0 || (sizeof(R) == 4 && v > -1e-6 && v < 1e-6);
}

/* Workgroups are dealt to the 8 XCDs round-robin; give each XCD one contiguous slice of the
 * (Hilbert-ordered) elements so that face neighbours mostly share that XCD's L2. */
__device__ __forceinline__ int element_of_thread(int nper)
{
    const int b = blockIdx.x;
    return ((b & (DGX_NXCD - 1)) * nper + (b >> 3)) * DGX_BLOCK + threadIdx.x;
}

/* Sponge and explosive source applied to a stress value of element q at node i (fused step). */
template <typename R>
__device__ __forceinline__ void condition_stress(const StepArgs<R> &a, unsigned code, int q, int i, int shot, int step,
                                                 R &txx, R &tzz, R &txz)
{
    if (code & DGX_CODE_SPONGE) {
        const R f = a.absorb[fidx(q, i)];
        txx = sponge(txx, f); tzz = sponge(tzz, f); txz = sponge(txz, f);
    }
    const unsigned ax = code >> DGX_CODE_AUX_SHIFT;
    if (ax) {
        const dgx_aux x = a.aux[ax - 1];
        for (int s = x.src_begin; s < x.src_begin + x.src_count; s++) {
            const dgx_src_dev &src = a.src[s];
            if (src.shot != shot || src.kind != 1) continue;
            const double w = a.wav[(long long)src.wavelet * a.nsteps + step];
            const R inc = (R)(src.fun[i] * w);
            txx += inc;
            tzz += inc;
        }
    }
}

/* condition_stress for all 6 nodes of element q: one sponge branch and one source branch per element */
template <typename R>
__device__ __forceinline__ void condition_stress6(const StepArgs<R> &a, unsigned code, int q, int shot, int step,
                                                  R *txx, R *tzz, R *txz)
{
    if (code & DGX_CODE_SPONGE) {
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const R f = a.absorb[fidx(q, i)];
            txx[i] = sponge(txx[i], f); tzz[i] = sponge(tzz[i], f); txz[i] = sponge(txz[i], f);
        }
    }
    const unsigned ax = code >> DGX_CODE_AUX_SHIFT;
    if (ax) {
        const dgx_aux x = a.aux[ax - 1];
        for (int s = x.src_begin; s < x.src_begin + x.src_count; s++) {
            const dgx_src_dev &src = a.src[s];
            if (src.shot != shot || src.kind != 1) continue;
            const double w = a.wav[(long long)src.wavelet * a.nsteps + step];
#pragma unroll
            for (int i = 0; i < 6; i++) {
                const R inc = (R)(src.fun[i] * w);
                txx[i] += inc;
                tzz[i] += inc;
            }
        }
    }
}

/* ---- arithmetic shared by the half-step and full-step updates ---- */

/* Own stresses of element q conditioned (sponge, explosive source) as the reference leaves them before
 * LU_Velocity reads them: per element (one branch each for sponge and sources) in FP32 and with
 * compiled-in operators; per node in FP64 with argument operators, where the per-element form needs more
 * registers than the 168 that keep 3 wavefronts per SIMD. */
template <typename R, bool P2>
__device__ __forceinline__ void condition_own_stresses(const StepArgs<R> &a, unsigned code, int q, int shot, int step,
                                                       R (&txx)[6], R (&tzz)[6], R (&txz)[6])
{
    if constexpr (sizeof(R) == 4 || P2) {
        condition_stress6(a, code, q, shot, step, txx, tzz, txz);
    } else {
#pragma unroll
        for (int i = 0; i < 6; i++) condition_stress(a, code, q, i, shot, step, txx[i], tzz[i], txz[i]);
    }
}

/* Neighbour stresses facing own face j, position k, from global memory.  STAGED: only for the faces whose
 * element lies outside the workgroup's elements q0 .. q0+DGX_BLOCK-1 (local[j] false; the others are read
 * from LDS later and their slots start at 0); otherwise for every face (local[j] = false).  A boundary
 * face reads the own element and its value is dropped, so that all loads are issued together. */
template <bool STAGED, typename R>
__device__ __forceinline__ void gather_stresses(unsigned code, const int (&nbs)[3], int q, int q0, bool (&local)[3],
                                                const R *Txx, const R *Tzz,
                                                const R *Txz, R (&gxx)[3][3], R (&gzz)[3][3],
                                                R (&gxz)[3][3])
{
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int ee = nbs[j] >= 0 ? nbs[j] : q;
        local[j] = STAGED && (unsigned)(ee - q0) < (unsigned)DGX_BLOCK;
        const unsigned m = (code >> (3 * j)) & 3u, rev = (code >> (3 * j + 2)) & 1u;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            gxx[j][k] = 0; gzz[j][k] = 0; gxz[j][k] = 0;
            if (!local[j]) {
                const int o = fidx(ee, nb_node(m, rev, k));
                gxx[j][k] = Txx[o]; gzz[j][k] = Tzz[o]; gxz[j][k] = Txz[o];
            }
        }
    }
}

/* Sponge and explosive source applied to the gathered stresses of neighbours that carry either (the code
 * word's NBSPECIAL bits); faces with local[j] true hold values their own thread conditioned already. */
template <typename R>
__device__ __forceinline__ void condition_neighbour_stresses(const StepArgs<R> &a, unsigned code, const int (&nbs)[3],
                                                             const bool (&local)[3], int shot, int step,
                                                             R (&gxx)[3][3], R (&gzz)[3][3], R (&gxz)[3][3])
{
    if (!(code & (DGX_CODE_NBSPECIAL(0) | DGX_CODE_NBSPECIAL(1) | DGX_CODE_NBSPECIAL(2)))) return;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        if (!(code & DGX_CODE_NBSPECIAL(j)) || nbs[j] < 0 || local[j]) continue;
        const unsigned m = (code >> (3 * j)) & 3u, rev = (code >> (3 * j + 2)) & 1u;
        const unsigned ncode = (unsigned)a.nbr[nbs[j]].w;
        for (int k = 0; k < 3; k++)
            condition_stress(a, ncode, nbs[j], nb_node(m, rev, k), shot, step, gxx[j][k], gzz[j][k], gxz[j][k]);
    }
}

/* The P2 derivatives in factored form.  The interpolant of six nodal values u on the FeketeNew nodes
 * (r, s) = (1,-1), (0,0), (-1,1), (0,-1), (-1,0), (-1,-1) is quadratic, so its derivatives are linear functions,
 * du/dr = r0 + r1 r + r2 s and du/ds = s0 + s1 r + s2 s, with
 *   r2 = s1 = u5 + u1 - u3 - u4;  X = u0 - u3 + r2, Y = u4 - u1: r0 = (X - Y)/2, r1 = X + Y;
 *   Z = u2 - u4 + r2, W = u3 - u1: s0 = (Z - W)/2, s2 = Z + W
 * (15 operations for both), and a linear function takes 6 operations to evaluate at the six nodes.  The
 * operator files' Dr = Vr inverV and Ds = Vs inverV are these exact derivatives (entries 0, 1/2, 1, 3/2, 2 and
 * their negatives) to within 4e-8, below FP32 resolution: rounded to FP32, 69 of their 72 entries equal the
 * exact values and three are one ulp off.  The FP32 blocked kernel with the shipped operators therefore applies
 * them in this form, 45 + 24 + 12 operations for the velocity volume term instead of 48 + 88 for the sparse
 * matrix form (the compiled-in matrices keep FP64, which must match the files' values, and compensated FP32). */
template <typename R>
struct P2Lin {
    R c0, c1, c2; /* c0 + c1 r + c2 s */
};

template <typename R>
__device__ __forceinline__ void p2_grad(const R (&u)[6], P2Lin<R> &dr, P2Lin<R> &ds)
{
    const R a4 = (u[5] + u[1]) - (u[3] + u[4]);
    const R X = (u[0] - u[3]) + a4, Y = u[4] - u[1];
    const R Z = (u[2] - u[4]) + a4, W = u[3] - u[1];
    dr = P2Lin<R>{(X - Y) * (R)0.5, X + Y, a4};
    ds = P2Lin<R>{(Z - W) * (R)0.5, a4, Z + W};
}

/* p x + q y + s z + t w, coefficient by coefficient */
template <typename R>
__device__ __forceinline__ P2Lin<R> p2_comb(R p, const P2Lin<R> &x, R q, const P2Lin<R> &y, R s,
                                            const P2Lin<R> &z, R t, const P2Lin<R> &w)
{
    return P2Lin<R>{p * x.c0 + q * y.c0 + s * z.c0 + t * w.c0, p * x.c1 + q * y.c1 + s * z.c1 + t * w.c1,
                    p * x.c2 + q * y.c2 + s * z.c2 + t * w.c2};
}

/* p x + q y */
template <typename R>
__device__ __forceinline__ P2Lin<R> p2_comb(R p, const P2Lin<R> &x, R q, const P2Lin<R> &y)
{
    return P2Lin<R>{p * x.c0 + q * y.c0, p * x.c1 + q * y.c1, p * x.c2 + q * y.c2};
}

/* c0 + c1 r + c2 s at the six nodes */
template <typename R>
__device__ __forceinline__ void p2_eval(const P2Lin<R> &c, R (&out)[6])
{
    const R m = c.c0 - c.c1;
    out[0] = (c.c0 + c.c1) - c.c2;
    out[1] = c.c0;
    out[2] = m + c.c2;
    out[3] = c.c0 - c.c2;
    out[4] = m;
    out[5] = m - c.c2;
}

/* Velocity volume term: rhx = Dr(rx Txx + rz Txz) + Ds(sx Txx + sz Txz), rhz = Dr(rx Txz + rz Tzz) +
 * Ds(sx Txz + sz Tzz).  FACTORED (FP32 with the compiled-in P2 operators): the same terms from the factored
 * derivatives, rhx = rx dTxx/dr + sx dTxx/ds + rz dTxz/dr + sz dTxz/ds and likewise rhz. */
template <typename R, bool P2, typename G = Geom, bool SPARSE = false, bool FACTORED = false>
__device__ __forceinline__ void velocity_volume(const StepArgs<R> &a, const G &g, const R (&txx)[6],
                                                const R (&tzz)[6], const R (&txz)[6], R (&rhx)[6], R (&rhz)[6])
{
    const R rx = (R)g.rx, rz = (R)g.rz, sx = (R)g.sx, sz = (R)g.sz;
    if constexpr (FACTORED && P2 && sizeof(R) == 4) {
        P2Lin<R> rxx, sxx, rxz, sxz, rzz, szz;
        p2_grad(txx, rxx, sxx);
        p2_grad(txz, rxz, sxz);
        p2_grad(tzz, rzz, szz);
        p2_eval(p2_comb(rx, rxx, sx, sxx, rz, rxz, sz, sxz), rhx);
        p2_eval(p2_comb(rx, rxz, sx, sxz, rz, rzz, sz, szz), rhz);
        return;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) { rhx[i] = 0; rhz[i] = 0; }
#pragma unroll
    for (int n = 0; n < 6; n++) {
        const R ux = rx * txx[n] + rz * txz[n], wx = sx * txx[n] + sz * txz[n];
        const R uz = rx * txz[n] + rz * tzz[n], wz = sx * txz[n] + sz * tzz[n];
#pragma unroll
        for (int i = 0; i < 6; i++) { /* one accumulation per product: two fused multiply-adds, no separate add */
            if (!p2_skip<R, SPARSE && P2>(i * 6 + n)) {
                rhx[i] += op_dr<P2>(a, i * 6 + n) * ux;
                rhz[i] += op_dr<P2>(a, i * 6 + n) * uz;
            }
            if (!p2_skip<R, SPARSE && P2>(36 + i * 6 + n)) {
                rhx[i] += op_ds<P2>(a, i * 6 + n) * wx;
                rhz[i] += op_ds<P2>(a, i * 6 + n) * wz;
            }
        }
    }
}

/* Face j of the velocity surface term: own stresses at the face's nodes (o*[k] at face_node(j, k)) and the
 * neighbour's facing them (n*[k]; ignored on the outer boundary, has false), accumulated into flx, flz. */
template <typename R, bool P2, typename G = Geom>
__device__ __forceinline__ void velocity_flux_face(const StepArgs<R> &a, const G &g, int j, bool has,
                                                   const R (&oxx)[3], const R (&ozz)[3], const R (&oxz)[3],
                                                   const R (&nxx)[3], const R (&nzz)[3], const R (&nxz)[3],
                                                   R (&flx)[6], R (&flz)[6])
{
    const R ex = (R)g.ex[j], ez = (R)g.ez[j];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        R dxx = -oxx[k], dzz = -ozz[k], dxz = -oxz[k];
        if (has) { dxx += nxx[k]; dzz += nzz[k]; dxz += nxz[k]; }
        const R gx = ez * dxx - ex * dxz;
        const R gz = ez * dxz - ex * dzz;
#pragma unroll
        for (int i = 0; i < 6; i++) {
            flx[i] += op_l<P2>(a, j * 18 + i * 3 + k) * gx;
            flz[i] += op_l<P2>(a, j * 18 + i * 3 + k) * gz;
        }
    }
}

/* Velocity surface term before its 1/D scale: sum_j S_j (ez_j dTxx_j - ex_j dTxz_j) and
 * sum_j S_j (ez_j dTxz_j - ex_j dTzz_j), the jump dT = T_nb - T_own (T_nb = 0 on the outer boundary). */
template <typename R, bool P2, typename G = Geom>
__device__ __forceinline__ void velocity_flux(const StepArgs<R> &a, const G &g, const int (&nbs)[3],
                                              const R (&txx)[6], const R (&tzz)[6], const R (&txz)[6],
                                              const R (&gxx)[3][3], const R (&gzz)[3][3], const R (&gxz)[3][3],
                                              R (&flx)[6], R (&flz)[6])
{
#pragma unroll
    for (int i = 0; i < 6; i++) { flx[i] = 0; flz[i] = 0; }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        R oxx[3], ozz[3], oxz[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int on = face_node(j, k);
            oxx[k] = txx[on]; ozz[k] = tzz[on]; oxz[k] = txz[on];
        }
        velocity_flux_face<R, P2>(a, g, j, nbs[j] >= 0, oxx, ozz, oxz, gxx[j], gzz[j], gxz[j], flx, flz);
    }
}

/* Sponge and force source applied to the 6 velocity nodes of element q (FP32 path: one branch each per
 * element). */
template <typename R>
__device__ __forceinline__ void condition_velocity6(const StepArgs<R> &a, unsigned code, int q, int shot, int step,
                                                    R (&vx)[6], R (&vz)[6])
{
    if (code & DGX_CODE_SPONGE) {
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const R f = a.absorb[fidx(q, i)];
            vx[i] = sponge(vx[i], f);
            vz[i] = sponge(vz[i], f);
        }
    }
    const unsigned ax = code >> DGX_CODE_AUX_SHIFT;
    if (ax) {
        const dgx_aux x = a.aux[ax - 1];
        for (int s = x.src_begin; s < x.src_begin + x.src_count; s++) {
            const dgx_src_dev &src = a.src[s];
            if (src.shot != shot || src.kind != 0) continue;
            const double w = a.wav[(long long)src.wavelet * a.nsteps + step];
#pragma unroll
            for (int i = 0; i < 6; i++) vz[i] += (R)(src.fun[i] * w);
        }
    }
}

/* Sponge and force source applied to velocity node i of element q, at offset o = fidx(q, i) (FP64 path,
 * per node). */
template <typename R>
__device__ __forceinline__ void condition_velocity(const StepArgs<R> &a, unsigned code, int o, int i, int shot,
                                                   int step, R &vx, R &vz)
{
    if (code & DGX_CODE_SPONGE) {
        const R f = a.absorb[o];
        vx = sponge(vx, f);
        vz = sponge(vz, f);
    }
    const unsigned ax = code >> DGX_CODE_AUX_SHIFT;
    if (ax) {
        const dgx_aux x = a.aux[ax - 1];
        for (int s = x.src_begin; s < x.src_begin + x.src_count; s++) {
            const dgx_src_dev &src = a.src[s];
            if (src.shot == shot && src.kind == 0)
                vz += (R)(src.fun[i] * a.wav[(long long)src.wavelet * a.nsteps + step]);
        }
    }
}

/* Neighbour velocities facing own face j, position k, from global memory (as gather_stresses). */
template <bool STAGED, typename R>
__device__ __forceinline__ void gather_velocities(unsigned code, const int (&nbs)[3], int q, int q0, bool (&local)[3],
                                                  const R *Vx, const R *Vz,
                                                  R (&gvx)[3][3], R (&gvz)[3][3])
{
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int ee = nbs[j] >= 0 ? nbs[j] : q;
        local[j] = STAGED && (unsigned)(ee - q0) < (unsigned)DGX_BLOCK;
        const unsigned m = (code >> (3 * j)) & 3u, rev = (code >> (3 * j + 2)) & 1u;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            gvx[j][k] = 0; gvz[j][k] = 0;
            if (!local[j]) {
                const int o = fidx(ee, nb_node(m, rev, k));
                gvx[j][k] = Vx[o]; gvz[j][k] = Vz[o];
            }
        }
    }
}

/* Strain-rate volume terms: exx = Dx Vx, ezz = Dz Vz, exz = Dx Vz + Dz Vx with Dx = rx Dr + sx Ds and
 * Dz = rz Dr + sz Ds (operators from the kernel argument in both variants). */
template <typename R, typename G = Geom, bool P2 = false, bool SPARSE = false, bool FACTORED = false>
__device__ __forceinline__ void strain_volume(const StepArgs<R> &a, const G &g, const R (&vx)[6],
                                              const R (&vz)[6], R (&exx)[6], R (&ezz)[6], R (&exz)[6])
{
    const R rx = (R)g.rx, rz = (R)g.rz, sx = (R)g.sx, sz = (R)g.sz;
    if constexpr (FACTORED && P2 && sizeof(R) == 4) { /* the factored P2 derivatives, as in velocity_volume */
        P2Lin<R> rvx, svx, rvz, svz;
        p2_grad(vx, rvx, svx);
        p2_grad(vz, rvz, svz);
        p2_eval(p2_comb(rx, rvx, sx, svx), exx);
        p2_eval(p2_comb(rz, rvz, sz, svz), ezz);
        p2_eval(p2_comb(rx, rvz, sx, svz, rz, rvx, sz, svx), exz);
        return;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) {
        R rvx = 0, svx = 0, rvz = 0, svz = 0;
#pragma unroll
        for (int n = 0; n < 6; n++) {
            if (!p2_skip<R, SPARSE && P2>(i * 6 + n)) {
                rvx += op_dr<P2>(a, i * 6 + n) * vx[n];
                rvz += op_dr<P2>(a, i * 6 + n) * vz[n];
            }
            if (!p2_skip<R, SPARSE && P2>(36 + i * 6 + n)) {
                svx += op_ds<P2>(a, i * 6 + n) * vx[n];
                svz += op_ds<P2>(a, i * 6 + n) * vz[n];
            }
        }
        exx[i] = rx * rvx + sx * svx;
        ezz[i] = rz * rvz + sz * svz;
        exz[i] = rx * rvz + sx * svz + rz * rvx + sz * svx;
    }
}

/* Strain-rate surface terms before their 1/D scale: sum_j S_j(ez_j dVx_j), sum_j S_j(-ex_j dVz_j) and
 * sum_j S_j(ez_j dVz_j - ex_j dVx_j). */
/* Face j of the strain-rate surface terms (own velocities vx, vz; the neighbour's facing face j's nodes in
 * gvx[k], gvz[k], ignored on the outer boundary), accumulated into fx, fz, fxz. */
template <typename R, bool P2, typename G = Geom>
__device__ __forceinline__ void stress_flux_face(const StepArgs<R> &a, const G &g, int j, bool has, const R (&vx)[6],
                                                 const R (&vz)[6], const R (&gvx)[3], const R (&gvz)[3], R (&fx)[6],
                                                 R (&fz)[6], R (&fxz)[6])
{
    const R ex = (R)g.ex[j], ez = (R)g.ez[j];
    R px[3], pz[3], pxz[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int on = face_node(j, k);
        R dvx = -vx[on], dvz = -vz[on];
        if (has) { dvx += gvx[k]; dvz += gvz[k]; }
        px[k] = ez * dvx;
        pz[k] = -ex * dvz;
        pxz[k] = ez * dvz - ex * dvx;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const R l = op_l<P2>(a, j * 18 + i * 3 + k);
            fx[i] += l * px[k];
            fz[i] += l * pz[k];
            fxz[i] += l * pxz[k];
        }
    }
}

template <typename R, bool P2, typename G = Geom>
__device__ __forceinline__ void stress_flux(const StepArgs<R> &a, const G &g, const int (&nbs)[3],
                                            const R (&vx)[6], const R (&vz)[6], const R (&gvx)[3][3],
                                            const R (&gvz)[3][3], R (&fx)[6], R (&fz)[6], R (&fxz)[6])
{
#pragma unroll
    for (int i = 0; i < 6; i++) { fx[i] = 0; fz[i] = 0; fxz[i] = 0; }
#pragma unroll
    for (int j = 0; j < 3; j++) stress_flux_face<R, P2>(a, g, j, nbs[j] >= 0, vx, vz, gvx[j], gvz[j], fx, fz, fxz);
}

/* Txx += dt[(l+2m) exx + l ezz], Tzz += dt[l exx + (l+2m) ezz], Txz += dt m exz for the 6 nodes of element
 * q (strain rates = volume + surface/D).  The old stresses come conditioned from `tload` (LDS,
 * (c*6 + i)*DGX_BLOCK + thread) when it is given, else from T, conditioned here when FUSED. */
template <typename R, bool FUSED>
__device__ __forceinline__ void stress_store(const StepArgs<R> &a, unsigned code, int q, int shot, int step,
                                             const Geom &g, const R (&exx)[6], const R (&ezz)[6],
                                             const R (&exz)[6], const R (&fx)[6], const R (&fz)[6],
                                             const R (&fxz)[6], const R *tload, R *Txx,
                                             R *Tzz, R *Txz)
{
    const R invD = (R)g.invD;
    const R lam = a.lam[q], mu = a.mu[q], l2m = lam + 2 * mu, dt = (R)a.dt;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int o = fidx(q, i);
        R txx, tzz, txz;
        if (tload) {
            txx = tload[(0 * 6 + i) * DGX_BLOCK + threadIdx.x];
            tzz = tload[(1 * 6 + i) * DGX_BLOCK + threadIdx.x];
            txz = tload[(2 * 6 + i) * DGX_BLOCK + threadIdx.x];
        } else {
            txx = Txx[o]; tzz = Tzz[o]; txz = Txz[o];
            if (FUSED) condition_stress(a, code, q, i, shot, step, txx, tzz, txz);
        }
        const R axx = exx[i] + fx[i] * invD, azz = ezz[i] + fz[i] * invD, axz = exz[i] + fxz[i] * invD;
        Txx[o] = (l2m * axx + lam * azz) * dt + txx;
        Tzz[o] = (lam * axx + l2m * azz) * dt + tzz;
        Txz[o] = (mu * axz) * dt + txz;
    }
}

/* Receivers of the element (OpenMP_Function.c:61-72 with per-node weights) from the new velocities it holds:
 * trace slot `step` of shot `shot`. */
template <typename R>
__device__ __forceinline__ void record_receivers(const StepArgs<R> &a, unsigned code, int shot, int step,
                                                 const R (&vx)[6], const R (&vz)[6])
{
    const unsigned ax = code >> DGX_CODE_AUX_SHIFT;
    if (!ax || step >= a.nsteps) return;
    const dgx_aux x = a.aux[ax - 1];
    for (int r = x.rec_begin; r < x.rec_begin + x.rec_count; r++) {
        const dgx_rec_dev &rc = a.rec[r];
        double sx_ = 0, sz_ = 0;
#pragma unroll
        for (int i = 0; i < 6; i++) {
            sx_ += rc.w[i] * (double)vx[i];
            sz_ += rc.w[i] * (double)vz[i];
        }
        double *t = a.trace + ((long long)shot * a.nsteps + step) * 2 * a.nrec;
        t[rc.index] = sx_;
        t[a.nrec + rc.index] = sz_;
    }
}

/* ---- the updates ---- */

/*
 * Velocity update of element q (full-step kernel, fused): stresses T and velocities Vi in, the new velocities
 * written to row `orow` of Vo; when `save`, the conditioned stresses of the element are also kept in LDS
 * (tsave[(c*6 + i)*DGX_BLOCK + threadIdx.x], c = xx, zz, xz) for the stress phase.
 */
template <typename R, bool P2>
__device__ __forceinline__ void velocity_update(const StepArgs<R> &a, int q, int shot, int step, int orow,
                                                const R *__restrict__ Txx, const R *__restrict__ Tzz,
                                                const R *__restrict__ Txz, const R *__restrict__ Vxi,
                                                const R *__restrict__ Vzi, R *__restrict__ Vxo,
                                                R *__restrict__ Vzo, R *tsave, bool save)
{
    const int4 nb = a.nbr[q];
    const unsigned code = (unsigned)nb.w;
    const Geom g = geometry(a.pt, a.vtx[q]);

    R txx[6], tzz[6], txz[6], vxe[6], vze[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int o = fidx(q, i);
        txx[i] = Txx[o]; tzz[i] = Tzz[o]; txz[i] = Txz[o];
        if (sizeof(R) == 4) { vxe[i] = Vxi[o]; vze[i] = Vzi[o]; } /* FP32: issue every load up front */
    }
    condition_own_stresses<R, P2>(a, code, q, shot, step, txx, tzz, txz);
    if (save) {
#pragma unroll
        for (int i = 0; i < 6; i++) {
            tsave[(0 * 6 + i) * DGX_BLOCK + threadIdx.x] = txx[i];
            tsave[(1 * 6 + i) * DGX_BLOCK + threadIdx.x] = tzz[i];
            tsave[(2 * 6 + i) * DGX_BLOCK + threadIdx.x] = txz[i];
        }
    }
    R rhx[6], rhz[6];
    velocity_volume<R, P2, Geom, false, true>(a, g, txx, tzz, txz, rhx, rhz);

    const int nbs[3] = {nb.x, nb.y, nb.z};
    bool global[3];
    R gxx[3][3], gzz[3][3], gxz[3][3];
    gather_stresses<false>(code, nbs, q, 0, global, Txx, Tzz, Txz, gxx, gzz, gxz);
    condition_neighbour_stresses(a, code, nbs, global, shot, step, gxx, gzz, gxz);
    R flx[6], flz[6];
    velocity_flux<R, P2>(a, g, nbs, txx, tzz, txz, gxx, gzz, gxz, flx, flz);
    /* V = (sponge(V) + force source) + dt/rho * (volume + surface/D), read at row q of Vi, written at row orow
     * of Vo.  FP32: velocities loaded up front, conditioned per element; FP64: loaded late (registers for
     * occupancy), conditioned per node. */
    const R invD = (R)g.invD;
    const R c = a.rho_inv[q] * (R)a.dt;
    if constexpr (sizeof(R) == 4) condition_velocity6(a, code, q, shot, step, vxe, vze);
#pragma unroll
    for (int i = 0; i < 6; i++) {
        R vx, vz;
        if constexpr (sizeof(R) == 4) {
            vx = vxe[i]; vz = vze[i];
        } else {
            const int o = fidx(q, i);
            vx = Vxi[o]; vz = Vzi[o];
            condition_velocity(a, code, o, i, shot, step, vx, vz);
        }
        const int oo = fidx(orow, i);
        Vxo[oo] = (rhx[i] + flx[i] * invD) * c + vx;
        Vzo[oo] = (rhz[i] + flz[i] * invD) * c + vz;
    }
}

/*
 * Velocity half-step with the workgroup's own stresses staged in LDS.  Each thread loads and conditions
 * its element's stresses and writes them to `wsh` ((c*6 + i)*DGX_BLOCK + thread, c = xx, zz, xz); face
 * neighbours inside the workgroup's 256 elements (about 91 % of interior faces in the locality order) are
 * then read from there after one barrier, already conditioned by their own thread, and only the others
 * are gathered from global memory, issued before the barrier.  Every thread of the workgroup reaches the
 * barrier: threads past the last element load element 0 and return after it.  In place: the velocities
 * are read and written at row q of Vx, Vz.
 */
template <typename R, bool FUSED, bool P2>
__device__ __forceinline__ void velocity_update_staged(const StepArgs<R> &a, int q, int q0, bool valid, int shot,
                                                       int step, const R *__restrict__ Txx,
                                                       const R *__restrict__ Tzz, const R *__restrict__ Txz,
                                                       R *__restrict__ Vx, R *__restrict__ Vz, R *wsh)
{
    const int qs = valid ? q : 0;
    const int4 nb = a.nbr[qs];
    const unsigned code = (unsigned)nb.w;
    const Geom g = geometry(a.pt, a.vtx[qs]);

    R txx[6], tzz[6], txz[6], vxe[6], vze[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int o = fidx(qs, i);
        txx[i] = Txx[o]; tzz[i] = Tzz[o]; txz[i] = Txz[o];
        if (sizeof(R) == 4) { vxe[i] = Vx[o]; vze[i] = Vz[o]; }
    }
    const int nbs[3] = {nb.x, nb.y, nb.z};
    bool inwg[3];
    R gxx[3][3], gzz[3][3], gxz[3][3];
    gather_stresses<true>(code, nbs, qs, q0, inwg, Txx, Tzz, Txz, gxx, gzz, gxz);
    if (FUSED) condition_own_stresses<R, P2>(a, code, qs, shot, step, txx, tzz, txz);
#pragma unroll
    for (int i = 0; i < 6; i++) {
        wsh[(0 * 6 + i) * DGX_BLOCK + threadIdx.x] = txx[i];
        wsh[(1 * 6 + i) * DGX_BLOCK + threadIdx.x] = tzz[i];
        wsh[(2 * 6 + i) * DGX_BLOCK + threadIdx.x] = txz[i];
    }
    R rhx[6], rhz[6];
    velocity_volume<R, P2, Geom, false, true>(a, g, txx, tzz, txz, rhx, rhz);

    __syncthreads();
#pragma unroll
    for (int j = 0; j < 3; j++) {
        if (!inwg[j]) continue;
        const unsigned m = (code >> (3 * j)) & 3u, rev = (code >> (3 * j + 2)) & 1u;
        const int r = (nbs[j] >= 0 ? nbs[j] : qs) - q0;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int nn = nb_node(m, rev, k);
            gxx[j][k] = wsh[(0 * 6 + nn) * DGX_BLOCK + r];
            gzz[j][k] = wsh[(1 * 6 + nn) * DGX_BLOCK + r];
            gxz[j][k] = wsh[(2 * 6 + nn) * DGX_BLOCK + r];
        }
    }
    if (!valid) return;
    if (FUSED) condition_neighbour_stresses(a, code, nbs, inwg, shot, step, gxx, gzz, gxz);
    R flx[6], flz[6];
    velocity_flux<R, P2>(a, g, nbs, txx, tzz, txz, gxx, gzz, gxz, flx, flz);
    /* V = (sponge(V) + force source) + dt/rho * (volume + surface/D), in place; loads and conditioning as in
     * velocity_update */
    const R invD = (R)g.invD;
    const R c = a.rho_inv[q] * (R)a.dt;
    if constexpr (sizeof(R) == 4) {
        if (FUSED) condition_velocity6(a, code, q, shot, step, vxe, vze);
    }
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int o = fidx(q, i);
        R vx, vz;
        if constexpr (sizeof(R) == 4) {
            vx = vxe[i]; vz = vze[i];
        } else {
            vx = Vx[o]; vz = Vz[o];
            if (FUSED) condition_velocity(a, code, o, i, shot, step, vx, vz);
        }
        Vx[o] = (rhx[i] + flx[i] * invD) * c + vx;
        Vz[o] = (rhz[i] + flz[i] * invD) * c + vz;
    }
}

/*
 * Stress update of element q in the full-step kernel: new velocities V (own values at row q, face neighbour
 * j's at row nrow.j; nrow.w is the element's code word), the element's conditioned stresses from LDS
 * (tload), the new stresses written to row q of T.  Records the receivers of q from the velocities it holds.
 */
template <typename R>
__device__ __forceinline__ void stress_update(const StepArgs<R> &a, int q, int shot, int step, int4 nrow,
                                              const R *__restrict__ Vx, const R *__restrict__ Vz, const R *tload,
                                              R *__restrict__ Txx, R *__restrict__ Tzz, R *__restrict__ Txz)
{
    const unsigned code = (unsigned)nrow.w;
    const Geom g = geometry(a.pt, a.vtx[q]);

    R vx[6], vz[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int o = fidx(q, i);
        vx[i] = Vx[o]; vz[i] = Vz[o];
    }
    R exx[6], ezz[6], exz[6];
    strain_volume(a, g, vx, vz, exx, ezz, exz);

    const int nbs[3] = {nrow.x, nrow.y, nrow.z};
    bool global[3];
    R gvx[3][3], gvz[3][3];
    gather_velocities<false>(code, nbs, q, 0, global, Vx, Vz, gvx, gvz);
    R fx[6], fz[6], fxz[6];
    stress_flux<R, false>(a, g, nbs, vx, vz, gvx, gvz, fx, fz, fxz);
    stress_store<R, true>(a, code, q, shot, step, g, exx, ezz, exz, fx, fz, fxz, tload, Txx, Tzz, Txz);
    record_receivers(a, code, shot, step, vx, vz);
}

/*
 * Stress half-step with the workgroup's new velocities staged in LDS (`wsh`, (c*6 + i)*DGX_BLOCK + thread,
 * c = x, z), as velocity_update_staged does for the stresses: face neighbours inside the workgroup are read
 * from LDS after one barrier, the others are gathered from global memory before it.  In place on T.
 */
template <typename R, bool FUSED, bool P2>
__device__ __forceinline__ void stress_update_staged(const StepArgs<R> &a, int q, int q0, bool valid, int shot,
                                                     int step, const R *__restrict__ Vx, const R *__restrict__ Vz,
                                                     R *__restrict__ Txx, R *__restrict__ Tzz, R *__restrict__ Txz,
                                                     R *wsh)
{
    const int qs = valid ? q : 0;
    const int4 nb = a.nbr[qs];
    const unsigned code = (unsigned)nb.w;
    const Geom g = geometry(a.pt, a.vtx[qs]);

    R vx[6], vz[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int o = fidx(qs, i);
        vx[i] = Vx[o]; vz[i] = Vz[o];
    }
    const int nbs[3] = {nb.x, nb.y, nb.z};
    bool inwg[3];
    R gvx[3][3], gvz[3][3];
    gather_velocities<true>(code, nbs, qs, q0, inwg, Vx, Vz, gvx, gvz);
#pragma unroll
    for (int i = 0; i < 6; i++) {
        wsh[(0 * 6 + i) * DGX_BLOCK + threadIdx.x] = vx[i];
        wsh[(1 * 6 + i) * DGX_BLOCK + threadIdx.x] = vz[i];
    }
    R exx[6], ezz[6], exz[6];
    /* FP64 reads Dr/Ds from the argument here (compiled in, they cost the kernel its occupancy); FP32 with the
     * shipped operators uses the factored derivatives */
    strain_volume<R, Geom, P2 && sizeof(R) == 4, false, true>(a, g, vx, vz, exx, ezz, exz);

    __syncthreads();
#pragma unroll
    for (int j = 0; j < 3; j++) {
        if (!inwg[j]) continue;
        const unsigned m = (code >> (3 * j)) & 3u, rev = (code >> (3 * j + 2)) & 1u;
        const int r = (nbs[j] >= 0 ? nbs[j] : qs) - q0;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int nn = nb_node(m, rev, k);
            gvx[j][k] = wsh[(0 * 6 + nn) * DGX_BLOCK + r];
            gvz[j][k] = wsh[(1 * 6 + nn) * DGX_BLOCK + r];
        }
    }
    if (!valid) return;
    R fx[6], fz[6], fxz[6];
    stress_flux<R, P2>(a, g, nbs, vx, vz, gvx, gvz, fx, fz, fxz);
    stress_store<R, FUSED>(a, code, q, shot, step, g, exx, ezz, exz, fx, fz, fxz, nullptr, Txx, Tzz, Txz);
    if (FUSED) record_receivers(a, code, shot, step, vx, vz);
}

/* The element kernels ask for at least 3 wavefronts per SIMD: this caps the FP64 kernels at 168 VGPRs
 * (unbounded they take 170-171, which leaves room for 2), without spilling. */
template <typename R, bool FUSED, bool P2>
__global__ void __launch_bounds__(DGX_BLOCK, 3) k_velocity(StepArgs<R> a)
{
    __shared__ R wsh[18 * DGX_BLOCK];
    const int q = element_of_thread(a.nper);
    const int q0 = q - (int)threadIdx.x;
    const int shot = blockIdx.y;
    const long long so = (long long)shot * a.fstride * 6;
    const int step = FUSED ? (*a.step_base + a.step_off) : 0;
    velocity_update_staged<R, FUSED, P2>(a, q, q0, q < a.tri, shot, step, a.f[2] + so, a.f[3] + so, a.f[4] + so,
                                         a.f[0] + so, a.f[1] + so, wsh);
}

template <typename R, bool FUSED, bool P2>
__global__ void __launch_bounds__(DGX_BLOCK, 3) k_stress(StepArgs<R> a)
{
    __shared__ R wsh[12 * DGX_BLOCK];
    const int q = element_of_thread(a.nper);
    const int q0 = q - (int)threadIdx.x;
    const int shot = blockIdx.y;
    const long long so = (long long)shot * a.fstride * 6;
    const int step = FUSED ? (*a.step_base + a.step_off) : 0;
    stress_update_staged<R, FUSED, P2>(a, q, q0, q < a.tri, shot, step, a.f[0] + so, a.f[1] + so, a.f[2] + so,
                                       a.f[3] + so, a.f[4] + so, wsh);
}

/*
 * One whole leap-frog step per launch (element layout), reading the fields in f and writing the step's
 * result to the other copy g.  Workgroup b owns tile t: `own` consecutive elements (in the locality
 * order) plus the `nring` elements outside the tile that share a face with them, together at most one
 * element per thread.
 *   phase 1: every thread computes the new velocity of its element (own or ring) from f, exactly as the
 *            velocity half-step does, and writes it to g: own elements at their row, ring elements at
 *            the tile's private rows vrow0 + (ring entry); own threads keep their conditioned stresses
 *            in LDS;
 *   barrier;
 *   phase 2: own threads update their stresses from the new velocities in g (in-tile neighbours at
 *            their rows, outside neighbours at the tile's ring rows, nb2) and write them to g.
 * Ring velocities are recomputed by every tile that borders them, from the same inputs with the same
 * code, so they equal the owner's values.  Since f is only read and every row of g is written by one
 * workgroup, tiles never wait on each other.  Fields move once per step: 5 reads + 5 writes per
 * element-node, against 7 + 8 for the two half-step kernels.
 */
template <typename R, bool P2>
__global__ void __launch_bounds__(DGX_BLOCK, 3) k_fullstep(StepArgs<R> a)
{
    __shared__ R tc[18 * DGX_BLOCK];
    const int b = blockIdx.x;
    const int t = (b & (DGX_NXCD - 1)) * a.tper + (b >> 3);
    if (t >= a.ntile) return; /* uniform over the workgroup */
    const int4 td = a.tiles[t]; /* {first, own, ring begin, nring} */
    const int lane = threadIdx.x;
    const int shot = blockIdx.y;
    const long long so = (long long)shot * a.fstride * 6;
    const int step = *a.step_base + a.step_off;

    int q = -1, orow = 0;
    if (lane < td.y) {
        q = td.x + lane;
        orow = q;
    } else if (lane < td.y + td.w) {
        const int k = td.z + lane - td.y;
        q = a.ring[k];
        orow = a.vrow0 + k;
    }
    if (q >= 0)
        velocity_update<R, P2>(a, q, shot, step, orow, a.f[2] + so, a.f[3] + so, a.f[4] + so, a.f[0] + so,
                                 a.f[1] + so, a.g[0] + so, a.g[1] + so, tc, lane < td.y);
    __syncthreads();
    /* argument operators in the stress phase: with the lifts compiled in, the FP64 kernel spills */
    if (lane < td.y)
        stress_update<R>(a, q, shot, step, a.nb2[q], a.g[0] + so, a.g[1] + so, tc, a.g[2] + so,
                               a.g[3] + so, a.g[4] + so);
}

/* Wavelet sample of a source at local step `lstep`: the solver's wavelet table row src.wavelet, or in a
 * serving batch (SERVE) the source's slot row. */
template <typename R, bool SERVE>
__device__ __forceinline__ double block_wavelet(const StepArgs<R> &a, const dgx_src_dev &src, int lstep)
{
    if constexpr (SERVE) return as_global(a.slot_wav[src.shot])[lstep];
    else return a.wav[(long long)src.wavelet * a.nsteps + lstep];
}

/* Compensated accumulation: the value h + l (l the low word) plus d, with the sum rounded into h and its
 * exact rounding error (TwoSum: Knuth, TAOCP vol. 2, 4.2.2) kept in l.  The increment absorbs the old low
 * word first; that rounding is second order (|l| <= ulp(h)/2). */
__device__ __forceinline__ void comp_add(float &h, float &l, float d)
{
    const float y = d + l;
    const float s = h + y;
    const float bb = s - h;
    l = (h - (s - bb)) + (y - bb);
    h = s;
}

/* Compensated sponge (h + l)(1 - d), d the FP32 deficit of the factor (sponge()): h d = p + e exactly (FMA), and
 * h - p is accumulated with TwoSum. */
__device__ __forceinline__ void comp_sponge(float &h, float &l, float d)
{
    const float p = h * d;
    const float e = __builtin_fmaf(h, d, -p);
    l = l - l * d - e;
    comp_add(h, l, -p);
}

/* Sponge factors (ab, loaded once per launch) and the element's sources (aux record ax, 0 = none) at local
 * step `lstep` applied to its stresses and, with VEL, its velocities: Absorb_Boundary then source_kernel
 * (OpenMP_Function.c:25-58); explosive sources add to Txx and Tzz, force sources to Vz. */
template <typename R, bool VEL, bool SERVE, bool COMP>
__device__ __forceinline__ void condition_block(const StepArgs<R> &a, unsigned code, unsigned ax, int shot, int lstep,
                                                const R (&ab)[6], R (&vx)[6], R (&vz)[6], R (&txx)[6],
                                                R (&tzz)[6], R (&txz)[6], R (&lo)[5][6])
{
    if (code & DGX_CODE_SPONGE) {
#pragma unroll
        for (int i = 0; i < 6; i++) {
            if constexpr (COMP) {
                comp_sponge(txx[i], lo[2][i], ab[i]); comp_sponge(tzz[i], lo[3][i], ab[i]); comp_sponge(txz[i], lo[4][i], ab[i]);
                if (VEL) { comp_sponge(vx[i], lo[0][i], ab[i]); comp_sponge(vz[i], lo[1][i], ab[i]); }
            } else {
                txx[i] = sponge(txx[i], ab[i]); tzz[i] = sponge(tzz[i], ab[i]); txz[i] = sponge(txz[i], ab[i]);
                if (VEL) { vx[i] = sponge(vx[i], ab[i]); vz[i] = sponge(vz[i], ab[i]); }
            }
        }
    }
    if (ax) {
        const dgx_aux x = a.aux[ax - 1];
        for (int k = x.src_begin; k < x.src_begin + x.src_count; k++) {
            const dgx_src_dev &src = a.src[k];
            if (src.shot != shot || (!VEL && src.kind != 1)) continue;
            const double w = block_wavelet<R, SERVE>(a, src, lstep);
#pragma unroll
            for (int i = 0; i < 6; i++) {
                const R inc = (R)(src.fun[i] * w);
                if constexpr (COMP) {
                    if (src.kind == 1) { comp_add(txx[i], lo[2][i], inc); comp_add(tzz[i], lo[3][i], inc); }
                    else if (VEL) comp_add(vz[i], lo[1][i], inc);
                } else {
                    if (src.kind == 1) { txx[i] += inc; tzz[i] += inc; }
                    else if (VEL) vz[i] += inc;
                }
            }
        }
    }
}

/* Receivers of the element (aux record ax) from the velocities it holds (Record, OpenMP_Function.c:61-72, with
 * interpolation weights): trace slot lstep of the solver's shot, or of the serving slot's trace buffer. */
template <typename R, bool SERVE, bool COMP>
__device__ __forceinline__ void record_block(const StepArgs<R> &a, unsigned ax, int shot, int lstep, const R (&vx)[6],
                                             const R (&vz)[6], const R (&lo)[5][6])
{
    if (!ax || (!SERVE && lstep >= a.nsteps)) return;
    const dgx_aux x = a.aux[ax - 1];
    for (int r = x.rec_begin; r < x.rec_begin + x.rec_count; r++) {
        const dgx_rec_dev &rc = a.rec[r];
        if (SERVE && rc.pad0 != shot) continue;
        double sx_ = 0, sz_ = 0;
#pragma unroll
        for (int i = 0; i < 6; i++) {
            if constexpr (COMP) { /* the compensated value h + l */
                sx_ += rc.w[i] * ((double)vx[i] + (double)lo[0][i]);
                sz_ += rc.w[i] * ((double)vz[i] + (double)lo[1][i]);
            } else {
                sx_ += rc.w[i] * (double)vx[i];
                sz_ += rc.w[i] * (double)vz[i];
            }
        }
        const int nrec = SERVE ? a.slot_nrec[shot] : a.nrec;
        double *t = SERVE ? as_global(a.slot_trace[shot]) + (long long)lstep * 2 * nrec
                          : a.trace + ((long long)shot * a.nsteps + lstep) * 2 * nrec;
        t[rc.index] = sx_;
        t[nrec + rc.index] = sz_;
    }
}

/* Velocity update of one element of a blocked tile (LU_Velocity, OpenMP_Function.c:75-185): the volume term and
 * the three face terms accumulate into one value per node (the face scalings carry 1/D), then V += c * rhs with
 * c = dt/rho.  The neighbour across face j is slot nsl[j] of the LDS stage sts (rows of NS slots); on the outer
 * boundary that is the workgroup's zero slot, so the jump is -U_own as in fun_DG.c:724-736. */
template <typename R, int NS, bool P2, bool COMP>
__device__ __forceinline__ void block_velocity(const StepArgs<R> &a, const GeomR<R> &g, unsigned code,
                                               const int (&nsl)[3], const R *sts, R c, const R (&txx)[6],
                                               const R (&tzz)[6], const R (&txz)[6], R (&vx)[6], R (&vz)[6],
                                               R (&lo)[5][6])
{
    R rhx[6], rhz[6];
    velocity_volume<R, P2, GeomR<R>, true, !COMP>(a, g, txx, tzz, txz, rhx, rhz);
#pragma unroll
    for (int fc = 0; fc < 3; fc++) {
        const unsigned m = (code >> (3 * fc)) & 3u, rev = (code >> (3 * fc + 2)) & 1u;
        const R ex = g.exd[fc], ez = g.ezd[fc];
        const R *nb = sts + nsl[fc];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int on = face_node(fc, k), nn = nb_node(m, rev, k);
            const R dxx = nb[(0 * 6 + nn) * NS] - txx[on];
            const R dzz = nb[(1 * 6 + nn) * NS] - tzz[on];
            const R dxz = nb[(2 * 6 + nn) * NS] - txz[on];
            const R gx = ez * dxx - ex * dxz;
            const R gz = ez * dxz - ex * dzz;
#pragma unroll
            for (int i = 0; i < 6; i++) {
                rhx[i] += op_l<P2>(a, fc * 18 + i * 3 + k) * gx;
                rhz[i] += op_l<P2>(a, fc * 18 + i * 3 + k) * gz;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 6; i++) {
        if constexpr (COMP) {
            comp_add(vx[i], lo[0][i], rhx[i] * c);
            comp_add(vz[i], lo[1][i], rhz[i] * c);
        } else {
            vx[i] = rhx[i] * c + vx[i];
            vz[i] = rhz[i] * c + vz[i];
        }
    }
}

/* Stress update of one element of a blocked tile (LU_Stress_Elastic, OpenMP_Function.c:188-306) in strain-rate
 * form: the volume strain rates, then each face's lifted velocity jumps S_j dVx, S_j dVz scaled by its edge
 * vector over D, then T += dt C e with the material constants already multiplied by dt (d2m = (l + 2m) dt,
 * dl = l dt, dm = m dt).  Neighbour velocities come from the LDS stage svl at slot nsl[j] (the zero slot on the
 * outer boundary). */
template <typename R, int NS, bool P2, bool COMP>
__device__ __forceinline__ void block_stress(const StepArgs<R> &a, const GeomR<R> &g, unsigned code,
                                             const int (&nsl)[3], const R *svl, R d2m, R dl, R dm,
                                             const R (&vx)[6], const R (&vz)[6], R (&txx)[6], R (&tzz)[6],
                                             R (&txz)[6], R (&lo)[5][6])
{
    R axx[6], azz[6], axz[6];
    strain_volume<R, GeomR<R>, P2, true, !COMP>(a, g, vx, vz, axx, azz, axz);
#pragma unroll
    for (int fc = 0; fc < 3; fc++) {
        const unsigned m = (code >> (3 * fc)) & 3u, rev = (code >> (3 * fc + 2)) & 1u;
        const R ex = g.exd[fc], ez = g.ezd[fc];
        const R *nb = svl + nsl[fc];
        R dvx[3], dvz[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int on = face_node(fc, k), nn = nb_node(m, rev, k);
            dvx[k] = nb[(0 * 6 + nn) * NS] - vx[on];
            dvz[k] = nb[(1 * 6 + nn) * NS] - vz[on];
        }
#pragma unroll
        for (int i = 0; i < 6; i++) {
            R lx = 0, lz = 0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                lx += op_l<P2>(a, fc * 18 + i * 3 + k) * dvx[k];
                lz += op_l<P2>(a, fc * 18 + i * 3 + k) * dvz[k];
            }
            axx[i] += ez * lx;
            azz[i] -= ex * lz;
            axz[i] += ez * lz - ex * lx;
        }
    }
#pragma unroll
    for (int i = 0; i < 6; i++) {
        if constexpr (COMP) {
            comp_add(txx[i], lo[2][i], d2m * axx[i] + dl * azz[i]);
            comp_add(tzz[i], lo[3][i], dl * axx[i] + d2m * azz[i]);
            comp_add(txz[i], lo[4][i], dm * axz[i]);
        } else {
            txx[i] = d2m * axx[i] + (dl * azz[i] + txx[i]);
            tzz[i] = dl * axx[i] + (d2m * azz[i] + tzz[i]);
            txz[i] = dm * axz[i] + txz[i];
        }
    }
}

/* LDS slots per stage row of the blocked kernel with NT threads: one per thread, then NT/4 for ghost elements,
 * the last of which is the zero slot (a multiple of 64, so that rows stay aligned for paired LDS stores). */
template <int NT>
__device__ __forceinline__ constexpr int block_slots()
{
    return NT + NT / 4;
}

/*
 * Several whole leap-frog steps of one tile (element layout), with the fields held on chip between them.
 * Tile t holds `own` consecutive elements in the locality order (level 0) and every element within 2*bkmax
 * face hops of them (level d = d hops).  Levels 0 .. 2*bkmax - 1 take one thread each; level 2*bkmax (ghosts)
 * only supplies stresses, conditioned at the first step, to the first velocity update of the level inside it:
 * thread g loads ghost g besides its own element and stages it in LDS slot NT + g.  Each thread reads its
 * element's fields from f once; nst steps follow (global steps step0.., local steps lstep0.. for the wavelet
 * and traces), step j being:
 *   every element conditions its values (sponge, sources: Absorb_Boundary and source_kernel,
 *   OpenMP_Function.c:25-58) and stages its stresses in LDS; barrier;
 *   elements of level <= 2 nst - 1 - 2j update their velocities (LU_Velocity, :75-185), face neighbours read
 *   from LDS, and stage them; barrier;
 *   elements of level <= 2 nst - 2 - 2j update their stresses (LU_Stress_Elastic, :188-306); level 0 records
 *   its receivers (Record, :61-72);
 * then the own elements' fields go to g.  The outer levels are recomputed from the same inputs with the same
 * code as their owning tile's, so tiles never wait on each other; f is only read and every row of g is written
 * by one workgroup.  Per launch each field is read once for every element of the tile and written once for
 * every own element, against 7 reads and writes per step and element for the half-step kernel pair.  Every
 * thread of the workgroup reaches every barrier.
 */
template <typename R, int NT, bool P2, bool SERVE, bool COMP = false>
__device__ __forceinline__ void block_tile(const StepArgs<R> &a, R *const f[5], R *const gout[5], int t, int shot,
                                           int nst, int lstep0, R *sts, R *svl)
{
    constexpr int NS = block_slots<NT>();
    const int4 td = a.tiles[t];
    const int l = threadIdx.x;
    const int ghost = 2 * a.bkmax;
    const long long so = (long long)shot * a.fstride * 6;
    /* the tile's level ends (uniform loads): lanes = elements with a thread, the rest of its records are ghosts */
    const int *le = a.blevels + (long long)t * DGX_BLOCK_LEVELS;
    const int lanes = le[ghost - 1], nghost = td.w - lanes;
    int lev = ghost + 1, q = 0; /* lev > ghost: no element (threads past the tile's lanes) */
    int4 nb = make_int4(-1, -1, -1, 0), lr = make_int4(0, 0, 0, 0);
    if (l < lanes) {
        /* the level of lane l: the number of level ends at or below it (the ends ascend) */
        lev = 0;
#pragma unroll
        for (int d = 0; d < DGX_BLOCK_LEVELS; d++) lev += (d < ghost && l >= le[d]) ? 1 : 0;
        lr = a.blrec[td.z + l];
        q = lr.x;
        nb = a.blnb[td.z + l];
    }
    const unsigned code = (unsigned)nb.w;
    unsigned ax = 0; /* aux record + 1 of the element's sources and receivers */
    if (lev < ghost) ax = SERVE ? ((unsigned)a.aux_of_elem[q] & DGX_DYN_AUX_MASK) : (code >> DGX_CODE_AUX_SHIFT);
    /* Slot NS - 1 is never an element (ghosts fill at most NT/4 - 1 slots): it holds zeros for the whole launch
     * and stands in for the neighbour across an outer-boundary face. */
    const int nsl[3] = {nb.x >= 0 ? nb.x : NS - 1, nb.y >= 0 ? nb.y : NS - 1, nb.z >= 0 ? nb.z : NS - 1};
    if (l == NT - 1) {
#pragma unroll
        for (int r = 0; r < 18; r++) sts[r * NS + NS - 1] = 0;
#pragma unroll
        for (int r = 0; r < 12; r++) svl[r * NS + NS - 1] = 0;
    }
    /* ghost l: its stresses, conditioned for the first step, into slot NT + l */
    if (l < nghost) {
        const int4 glr = a.blrec[td.z + lanes + l];
        const unsigned gcode = (unsigned)a.blnb[td.z + lanes + l].w;
        const int gq = glr.x;
        const unsigned gax = SERVE ? ((unsigned)a.aux_of_elem[gq] & DGX_DYN_AUX_MASK) : (gcode >> DGX_CODE_AUX_SHIFT);
        R gxx[6], gzz[6], gxz[6], gab[6], gv[6], glo[5][6];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const int o = fidx(gq, i);
            gxx[i] = f[2][so + o]; gzz[i] = f[3][so + o]; gxz[i] = f[4][so + o];
            gab[i] = (gcode & DGX_CODE_SPONGE) ? a.absorb[o] : (R)0; /* used only with the sponge bit */
            if constexpr (COMP) { glo[2][i] = a.fl[2][so + o]; glo[3][i] = a.fl[3][so + o]; glo[4][i] = a.fl[4][so + o]; }
        }
        condition_block<R, false, SERVE, COMP>(a, gcode, gax, shot, lstep0, gab, gv, gv, gxx, gzz, gxz, glo);
#pragma unroll
        for (int i = 0; i < 6; i++) {
            sts[(0 * 6 + i) * NS + NT + l] = gxx[i];
            sts[(1 * 6 + i) * NS + NT + l] = gzz[i];
            sts[(2 * 6 + i) * NS + NT + l] = gxz[i];
        }
    }
    const bool upd = lev < ghost; /* element updated by this tile */

    /* the stresses live in registers from their update to the next step's conditioning and in LDS (sts) in
     * between; the velocities stay in registers */
    R vx[6], vz[6], txx[6], tzz[6], txz[6], ab[6];
    R lo[5][6]; /* compensated FP32: the low words of Vx, Vz, Txx, Tzz, Txz (unused otherwise) */
    GeomR<R> g;
    R c = 0, d2m = 0, dl = 0, dm = 0; /* dt/rho, (l + 2m) dt, l dt, m dt */
    if (upd) {
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const int o = fidx(q, i);
            txx[i] = f[2][so + o]; tzz[i] = f[3][so + o]; txz[i] = f[4][so + o];
            ab[i] = (code & DGX_CODE_SPONGE) ? a.absorb[o] : (R)0; /* used only with the sponge bit */
            if constexpr (COMP) { lo[2][i] = a.fl[2][so + o]; lo[3][i] = a.fl[3][so + o]; lo[4][i] = a.fl[4][so + o]; }
        }
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const int o = fidx(q, i);
            vx[i] = f[0][so + o]; vz[i] = f[1][so + o];
            if constexpr (COMP) { lo[0][i] = a.fl[0][so + o]; lo[1][i] = a.fl[1][so + o]; }
        }
        g = geometry_r<R>(geometry(a.pt, make_int4(lr.y, lr.z, lr.w, 0)));
        const R dt = (R)a.dt, lam = a.lam[q], mu = a.mu[q];
        c = a.rho_inv[q] * dt;
        d2m = (lam + 2 * mu) * dt;
        dl = lam * dt;
        dm = mu * dt;
    }
    for (int j = 0; j < nst; j++) {
        const int lstep = lstep0 + j;
        if (upd) {
            condition_block<R, true, SERVE, COMP>(a, code, ax, shot, lstep, ab, vx, vz, txx, tzz, txz, lo);
#pragma unroll
            for (int i = 0; i < 6; i++) {
                sts[(0 * 6 + i) * NS + l] = txx[i];
                sts[(1 * 6 + i) * NS + l] = tzz[i];
                sts[(2 * 6 + i) * NS + l] = txz[i];
            }
        }
        __syncthreads();
        if (lev <= 2 * nst - 1 - 2 * j)
            block_velocity<R, NS, P2, COMP>(a, g, code, nsl, sts, c, txx, tzz, txz, vx, vz, lo);
        if (upd) {
#pragma unroll
            for (int i = 0; i < 6; i++) {
                svl[(0 * 6 + i) * NS + l] = vx[i];
                svl[(1 * 6 + i) * NS + l] = vz[i];
            }
        }
        __syncthreads();
        if (lev <= 2 * nst - 2 - 2 * j) {
            block_stress<R, NS, P2, COMP>(a, g, code, nsl, svl, d2m, dl, dm, vx, vz, txx, tzz, txz, lo);
            if (lev == 0) record_block<R, SERVE, COMP>(a, ax, shot, lstep, vx, vz, lo);
        }
    }
    if (lev == 0) {
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const int o = fidx(q, i);
            gout[0][so + o] = vx[i]; gout[1][so + o] = vz[i];
            gout[2][so + o] = txx[i]; gout[3][so + o] = tzz[i]; gout[4][so + o] = txz[i];
            if constexpr (COMP) {
#pragma unroll
                for (int k = 0; k < 5; k++) a.gl[k][so + o] = lo[k][i];
            }
        }
    }
}

/* The solver's blocked kernel: workgroup (b, shot) runs tile (b mod 8) * tper + b / 8 (XCD-contiguous runs of
 * tiles) of that shot for bsteps steps.  MINW: at least that many wavefronts per SIMD (512-thread tiles in
 * FP32: 4, two workgroups per CU). */
template <typename R, int NT, int MINW, bool P2, bool COMP = false>
__global__ void __launch_bounds__(NT, MINW) k_block(StepArgs<R> a)
{
    __shared__ R sts[18 * block_slots<NT>()]; /* conditioned stresses, (c*6 + i)*slots + slot */
    __shared__ R svl[12 * block_slots<NT>()]; /* new velocities */
    const int b = blockIdx.x;
    const int t = (b & (DGX_NXCD - 1)) * a.tper + (b >> 3);
    if (t >= a.ntile) return; /* uniform over the workgroup */
    const int step0 = *a.step_base + a.step_off;
    block_tile<R, NT, P2, false, COMP>(a, a.f, a.g, t, blockIdx.y, a.bsteps, step0, sts, svl);
}

/*
 * Serving with single-shot slots: one launch advances up to bkmax steps of every active slot (a batch of one
 * shot in the element layout) over the concatenated tiles of all batches: sched = {nbatch, total tiles, tile
 * offset of batch 0, ..., of batch nbatch}.  Workgroup b takes tile (b mod 8) * nper + b / 8 of that list, its
 * batch by a scalar binary search; the batch's slot runs min(bkmax, steps left) steps from its local step.
 * Launch `launch` of a block reads copy (launch & 1) of the batch's fields and writes the other.
 */
template <typename R, int NT, int MINW, bool P2>
__global__ void __launch_bounds__(NT, MINW) ks_block(const StepArgs<R> *__restrict__ table,
                                                    const int *__restrict__ sched, const int *step_base,
                                                    int step_off, int launch, int nb, int total, int nper)
{
    __shared__ R sts[18 * block_slots<NT>()];
    __shared__ R svl[12 * block_slots<NT>()];
    const int tile = (int)(blockIdx.x & (DGX_NXCD - 1)) * nper + (int)(blockIdx.x >> 3);
    if (tile >= total) return;
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sched[2 + mid] <= tile) lo = mid;
        else hi = mid - 1;
    }
    /* the batch's arguments, with the pointers the tile reads made global (device memory) */
    StepArgs<R> a = table[lo];
    a.tiles = as_global(a.tiles);
    a.blevels = as_global(a.blevels);
    a.blrec = as_global(a.blrec);
    a.blnb = as_global(a.blnb);
    a.vtx = as_global(a.vtx);
    a.pt = as_global(a.pt);
    a.rho_inv = as_global(a.rho_inv);
    a.lam = as_global(a.lam);
    a.mu = as_global(a.mu);
    a.absorb = as_global(a.absorb);
    a.aux = as_global(a.aux);
    a.src = as_global(a.src);
    a.rec = as_global(a.rec);
    a.aux_of_elem = as_global(a.aux_of_elem);
    a.slot_wav = as_global(a.slot_wav);
    a.slot_trace = as_global(a.slot_trace);
    a.slot_nrec = as_global(a.slot_nrec);
    const int lstep0 = *step_base + step_off - as_global(a.shot_start)[0];
    const int left = as_global(a.shot_len)[0] - lstep0;
    const int nst = left < a.bsteps ? left : a.bsteps;
    if (lstep0 < 0 || nst <= 0) return; /* slot not running in this launch: uniform over the workgroup */
    R *f[5], *gout[5];
#pragma unroll
    for (int c = 0; c < 5; c++) {
        f[c] = as_global((launch & 1) ? a.g[c] : a.f[c]);
        gout[c] = as_global((launch & 1) ? a.f[c] : a.g[c]);
    }
    block_tile<R, NT, P2, true>(a, f, gout, tile - sched[2 + lo], 0, nst, lstep0, sts, svl);
}

/* ---- split entry points (unfused reference step functions) ---- */

template <typename R>
__global__ void __launch_bounds__(DGX_BLOCK) k_absorb(StepArgs<R> a)
{
    const int q = element_of_thread(a.nper);
    if (q >= a.tri) return;
    if (!((unsigned)a.nbr[q].w & DGX_CODE_SPONGE)) return;
    const long long so = (long long)blockIdx.y * a.fstride * 6;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int o = fidx(q, i);
        const R f = a.absorb[o];
#pragma unroll
        for (int c = 0; c < 5; c++) a.f[c][so + o] = sponge(a.f[c][so + o], f);
    }
}

/* All sources in one thread, in list order, so that two sources on one element add in a fixed order. */
template <typename R>
__global__ void k_source(StepArgs<R> a, int nsrc, double w)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int s = 0; s < nsrc; s++) {
        const dgx_src_dev &src = a.src[s];
        const long long so = (long long)src.shot * a.fstride * 6;
        for (int i = 0; i < 6; i++) {
            const R inc = (R)(src.fun[i] * w);
            const int o = fidx(src.elem, i);
            if (src.kind == 0) a.f[1][so + o] += inc;
            else { a.f[2][so + o] += inc; a.f[3][so + o] += inc; }
        }
    }
}

template <typename R>
__global__ void k_record(StepArgs<R> a, int t, int nshots)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nrec * nshots) return;
    const int shot = k / a.nrec, r = k % a.nrec;
    const dgx_rec_dev &rc = a.rec[r];
    const long long so = (long long)shot * a.fstride * 6;
    double sx_ = 0, sz_ = 0;
    for (int i = 0; i < 6; i++) {
        const int o = fidx(rc.elem, i);
        sx_ += rc.w[i] * (double)a.f[0][so + o];
        sz_ += rc.w[i] * (double)a.f[1][so + o];
    }
    double *tr = a.trace + ((long long)shot * a.nsteps + t) * 2 * a.nrec;
    tr[rc.index] = sx_;
    tr[a.nrec + rc.index] = sz_;
}

/* Plain streaming copies, the reference point for HBM counters and achievable bandwidth.
 * variant 0: grid-stride loop, one 32-byte vector per iteration;
 * variant 1: one workgroup per 32 KiB, four 32-byte vectors per thread, non-temporal. */
__global__ void __launch_bounds__(256) k_copy(const double4 *__restrict__ src, double4 *__restrict__ dst, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

typedef double v4d __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_copy4(const double4 *__restrict__ src, double4 *__restrict__ dst, long long n)
{
    const v4d *s = (const v4d *)src;
    v4d *d = (v4d *)dst;
    const long long base = (long long)blockIdx.x * 1024 + threadIdx.x;
    v4d v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const long long i = base + k * 256;
        if (i < n) v[k] = __builtin_nontemporal_load(&s[i]);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const long long i = base + k * 256;
        if (i < n) __builtin_nontemporal_store(v[k], &d[i]);
    }
}

/* variant 2: read-only stream (sum of 32 KiB per workgroup, one value written per workgroup) */
__global__ void __launch_bounds__(256) k_read4(const double4 *__restrict__ src, double *__restrict__ out, long long n)
{
    const v4d *s = (const v4d *)src;
    const long long base = (long long)blockIdx.x * 1024 + threadIdx.x;
    v4d acc = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const long long i = base + k * 256;
        if (i < n) acc += __builtin_nontemporal_load(&s[i]);
    }
    double t = acc.x + acc.y + acc.z + acc.w;
    for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o);
    __shared__ double part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

/* Snapshot of one field of one shot in the caller's element order, as doubles: out[6*perm[q] + i] is node i
 * of device element q, read from the element layout's blocked row or, when `spad` > 0, from the
 * shot-innermost layout (node rows of spad shots interleaving 2^il elements:
 * ((((q >> il)*6 + i) << il) + q mod 2^il)*spad + shot). */
template <typename R>
__global__ void __launch_bounds__(256) k_unpermute(const R *__restrict__ f, const R *__restrict__ lo,
                                                   const int *__restrict__ perm, double *__restrict__ out, int tri,
                                                   long long spad, int il, long long shot)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= tri) return;
    double *o = out + 6 * (long long)perm[q];
    const long long row0 = (long long)(q >> il) * 6, sub = q & ((1 << il) - 1);
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const long long k = spad > 0 ? (((row0 + i) << il) + sub) * spad + shot : fidx(q, i);
        o[i] = lo ? (double)f[k] + (double)lo[k] : (double)f[k]; /* compensated FP32: high + low word */
    }
}

__global__ void __launch_bounds__(256) k_lane_records(const int *__restrict__ gid, const int4 *__restrict__ vtx,
                                                     int4 *__restrict__ rec, long long n)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int q = gid[k];
    const int4 v = vtx[q];
    rec[k] = make_int4(q, v.x, v.y, v.z);
}

__global__ void k_advance(int *step_base, int n)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *step_base += n;
}

} // namespace dgx

/* ---- host-side launchers (C linkage inside the library only) ---- */
using dgx::StepArgs;

template <typename R, bool P2>
static void launch_step_kernels(const StepArgs<R> &a, int nshots, hipStream_t st, int which)
{
    dim3 grid(DGX_NXCD * a.nper, nshots), block(DGX_BLOCK);
    switch (which) {
    case 0: dgx::k_velocity<R, true, P2><<<grid, block, 0, st>>>(a); break;
    case 1: dgx::k_stress<R, true, P2><<<grid, block, 0, st>>>(a); break;
    case 2: dgx::k_velocity<R, false, P2><<<grid, block, 0, st>>>(a); break;
    case 3: dgx::k_stress<R, false, P2><<<grid, block, 0, st>>>(a); break;
    case 4: dgx::k_absorb<R><<<grid, block, 0, st>>>(a); break;
    case 5: dgx::k_fullstep<R, P2><<<dim3(DGX_NXCD * a.tper, nshots), block, 0, st>>>(a); break;
    case 6:
        /* one element per thread, a.bthreads threads (the tiles are built for that count): FP32 1024 (one
         * workgroup per CU) or 512 (two), FP64 256 (two) or 512 (one) */
        {
        const dim3 bg(DGX_NXCD * a.tper, nshots);
        if constexpr (sizeof(R) == 4) {
            /* compensated FP32: twice the registers for the low words, one 512-thread workgroup per CU */
            if (a.fl[0]) dgx::k_block<R, 512, 2, P2, true><<<bg, 512, 0, st>>>(a);
            else if (a.bthreads == 1024) dgx::k_block<R, 1024, 1, P2><<<bg, 1024, 0, st>>>(a);
            else dgx::k_block<R, 512, 4, P2><<<bg, 512, 0, st>>>(a);
        } else if (a.bthreads == 256) {
            dgx::k_block<R, 256, 2, P2><<<bg, 256, 0, st>>>(a);
        } else {
            dgx::k_block<R, 512, 1, P2><<<bg, 512, 0, st>>>(a);
        }
        }
        break;
    }
}

template <typename R>
static hipError_t launch_step_kernels(const StepArgs<R> &a, int nshots, hipStream_t st, int which)
{
    if (a.p2) launch_step_kernels<R, true>(a, nshots, st, which);
    else launch_step_kernels<R, false>(a, nshots, st, which);
    return hipGetLastError();
}

extern "C" {
hipError_t dgx_launch_f64(const void *args, int nshots, hipStream_t st, int which)
{
    return launch_step_kernels(*(const StepArgs<double> *)args, nshots, st, which);
}
hipError_t dgx_launch_f32(const void *args, int nshots, hipStream_t st, int which)
{
    return launch_step_kernels(*(const StepArgs<float> *)args, nshots, st, which);
}
hipError_t dgx_launch_source(const void *args, int fp32, int nsrc, double w, hipStream_t st)
{
    if (fp32) dgx::k_source<float><<<1, 64, 0, st>>>(*(const StepArgs<float> *)args, nsrc, w);
    else dgx::k_source<double><<<1, 64, 0, st>>>(*(const StepArgs<double> *)args, nsrc, w);
    return hipGetLastError();
}
hipError_t dgx_launch_record(const void *args, int fp32, int t, int nshots, int nrec, hipStream_t st)
{
    int n = nrec * nshots;
    if (n <= 0) return hipSuccess;
    int blocks = (n + 255) / 256;
    if (fp32) dgx::k_record<float><<<blocks, 256, 0, st>>>(*(const StepArgs<float> *)args, t, nshots);
    else dgx::k_record<double><<<blocks, 256, 0, st>>>(*(const StepArgs<double> *)args, t, nshots);
    return hipGetLastError();
}
hipError_t dgx_launch_copy(const void *src, void *dst, long long n32, int variant, hipStream_t st)
{
    if (variant == 0) dgx::k_copy<<<256 * 8, 256, 0, st>>>((const double4 *)src, (double4 *)dst, n32);
    else if (variant == 1) dgx::k_copy4<<<(unsigned)((n32 + 1023) / 1024), 256, 0, st>>>((const double4 *)src, (double4 *)dst, n32);
    else dgx::k_read4<<<(unsigned)((n32 + 1023) / 1024), 256, 0, st>>>((const double4 *)src, (double *)dst, n32);
    return hipGetLastError();
}
/* Staging of dgx_get_fields_noalloc: the requested fields of one shot, reordered into the caller's element
 * order, before they are copied to the host.  One copy per GPU the library runs on, held until the process
 * ends; host code reaches it only through the address hipGetSymbolAddress returns for the current device.
 * Every entry a call reads is written by that call's reorder first, so nothing needs clearing. */
static __device__ __align__(256) double scratch[DGX_NOALLOC_SCRATCH_BYTES / sizeof(double)];

#define DGX_MAX_DEVICES 64
static std::mutex scratch_mu;
static double *scratch_dev[DGX_MAX_DEVICES]; /* per device: the address of scratch on it, once looked up */

/* Address of scratch on the current device (looked up on the first call on that device, with it current). */
hipError_t dgx_noalloc_scratch(double **ptr)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= DGX_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_dev[dev] = (double *)p;
    }
    *ptr = scratch_dev[dev];
    return hipSuccess;
}

/* Lane records of the blocked kernel, {element, v0, v1, v2}, from the lanes' element ids (host) and the
 * device vertex records: the kernel then reads an element's vertices without first reading its element id.
 * *out is allocated here (hipMalloc) and owned by the caller. */
hipError_t dgx_upload_lane_records(const int *gid, long n, const int4 *vtx, int4 **out)
{
    *out = nullptr;
    if (n <= 0) return hipSuccess;
    int *d_gid = nullptr;
    hipError_t e = hipMalloc((void **)&d_gid, sizeof(int) * n);
    if (e == hipSuccess) e = hipMemcpy(d_gid, gid, sizeof(int) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void **)out, sizeof(int4) * n);
    if (e == hipSuccess) {
        dgx::k_lane_records<<<(unsigned)((n + 255) / 256), 256>>>(d_gid, vtx, *out, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(d_gid);
    if (e != hipSuccess && *out) {
        (void)hipFree(*out);
        *out = nullptr;
    }
    return e;
}

hipError_t dgx_launch_unpermute(const void *f, const void *lo, int fp32, const int *perm, double *out, long tri,
                                long spad, int il, long shot, hipStream_t st)
{
    const unsigned blocks = (unsigned)((tri + 255) / 256);
    if (fp32) dgx::k_unpermute<float><<<blocks, 256, 0, st>>>((const float *)f, (const float *)lo, perm, out, (int)tri, spad, il, shot);
    else dgx::k_unpermute<double><<<blocks, 256, 0, st>>>((const double *)f, (const double *)lo, perm, out, (int)tri, spad, il, shot);
    return hipGetLastError();
}
/* Serving launch of the blocked kernel over the tiles of nb single-shot batches (ks_block): 512 threads per
 * workgroup in FP32, 256 in FP64 (the tiles are built for that count). */
hipError_t dgx_launch_serve_block(int fp32, int p2, const void *table, const int *sched, const int *step_base,
                                  int step_off, int launch, int nb, int total, hipStream_t st)
{
    if (total <= 0) return hipSuccess;
    const int nper = (total + DGX_NXCD - 1) / DGX_NXCD;
    const dim3 grid(DGX_NXCD * nper);
    if (fp32) {
        const auto *t = (const dgx::StepArgs<float> *)table;
        if (p2) dgx::ks_block<float, 512, 4, true><<<grid, 512, 0, st>>>(t, sched, step_base, step_off, launch, nb, total, nper);
        else dgx::ks_block<float, 512, 4, false><<<grid, 512, 0, st>>>(t, sched, step_base, step_off, launch, nb, total, nper);
    } else { /* FP64: 256-thread tiles, two workgroups per CU */
        const auto *t = (const dgx::StepArgs<double> *)table;
        if (p2) dgx::ks_block<double, 256, 2, true><<<grid, 256, 0, st>>>(t, sched, step_base, step_off, launch, nb, total, nper);
        else dgx::ks_block<double, 256, 2, false><<<grid, 256, 0, st>>>(t, sched, step_base, step_off, launch, nb, total, nper);
    }
    return hipGetLastError();
}
hipError_t dgx_launch_advance(int *step_base, int n, hipStream_t st)
{
    dgx::k_advance<<<1, 64, 0, st>>>(step_base, n);
    return hipGetLastError();
}
size_t dgx_step_args_size(int fp32) { return fp32 ? sizeof(StepArgs<float>) : sizeof(StepArgs<double>); }
}
