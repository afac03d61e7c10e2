/*
 * dgx_reference.hip -- the FP64 step of DGX_FP64_REFERENCE: the reference's arithmetic, operation for operation.
 *
 * The fast FP64 kernels regroup the reference's sums (metric terms factored out of Dx/Dz, 1/D folded into the face
 * scalings, fused multiply-adds); their results differ from the reference's in the last bits, and over thousands
 * of steps by about 1e-12 of the field maxima.  These kernels instead use the reference's own per-element values
 * (Dx, Dz, Q, A, B as its setup forms them; Q, A and B computed on the host by dgx_reference_operators) and
 * evaluate every sum in the reference's order with every product rounded on its own, so that each value equals
 * the reference's bit for bit (up to the sign of exact zeros in the face jumps, which its additions erase):
 *   velocity  OpenMP_Function.c:75-185    stress  OpenMP_Function.c:188-306
 *   source    OpenMP_Function.c:49-58     record  OpenMP_Function.c:61-72 (P2 weights; a vertex receiver is the
 *                                                  reference's nodal pick)
 * Contraction into fused multiply-adds is switched off for this whole file.
 *
 * One thread per element and shot, fields in the element layout of the other kernels (64-element blocks,
 * node-major).  Dx and Dz (72 values per element) are not stored: each thread forms them row by row from its
 * vertex coordinates and the reference-element matrices Vr, Vs, inverV exactly as the reference's setup does
 * (Jacobian, fun_DG.c:421-455; Dx_Dz, fun_DG.c:1137-1192: the same products, quotients and sums in the same
 * order, all of them correctly rounded IEEE operations, so the values are its values), which costs about 1100
 * FP64 operations per element and half-step instead of 576 bytes of memory traffic.  The rows run in a loop with
 * the matrices read from LDS per row, which keeps the kernels at 2-3 wavefronts per SIMD.  Q and the materials are
 * read: entry k of element q lives at ((q/64)*34 + k)*64 + q%64, k = Q 0..29 | 1/rho, lambda + 2 mu, lambda, mu
 * 30..33 (the values of the reference's A and B: A0 = A1 = B0 = B1 = 1/rho, A2 = B3 = lambda + 2 mu,
 * A3 = B2 = lambda, A4 = B4 = mu; Elastic_A_B, fun_DG.c:1071-1090), so a wavefront reading entry k of its 64
 * elements reads 512 contiguous bytes.
 *
 * A step (dgx_step) is two kernels, FUSED: the velocity kernel applies the sponge and the step's sources to every
 * value it loads, and writes the new velocities; the stress kernel applies them again to the stresses it updates
 * (the same operations on the same stored values, so the same results) and records the receivers.  That is the
 * reference's order, absorb, sources, velocity, stress, record, without separate passes.  The split entry points
 * run the unfused kernels and the one-thread source and record kernels, one per reference function.
 */
#include <hip/hip_runtime.h>

#include "dgx.h"
#include "dgx_device.h"
#include "dgx_kernels.h"

#pragma clang fp contract(off)

namespace dgx {
namespace exact {

constexpr int kWidth = 34; /* per-element reals: Q 30 | 1/rho | lambda + 2 mu | lambda | mu */
constexpr int kQ = 0, kRinv = 30, kL2m = 31, kLam = 32, kMu = 33;

/* The reference-element matrices the reference's setup forms Dx and Dz from (row-major, as read from its
 * operator files). */
struct RefMats {
    double Vr[36], Vs[36], iV[36];
};

__device__ __forceinline__ long long fidx(int q, int i) { return ((long long)(q >> 6) * 6 + i) * 64 + (q & 63); }
__device__ __forceinline__ long long ridx(int q, int k) { return ((long long)(q >> 6) * kWidth + k) * 64 + (q & 63); }

/* Elements dealt to the 8 XCDs in contiguous slices, as in the other element-layout kernels. */
__device__ __forceinline__ int element_of_thread(int nper)
{
    const int b = blockIdx.x;
    return ((b & (DGX_NXCD - 1)) * nper + (b >> 3)) * DGX_BLOCK + threadIdx.x;
}

/* Face j's neighbour (device index, -1 on the outer boundary) from the neighbour record. */
__device__ __forceinline__ int neighbour(const int4 &nb, int j) { return j == 0 ? nb.x : j == 1 ? nb.y : nb.z; }

/* Metric terms as the reference computes them: half edge vectors, J = xr ys - xs yr (Jacobian, fun_DG.c:421-455),
 * Rx = ys/J, Rz = -xs/J, Sx = -yr/J, Sz = xr/J (Dx_Dz, fun_DG.c:1137-1192). */
struct Metric {
    double Rx, Rz, Sx, Sz;
};

__device__ __forceinline__ Metric metric(const double2 *pt, int4 v)
{
    const double2 p0 = pt[v.x], p1 = pt[v.y], p2 = pt[v.z];
    const double xr = (p1.x - p0.x) * 0.5, yr = (p1.y - p0.y) * 0.5;
    const double xs = (p2.x - p0.x) * 0.5, ys = (p2.y - p0.y) * 0.5;
    const double J = xr * ys - xs * yr;
    return Metric{ys / J, -xs / J, -yr / J, xr / J};
}

/* Row p of Dx and Dz: g = R Vr + S Vs elementwise, then the row of g times inverV, summed from zero in k order
 * (Dx_Dz, fun_DG.c:1137-1192). */
__device__ __forceinline__ void dxdz_row(const double *m, const Metric &g, int p, double (&dx)[6], double (&dz)[6])
{
    double gx[6], gz[6];
    int o = 0;
    asm volatile("" : "+v"(o)); /* an offset the compiler cannot see through: the matrices are read in each row,
                                 * not held in registers across the row loop */
    m += o;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        gx[k] = g.Rx * m[p * 6 + k] + g.Sx * m[36 + p * 6 + k];
        gz[k] = g.Rz * m[p * 6 + k] + g.Sz * m[36 + p * 6 + k];
    }
#pragma unroll
    for (int q = 0; q < 6; q++) {
        double f1 = 0, f2 = 0;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            f1 = f1 + gx[k] * m[72 + k * 6 + q];
            f2 = f2 + gz[k] * m[72 + k * 6 + q];
        }
        dx[q] = f1;
        dz[q] = f2;
    }
}

/* a[0..4] = a[1..5], a[5] = v */
__device__ __forceinline__ void shift_in(double (&a)[6], double v)
{
#pragma unroll
    for (int k = 0; k < 5; k++) a[k] = a[k + 1];
    a[5] = v;
}

/* Vr | Vs | inverV into LDS: read there with one address per wavefront, they take no registers between uses (as
 * arguments, the 108 values overflow the scalar registers and the spills cost vector registers). */
__device__ __forceinline__ void stage_matrices(const RefMats &mats, double *sm)
{
    const double *src = &mats.Vr[0];
    for (int k = threadIdx.x; k < 108; k += blockDim.x) sm[k] = src[k];
    __syncthreads();
}

/* The sponge and the sources of the step, applied to values as a kernel loads them (fused step), with the
 * reference's operations in its order: x f (Absorb_Boundary, OpenMP_Function.c:25-46; elements outside the sponge
 * have f = 1 and are skipped), then x + fun w for the element's sources in list order (source_kernel, :49-58).
 * The aux record lists an element's sources in the caller's order.  Both half-step kernels condition the same
 * stored values with the same operations, so they see the same conditioned values. */
/* Stresses of N nodes (node[k]) of element q conditioned in place: the sponge factor of each node, then the
 * element's explosive sources in list order. */
template <int N>
__device__ __forceinline__ void cond_stress(const StepArgs<double> &a, unsigned code, int q, const int (&node)[N],
                                            int shot, int step, double (&txx)[N], double (&tzz)[N], double (&txz)[N])
{
    if (code & DGX_CODE_SPONGE) {
#pragma unroll
        for (int k = 0; k < N; k++) {
            const double f = a.absorb[fidx(q, node[k])];
            txx[k] = txx[k] * f; tzz[k] = tzz[k] * f; txz[k] = txz[k] * f;
        }
    }
    const unsigned ax = code >> DGX_CODE_AUX_SHIFT;
    if (ax) {
        const dgx_aux x = a.aux[ax - 1];
        for (int s = x.src_begin; s < x.src_begin + x.src_count; s++) {
            const dgx_src_dev &src = a.src[s];
            if (src.shot != shot || src.kind != DGX_SRC_EXPLOSIVE) continue;
            const double w = a.wav[(long long)src.wavelet * a.nsteps + step];
#pragma unroll
            for (int k = 0; k < N; k++) {
                const double inc = src.fun[node[k]] * w;
                txx[k] = txx[k] + inc;
                tzz[k] = tzz[k] + inc;
            }
        }
    }
}

/* Velocities of element q conditioned in place: the sponge, then the element's force sources in list order. */
__device__ __forceinline__ void cond_velocity(const StepArgs<double> &a, unsigned code, int q, int shot, int step,
                                              double (&vx)[6], double (&vz)[6])
{
    if (code & DGX_CODE_SPONGE) {
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const double f = a.absorb[fidx(q, i)];
            vx[i] = vx[i] * f; vz[i] = vz[i] * f;
        }
    }
    const unsigned ax = code >> DGX_CODE_AUX_SHIFT;
    if (ax) {
        const dgx_aux x = a.aux[ax - 1];
        for (int s = x.src_begin; s < x.src_begin + x.src_count; s++) {
            const dgx_src_dev &src = a.src[s];
            if (src.shot != shot || src.kind != DGX_SRC_FORCE_Z) continue;
            const double w = a.wav[(long long)src.wavelet * a.nsteps + step];
#pragma unroll
            for (int i = 0; i < 6; i++) vz[i] = vz[i] + src.fun[i] * w;
        }
    }
}

constexpr int kAllNodes[6] = {0, 1, 2, 3, 4, 5};

/* LU_Velocity (OpenMP_Function.c:75-185):
 *   Vx = (A0 (Dx Txx) + B0 (Dz Txz) + sum_j (Q_j0 S_j dTxx_j + Q_j1 S_j dTxz_j)) dt + Vx
 *   Vz = (A1 (Dx Txz) + B1 (Dz Tzz) + sum_j (Q_j2 S_j dTzz_j + Q_j3 S_j dTxz_j)) dt + Vz
 * with the jump dU = flag U_nb - U_own (flag 0 on the boundary). */
template <bool FUSED>
__global__ void __launch_bounds__(DGX_BLOCK, FUSED ? 2 : 3)
    k_ref_velocity(StepArgs<double> a, const double *__restrict__ ops, RefMats mats)
{
    __shared__ double sm[108];
    stage_matrices(mats, sm);
    const int q = element_of_thread(a.nper);
    if (q >= a.tri) return;
    const long long so = (long long)blockIdx.y * a.fstride * 6;
    double *__restrict__ Vx = a.f[0] + so;
    double *__restrict__ Vz = a.f[1] + so;
    const double *__restrict__ Txx = a.f[2] + so;
    const double *__restrict__ Tzz = a.f[3] + so;
    const double *__restrict__ Txz = a.f[4] + so;
    const int shot = blockIdx.y, step = FUSED ? *a.step_base + a.step_off : 0;
    const int4 nb = a.nbr[q];
    const unsigned code = (unsigned)nb.w;
    double tx[6], tz[6], tc[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const long long o = fidx(q, i);
        tx[i] = Txx[o]; tz[i] = Tzz[o]; tc[i] = Txz[o];
    }
    if (FUSED) cond_stress(a, code, q, kAllNodes, shot, step, tx, tz, tc);
    const Metric g = metric(a.pt, a.vtx[q]);
    const double rinv = ops[ridx(q, kRinv)]; /* A0 = A1 = B0 = B1 */
    /* rows one at a time (a loop, not unrolled: the unrolled rows are scheduled together and need more registers
     * than a SIMD has), each result shifted into the last slot so that row i ends in slot i */
    double accx[6] = {}, accz[6] = {};
#pragma unroll 1
    for (int i = 0; i < 6; i++) {
        double dx[6], dz[6];
        dxdz_row(sm, g, i, dx, dz);
        double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
#pragma unroll
        for (int n = 0; n < 6; n++) {
            s1 = s1 + dx[n] * tx[n];
            s2 = s2 + dx[n] * tc[n];
            s3 = s3 + dz[n] * tz[n];
            s4 = s4 + dz[n] * tc[n];
        }
        shift_in(accx, rinv * s1 + rinv * s4);
        shift_in(accz, rinv * s2 + rinv * s3);
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int e = neighbour(nb, j);
        const unsigned m = (code >> (3 * j)) & 3u, rev = (code >> (3 * j + 2)) & 1u;
        double nxx[3] = {}, nzz[3] = {}, nxz[3] = {};
        int nn[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            nn[k] = nb_node(m, rev, k);
            if (e >= 0) {
                const long long o = fidx(e, nn[k]);
                nxx[k] = Txx[o]; nzz[k] = Tzz[o]; nxz[k] = Txz[o];
            }
        }
        if (FUSED && e >= 0 && (code & DGX_CODE_NBSPECIAL(j)))
            cond_stress(a, (unsigned)a.nbr[e].w, e, nn, shot, step, nxx, nzz, nxz);
        double jxx[3], jzz[3], jxz[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int on = face_node(j, k);
            jxx[k] = nxx[k] - tx[on];
            jzz[k] = nzz[k] - tz[on];
            jxz[k] = nxz[k] - tc[on];
        }
        const double q0 = ops[ridx(q, kQ + j * 10 + 0)], q1 = ops[ridx(q, kQ + j * 10 + 1)];
        const double q2 = ops[ridx(q, kQ + j * 10 + 2)], q3 = ops[ridx(q, kQ + j * 10 + 3)];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            double sa = 0, sb = 0, sc = 0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double sv = a.op.L[j * 18 + i * 3 + k];
                sa = sa + sv * jxx[k];
                sb = sb + sv * jzz[k];
                sc = sc + sv * jxz[k];
            }
            accx[i] = accx[i] + (q0 * sa + q1 * sc);
            accz[i] = accz[i] + (q2 * sb + q3 * sc);
        }
    }
    double vx[6], vz[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const long long o = fidx(q, i);
        vx[i] = Vx[o]; vz[i] = Vz[o];
    }
    if (FUSED) cond_velocity(a, code, q, shot, step, vx, vz);
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const long long o = fidx(q, i);
        Vx[o] = accx[i] * a.dt + vx[i];
        Vz[o] = accz[i] * a.dt + vz[i];
    }
}

/* LU_Stress_Elastic (OpenMP_Function.c:188-306):
 *   Txx = (A2 (Dx Vx) + B2 (Dz Vz) + sum_j (Q_j4 S_j dVx_j + Q_j5 S_j dVz_j)) dt + Txx
 *   Tzz = (A3 (Dx Vx) + B3 (Dz Vz) + sum_j (Q_j6 S_j dVx_j + Q_j7 S_j dVz_j)) dt + Tzz
 *   Txz = (A4 (Dx Vz) + B4 (Dz Vx) + sum_j (Q_j8 S_j dVx_j + Q_j9 S_j dVz_j)) dt + Txz */
template <bool FUSED>
__global__ void __launch_bounds__(DGX_BLOCK, FUSED ? 2 : 3)
    k_ref_stress(StepArgs<double> a, const double *__restrict__ ops, RefMats mats)
{
    __shared__ double sm[108];
    stage_matrices(mats, sm);
    const int q = element_of_thread(a.nper);
    if (q >= a.tri) return;
    const long long so = (long long)blockIdx.y * a.fstride * 6;
    const double *__restrict__ Vx = a.f[0] + so;
    const double *__restrict__ Vz = a.f[1] + so;
    double *__restrict__ Txx = a.f[2] + so;
    double *__restrict__ Tzz = a.f[3] + so;
    double *__restrict__ Txz = a.f[4] + so;
    double vx[6], vz[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const long long o = fidx(q, i);
        vx[i] = Vx[o]; vz[i] = Vz[o];
    }
    const Metric g = metric(a.pt, a.vtx[q]);
    const double l2m = ops[ridx(q, kL2m)], lam = ops[ridx(q, kLam)], mu = ops[ridx(q, kMu)];
    const int shot = blockIdx.y, step = FUSED ? *a.step_base + a.step_off : 0;
    double axx[6] = {}, azz[6] = {}, axz[6] = {};
#pragma unroll 1
    for (int i = 0; i < 6; i++) { /* rows one at a time, as in the velocity kernel */
        double dx[6], dz[6];
        dxdz_row(sm, g, i, dx, dz);
        double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
#pragma unroll
        for (int n = 0; n < 6; n++) {
            s1 = s1 + dx[n] * vx[n];
            s2 = s2 + dx[n] * vz[n];
            s3 = s3 + dz[n] * vx[n];
            s4 = s4 + dz[n] * vz[n];
        }
        shift_in(axx, l2m * s1 + lam * s4); /* A2, B2 */
        shift_in(azz, lam * s1 + l2m * s4); /* A3, B3 */
        shift_in(axz, mu * s2 + mu * s3);   /* A4, B4 */
    }
    const int4 nb = a.nbr[q];
    const unsigned code = (unsigned)nb.w;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int e = neighbour(nb, j);
        const unsigned m = (code >> (3 * j)) & 3u, rev = (code >> (3 * j + 2)) & 1u;
        double jx[3], jz[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int on = face_node(j, k);
            double nx = 0, nz = 0;
            if (e >= 0) {
                const long long o = fidx(e, nb_node(m, rev, k));
                nx = Vx[o]; nz = Vz[o];
            }
            jx[k] = nx - vx[on];
            jz[k] = nz - vz[on];
        }
        double qq[6];
#pragma unroll
        for (int c = 0; c < 6; c++) qq[c] = ops[ridx(q, kQ + j * 10 + 4 + c)];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            double sa = 0, sb = 0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double sv = a.op.L[j * 18 + i * 3 + k];
                sa = sa + sv * jx[k];
                sb = sb + sv * jz[k];
            }
            axx[i] = axx[i] + (qq[0] * sa + qq[1] * sb);
            azz[i] = azz[i] + (qq[2] * sa + qq[3] * sb);
            axz[i] = axz[i] + (qq[4] * sa + qq[5] * sb);
        }
    }
    double txx[6], tzz[6], txz[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const long long o = fidx(q, i);
        txx[i] = Txx[o]; tzz[i] = Tzz[o]; txz[i] = Txz[o];
    }
    if (FUSED) cond_stress(a, code, q, kAllNodes, shot, step, txx, tzz, txz);
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const long long o = fidx(q, i);
        Txx[o] = axx[i] * a.dt + txx[i];
        Tzz[o] = azz[i] * a.dt + tzz[i];
        Txz[o] = axz[i] * a.dt + txz[i];
    }
    /* Record (OpenMP_Function.c:61-72) of this element's receivers: the velocities are final for the step */
    const unsigned ax = code >> DGX_CODE_AUX_SHIFT;
    if (FUSED && ax) {
        const dgx_aux x = a.aux[ax - 1];
        double *tr = a.trace + ((long long)shot * a.nsteps + step) * 2 * a.nrec;
        for (int r = x.rec_begin; r < x.rec_begin + x.rec_count; r++) {
            const dgx_rec_dev &rc = a.rec[r];
            double sx = 0, sz = 0;
            for (int i = 0; i < 6; i++) {
                sx = sx + rc.w[i] * vx[i];
                sz = sz + rc.w[i] * vz[i];
            }
            tr[rc.index] = sx;
            tr[a.nrec + rc.index] = sz;
        }
    }
}

/* source_kernel (OpenMP_Function.c:49-58) for every source, in list order, in one thread: V = V + fun w, the
 * product rounded first.  w: the sample of the source's wavelet at the device step (*step_base + step_off), or
 * the given value when `fixed` is set. */
__global__ void k_ref_source(StepArgs<double> a, int nsrc, int fixed, double wfix)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int step = *a.step_base + a.step_off;
    for (int s = 0; s < nsrc; s++) {
        const dgx_src_dev &src = a.src[s];
        const double w = fixed ? wfix : a.wav[(long long)src.wavelet * a.nsteps + step];
        const long long so = (long long)src.shot * a.fstride * 6;
        for (int i = 0; i < 6; i++) {
            const double inc = src.fun[i] * w;
            const long long o = so + fidx(src.elem, i);
            if (src.kind == 0) {
                a.f[1][o] = a.f[1][o] + inc;
            } else {
                a.f[2][o] = a.f[2][o] + inc;
                a.f[3][o] = a.f[3][o] + inc;
            }
        }
    }
}

/* Record (OpenMP_Function.c:61-72): trace = sum_i w_i V_i in node order, into slot t, or into the device step's
 * slot when t < 0. */
__global__ void k_ref_record(StepArgs<double> a, int t, int nshots)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nrec * nshots) return;
    if (t < 0) t = *a.step_base + a.step_off;
    const int shot = k / a.nrec, r = k % a.nrec;
    const dgx_rec_dev &rc = a.rec[r];
    const long long so = (long long)shot * a.fstride * 6;
    double sx = 0, sz = 0;
    for (int i = 0; i < 6; i++) {
        const long long o = so + fidx(rc.elem, i);
        sx = sx + rc.w[i] * a.f[0][o];
        sz = sz + rc.w[i] * a.f[1][o];
    }
    double *tr = a.trace + ((long long)shot * a.nsteps + t) * 2 * a.nrec;
    tr[rc.index] = sx;
    tr[a.nrec + rc.index] = sz;
}

} // namespace exact
} // namespace dgx

extern "C" {
/* which: 0 velocity, 1 stress, 2 sources at the device step, 3 sources with the sample w, 4 record into the
 * device step's slot, 5 record into slot t; 6 and 7 the velocity and stress kernels of the fused step (sponge and
 * sources applied as values are loaded, receivers recorded by the stress kernel), sources read from the aux-ordered
 * list.  ops: the per-element Q and materials; mats: Vr | Vs | inverV, 108
 * doubles. */
hipError_t dgx_launch_reference(const void *args, const double *ops, const double *mats, int which, int nshots,
                                int nsrc, int nrec, int t, double w, hipStream_t st)
{
    const dgx::StepArgs<double> &a = *(const dgx::StepArgs<double> *)args;
    const dim3 grid(DGX_NXCD * a.nper, nshots);
    dgx::exact::RefMats m;
    static_assert(sizeof m == 108 * sizeof(double), "RefMats is Vr | Vs | inverV, packed");
    if (which <= 1 || which >= 6) __builtin_memcpy(&m, mats, sizeof m);
    switch (which) {
    case 0: dgx::exact::k_ref_velocity<false><<<grid, DGX_BLOCK, 0, st>>>(a, ops, m); break;
    case 1: dgx::exact::k_ref_stress<false><<<grid, DGX_BLOCK, 0, st>>>(a, ops, m); break;
    case 6: dgx::exact::k_ref_velocity<true><<<grid, DGX_BLOCK, 0, st>>>(a, ops, m); break;
    case 7: dgx::exact::k_ref_stress<true><<<grid, DGX_BLOCK, 0, st>>>(a, ops, m); break;
    case 2:
    case 3:
        if (nsrc > 0) dgx::exact::k_ref_source<<<1, 64, 0, st>>>(a, nsrc, which == 3, w);
        break;
    case 4:
    case 5:
        if (nrec > 0) dgx::exact::k_ref_record<<<(nrec * nshots + 255) / 256, 256, 0, st>>>(a, which == 5 ? t : -1,
                                                                                         nshots);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
}
