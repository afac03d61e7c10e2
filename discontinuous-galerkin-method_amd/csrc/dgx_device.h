/*
 * dgx_device.h -- element geometry and face numbering shared by the element-layout and shot-innermost
 * kernels (device code only).
 *
 * Geometry is recomputed in FP64 from the three vertex coordinates; with D = 2*area = 4J (J as in
 * fun_DG.c:447):
 *   rx = 2(z2-z0)/D, rz = -2(x2-x0)/D, sx = -2(z1-z0)/D, sz = 2(x1-x0)/D     (fun_DG.c:1157-1165)
 *   face j with counter-clockwise edge vector (ex_j, ez_j):
 *   0.5 * n_j * (L_j/2) / J = (ez_j, -ex_j) / D                              (fun_DG.c:1118-1131)
 */
#ifndef DGX_DEVICE_H
#define DGX_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dgx {

/* Pointers read from a device-side table (the serving batch table, the per-slot wavelet and trace tables) are
 * generic, so every access through them would be a flat memory instruction.  They all point to device memory
 * (hipMalloc), where a global pointer has the same value as the generic one: rebuilding each from its address
 * as a global pointer lets the compiler use global instructions for every access derived from it. */
template <typename T>
__device__ __forceinline__ T *as_global(T *p)
{
    return (T *)(__attribute__((address_space(1))) T *)(uintptr_t)p;
}

/* Sponge factor applied to a value (Absorb_Boundary, OpenMP_Function.c:25-46).  FP64 arrays hold the factor
 * f itself, so the product is the reference's bit for bit.  FP32 arrays hold its deficit d = 1 - f (f lies in
 * [0.9755, 1] for the reference's beta = 0.1, so d in [0, 0.0245] is represented about 40x more accurately
 * than f), applied as x - x d in one fused operation: the factor's representation error, which compounds
 * once per step, stays below 2e-9 instead of 3e-8. */
template <typename R>
__device__ __forceinline__ R sponge(R x, R a)
{
    if constexpr (sizeof(R) == 8) return x * a;
    else return __builtin_fmaf(-x, a, x);
}

/* Local node of face j, position k, in the reference's face order (ibe, fun_DG.c:1048-1056). */
__device__ __forceinline__ constexpr int face_node(int j, int k)
{
    return (j == 0) ? (k == 0 ? 5 : k == 1 ? 3 : 0) : (j == 1) ? k : (k == 0 ? 5 : k == 1 ? 4 : 2);
}

/* Neighbour's node facing own face position k, given the neighbour's face m and orientation. */
__device__ __forceinline__ int nb_node(unsigned m, unsigned rev, int k)
{
    const unsigned long long tab = 0x245210035ull; /* nibbles: 5,3,0, 0,1,2, 5,4,2 (low first) */
    const int kk = rev ? 2 - k : k;
    return (int)((tab >> (4 * (3 * m + kk))) & 15);
}

struct Geom {
    double rx, rz, sx, sz, invD;
    double ex[3], ez[3];
};

__device__ __forceinline__ Geom geometry(const double2 *pt, int4 v)
{
    const double2 p0 = pt[v.x], p1 = pt[v.y], p2 = pt[v.z];
    Geom g;
    g.ex[0] = p1.x - p0.x; g.ez[0] = p1.y - p0.y;
    g.ex[1] = p2.x - p1.x; g.ez[1] = p2.y - p1.y;
    g.ex[2] = p0.x - p2.x; g.ez[2] = p0.y - p2.y;
    const double D = g.ex[0] * (p2.y - p0.y) - (p2.x - p0.x) * g.ez[0];
    g.invD = 1.0 / D;
    const double t = 2.0 * g.invD;
    g.rx = (p2.y - p0.y) * t;
    g.rz = -(p2.x - p0.x) * t;
    g.sx = -g.ez[0] * t;
    g.sz = g.ex[0] * t;
    return g;
}

/* The geometric factors rounded once to the working precision R, for kernels that keep them in registers over
 * several steps: rx, rz, sx, sz, and each face's edge vector divided by D, (ex_j, ez_j)/D, so that the face
 * terms carry their 1/D scale and accumulate into the volume terms directly. */
template <typename R>
struct GeomR {
    R rx, rz, sx, sz;
    R exd[3], ezd[3];
};

template <typename R>
__device__ __forceinline__ GeomR<R> geometry_r(const Geom &g)
{
    GeomR<R> r;
    r.rx = (R)g.rx; r.rz = (R)g.rz; r.sx = (R)g.sx; r.sz = (R)g.sz;
#pragma unroll
    for (int j = 0; j < 3; j++) { r.exd[j] = (R)(g.ex[j] * g.invD); r.ezd[j] = (R)(g.ez[j] * g.invD); }
    return r;
}

} // namespace dgx
#endif
