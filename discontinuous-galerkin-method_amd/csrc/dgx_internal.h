/*
 * dgx_internal.h -- host-side pieces shared by the solver (dgx_api.cpp) and the server (dgx_serve.cpp).
 * Not part of the C-ABI.
 */
#ifndef DGX_INTERNAL_H
#define DGX_INTERNAL_H

#include <hip/hip_runtime.h>

#include <vector>

#include "dgx.h"

/* Mesh as uploaded: device element order, face codes, renumbered vertices, materials, sponge. */
struct dgx_mesh_host {
    long tri = 0, tri_pad = 0;
    std::vector<long> perm; /* device element -> caller element */
    std::vector<long> inv;  /* caller element -> device element */
    std::vector<int4> nbr;  /* {nb0, nb1, nb2, code} */
    std::vector<int4> vtx;
    std::vector<double2> pt;
    std::vector<double> mat;    /* 1/rho | lambda | mu, tri_pad each */
    std::vector<double> absorb; /* 6 * tri_pad */
};

/* Sponge factors as the FP32 kernels read them: the deficit 1 - f, rounded once (the device helper sponge()
 * applies x - x (1 - f)).  FP64 kernels read f itself. */
inline std::vector<float> dgx_sponge_deficit_f32(const std::vector<double> &f)
{
    std::vector<float> d(f.size());
    for (size_t k = 0; k < f.size(); k++) d[k] = (float)(1.0 - f[k]);
    return d;
}

int dgx_build_mesh_host(const dgx_problem *p, int shot_inner, long tri_pad, const std::vector<char> *explosive,
                        dgx_mesh_host *out);

extern "C" void dgx_set_error(const char *fmt, ...);

/* Host-only view of the full-step tiling of a mesh in the element layout (no device needed): counts
 * when the array pointers are NULL, else also tiles [ntile], ring [nring], nb2 [tri] and the device-order
 * neighbour records nbr [tri].  For tests of the tiling's invariants; not part of the C-ABI. */
extern "C" int dgx_full_step_tiling(const dgx_problem *p, long *ntile, long *nring, long *vrow0, long *vrows,
                                    int4 *tiles, int *ring, int4 *nb2, int4 *nbr);

/* Host-only view of the blocked kernel's tiling of a mesh (element layout, K steps per launch, `cap` threads
 * per workgroup): counts when the array pointers are NULL, else also tiles [ntile], levels
 * [ntile * DGX_BLOCK_LEVELS], gid and lnb [nlanes] and the device-order neighbour records nbr [tri].  For tests
 * of the tiling's invariants; not part of the C-ABI. */
extern "C" int dgx_block_tiling(const dgx_problem *p, int K, int cap, long *ntile, long *nlanes, int4 *tiles,
                                int *levels, int *gid, int4 *lnb, int4 *nbr);

#endif
