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

int dgx_build_mesh_host(const dgx_problem *p, int shot_inner, long tri_pad, const std::vector<char> *explosive,
                        dgx_mesh_host *out);

/* The blocked kernel's lane records {element, v0, v1, v2}, built on the device from the lanes' element ids (n, host)
 * and the device vertex records (dgx_plain.hip); *out is allocated there and owned by the caller. */
extern "C" hipError_t dgx_upload_lane_records(const int *gid, long n, const int4 *vtx, int4 **out);

/* The same mesh built on the device (dgx_meshdev.hip), bit for bit the arrays of dgx_build_mesh_host: nbr, vtx,
 * pt (npt used vertices), mat and absorb (FP64, in the layout of shot_inner) stay in device memory, owned by the
 * caller (dgx_mesh_dev_free); perm, inv and a copy of nbr come back to the host. */
struct dgx_mesh_dev {
    long tri = 0, tri_pad = 0, npt = 0;
    std::vector<long> perm, inv;
    std::vector<int4> nbr;
    int4 *d_nbr = nullptr, *d_vtx = nullptr;
    double2 *d_pt = nullptr;
    double *d_mat = nullptr, *d_absorb = nullptr;
};

int dgx_build_mesh_device(const dgx_problem *p, int shot_inner, long tri_pad, const std::vector<char> *explosive,
                          dgx_mesh_dev *out);
void dgx_mesh_dev_free(dgx_mesh_dev *m);
/* FP32 copies on the device for the FP32 kernels: materials rounded once, sponge factors as their deficit 1 - f
 * rounded once (the device helper sponge() applies x - x (1 - f)). */
hipError_t dgx_mesh_f32(const double *mat, long nmat, const double *absorb, long nab, float **mat32, float **absorb32);
/* A mesh's materials 1/rho | lambda | mu with stride tri, from the build's stride tri_pad, in FP64 or FP32. */
hipError_t dgx_mesh_mat_packed(const double *mat, long tri_pad, long tri, int fp32, void **out);

/* Both builds of one mesh compared array by array, bit for bit (tests): *mismatch = entries that differ.  Not
 * part of the C-ABI. */
extern "C" int dgx_mesh_build_compare(const dgx_problem *p, int shot_inner, long tri_pad, const char *explosive,
                                      long *mismatch);

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
