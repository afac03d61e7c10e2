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

extern "C" void dgx_set_error(const char *fmt, ...);

#endif
