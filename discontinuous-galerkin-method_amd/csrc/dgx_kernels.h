/*
 * dgx_kernels.h -- device-side data layout shared by the HIP kernels and the host code of libdgx.
 *
 * Element order on the device is a locality order (Hilbert order of element centroids) chosen at
 * dgx_create; all per-element arrays use that order.  Fields of one shot are stored in 64-element
 * blocks, node-major inside a block: value (q, i) lives at ((q/64)*6 + i)*64 + q%64, so one
 * wavefront loading node i of its 64 elements reads 64 consecutive values.
 */
#ifndef DGX_KERNELS_H
#define DGX_KERNELS_H

#include <stdint.h>

#define DGX_BLOCK 256          /* threads (= elements) per workgroup in the plain kernels */
#define DGX_BLOCK_LEVELS 8     /* level table entries per tile of the blocked kernel (2 * steps + 1 <= 8) */
#define DGX_BLOCK_KMAX 3       /* most steps per launch of the blocked kernel */
#define DGX_NXCD 8             /* MI355X accelerator complex dies; workgroups are dealt round-robin */

/* Per-element code word (the .w of the int4 neighbour record):
 *   bits 3j..3j+2  face j: matched neighbour face (2 bits) | reversed (1 bit)
 *   bit  9+j       neighbour across face j needs conditioning (sponge or explosive source)
 *   bit  12        own element lies in the sponge (some factor != 1)
 *   bits 13..31    aux record index + 1 (0 = no source / receiver on this element) */
#define DGX_CODE_NBSPECIAL(j) (1u << (9 + (j)))
#define DGX_CODE_SPONGE (1u << 12)
#define DGX_CODE_AUX_SHIFT 13
#define DGX_MAX_AUX ((1u << 19) - 2)
/* serving: per-element word of a batch (StepArgs::aux_of_elem) */
#define DGX_DYN_AUX_MASK ((1u << 29) - 1)
#define DGX_DYN_NB_SHIFT 29

/* Aux record of one element: its sources and receivers as index ranges. */
typedef struct {
    int src_begin, src_count, rec_begin, rec_count;
} dgx_aux;

typedef struct {
    int elem;    /* device element index */
    int kind;    /* 0 force (Vz), 1 explosive (Txx, Tzz) */
    int shot;
    int wavelet; /* wavelet row */
    double fun[6];
} dgx_src_dev;

typedef struct {
    int elem;  /* device element index */
    int index; /* receiver index in the caller's list */
    int pad0, pad1;
    double w[6];
} dgx_rec_dev;


#ifdef __cplusplus
#include <hip/hip_runtime.h>
namespace dgx {

template <typename R>
struct Ops {
    R Dr[36], Ds[36];
    R L[54]; /* L[j*18 + i*3 + k] = S_{j+1}[i][k] */
};

template <typename R>
struct StepArgs {
    Ops<R> op;
    R *f[5];               /* Vx, Vz, Txx, Tzz, Txz of shot 0 */
    long long fstride;     /* elements between shots in each field array */
    const int4 *nbr;       /* {nb0, nb1, nb2, code} */
    const int4 *vtx;       /* {v0, v1, v2, -} */
    const double2 *pt;     /* vertex coordinates */
    const R *rho_inv, *lam, *mu;
    const R *absorb;       /* blocked like a field, read only where the sponge bit is set */
    const dgx_aux *aux;
    const dgx_src_dev *src;
    const dgx_rec_dev *rec;
    const double *wav;     /* [nwavelets][nsteps] */
    double *trace;         /* [shot][step][2][nrec] */
    const int *step_base;  /* device step counter; this launch works on step *step_base + step_off */
    int step_off;
    int nsteps, nrec;
    int tri, nper;         /* elements; blocks per XCD slice */
    double dt;
    /* shot-innermost layout only */
    int spad;              /* padded shot count: stride between nodes */
    int nshots;            /* real shots; padded lanes s >= nshots never record */
    int shot_tile;         /* shots per lane group (ST); a wavefront covers 64/ST elements */
    int wg_waves;          /* wavefronts per workgroup of the batched kernels (2, 4 or 8) */
    int il;                /* shot-innermost rows interleave 2^il elements (0: one element per row) */
    const int *shot_start; /* optional per-shot first global step (NULL: all shots start at 0) */
    const int *shot_len;   /* optional per-shot step count */
    /* serving only (NULL otherwise): sources/receivers change as requests come and go, so they are
     * indexed per element by aux_of_elem instead of the code word (bits 0..28: aux record index + 1,
     * 0 = none; bit 29+j: the neighbour across face j has an explosive source of this batch, so its
     * stresses need conditioning when gathered), every receiver record
     * names its slot (dgx_rec_dev.pad0), and each slot has its own wavelet and trace buffers */
    const int *aux_of_elem;
    const double *const *slot_wav;
    double *const *slot_trace;
    const int *slot_nrec;
    /* full-step kernel only (element layout): the step reads f and writes the other copy g */
    R *g[5];
    const int4 *nb2;       /* per element {row in g of face j's neighbour velocities (j = 0..2), code} */
    const int4 *tiles;     /* per tile {first element, own elements, first ring entry, ring elements} */
    const int *ring;       /* ring entries: element ids, tile after tile */
    int ntile, tper;       /* tiles; tiles per XCD slice */
    int vrow0;             /* velocity row of ring entry 0 (rows below it are elements) */
    int p2;                /* element layout: op equals the compiled-in P2 table (dgx_p2ops.h) */
    /* blocked kernel only (element layout, several steps per launch; tiles/ntile/tper above describe its
     * tiles {first own element, own elements, first lane record, lanes}) */
    const int *blevels;    /* per tile DGX_BLOCK_LEVELS ints: end lane of each level (hops from the own elements) */
    const int4 *blrec;     /* per lane record: {element, its vertices v0, v1, v2} */
    const int4 *blnb;      /* per lane record: {lane of face j's neighbour (j = 0..2), code}; -1 boundary */
    R *fl[5], *gl[5];      /* compensated FP32 (blocked kernel): the low words of f and g; NULL otherwise */
    int bkmax;             /* steps the tiles' levels allow per launch */
    int bthreads;          /* threads per workgroup (lanes per tile at most) */
    int bsteps;            /* steps of this launch (1..bkmax) */
};

} // namespace dgx
#endif

#endif
