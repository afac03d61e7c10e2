/*
 * dgx_serve.h -- forward-modelling request loop on one GPU (libdgx.so).
 *
 * No counterpart in the reference, which runs one shot per process (DG_leap_Frog_OpenMP.c:397 loops over
 * a single shot).  The server keeps meshes, materials and sponge factors resident in HBM and runs many
 * shots of many requests at once: every request is 1..DGX_SERVE_MAX_SHOTS shots on one resident mesh, each
 * shot with its own source, FP64 wavelet samples and start step, all shots of a request sharing one
 * receiver spread and one material model (the mesh's own, or the request's).  Shots run in slot groups of
 * group_shots lanes (dgx_server_config): 4, 8, 16, 32 or 64 shots in the shot-innermost layout, stepped by one
 * kernel per half-step, or single-shot slots in the element layout, stepped two steps per launch by the
 * blocked kernel; a request's shots may spread over several groups of its mesh.
 * Requests are admitted as slots and memory free up (continuous batching), stepped by launches over the
 * tiles of all active (mesh, slot group) batches, in hipGraph-captured blocks of steps, and retired with
 * their traces copied back.  The mix of requests does not change a result: every shot's traces are bit for
 * bit those of the same request served alone.
 */
#ifndef DGX_SERVE_H
#define DGX_SERVE_H

#include "dgx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dgx_server dgx_server;

#define DGX_SERVE_MAX_SHOTS 64 /* shots per request */

typedef struct {
    int precision;         /* DGX_FP32 (default for serving) or DGX_FP64 */
    long long hbm_budget;  /* bytes for wavefields, wavelets and traces; 0 = 80 % of free HBM */
    int steps_per_block;   /* steps between scheduling points (one graph replay); 0 = 100 */
    int group_shots;       /* shots per slot group, 1, 4, 8, 16, 32 or 64 (0 = 16).  4-64: the shot-innermost
                            * serving kernel's shot tile, one launch per half-step; each node row of a group
                            * holds 64/G elements x G shots, so every G moves whole lines; 64 is fastest per
                            * lane when groups fill, 4-16 waste fewer lanes when requests are small and meshes
                            * many.  1: every shot is a slot of its own in the element layout, stepped by
                            * the blocked kernel (tiles of each mesh with recomputed rings, two copies of the
                            * fields per slot), three steps per launch in FP32 when steps_per_block is a
                            * multiple of 6, else two; steps_per_block must then be a multiple of 4 or 6 */
} dgx_server_config;

enum dgx_request_state { DGX_REQ_QUEUED = 0, DGX_REQ_RUNNING = 1, DGX_REQ_DONE = 2 };

typedef struct {
    int mesh;                /* id from dgx_server_add_mesh */
    int nshots;              /* 1..DGX_SERVE_MAX_SHOTS */
    long nsteps;             /* time steps of every shot */
    const dgx_source *src;   /* nshots sources; src[k].shot and .wavelet are ignored (shot k uses row k) */
    const double *wavelet;   /* [nshots][nsteps] FP64 source samples */
    long nrec;
    const dgx_receiver *rec; /* receivers shared by the request's shots (1-based caller element ids) */
    double arrival;          /* caller's clock (seconds) at which the request arrived, for latency */
    /* The request's material model, per element of the mesh in the caller's numbering (the arrays the
     * reference fills per element at DG_leap_Frog_OpenMP.c:236-265), and the time step its wavelet samples
     * were taken at (the reference's Calculate_Time_Step, fun_DG.c:179-199, from this model's Vp).  All NULL:
     * the mesh's own model and dt from dgx_server_add_mesh.  Requests with equal arrays and dt share one
     * resident copy, which is on the device while one of them runs and counts against the HBM budget; the
     * shots of one slot group all run on one model. */
    const double *lambda, *miu, *den;
    double dt;
} dgx_request;

typedef struct {
    long queued, running, done;
    long batches;               /* allocated (mesh, slot group) batches */
    long long bytes_in_use;     /* wavefields + per-request buffers */
    long long dof_updates;      /* useful DOF-updates advanced so far (active shots only) */
    double busy_seconds;        /* wall time spent in dgx_server_advance */
    long blocks;                /* step blocks run */
    long global_step;
    long long lane_updates;     /* DOF-updates of every slot lane stepped, idle slots of live groups
                                 * included: dof_updates / lane_updates is the slot occupancy */
    double gpu_seconds;         /* execution time of the step blocks on the device (HIP events); the
                                 * rest of busy_seconds is host work between blocks */
    long long moved;            /* shots moved into fuller groups of their mesh by compaction */
} dgx_server_stats;

int dgx_server_create(const dgx_server_config *cfg, const dgx_operators *ops, dgx_server **out);
int dgx_server_destroy(dgx_server *sv);
/* Make a mesh resident.  Uses tri, point_num, point, triangle, neighbor, lambda, miu, den, absorb_fun
 * and dt of `mesh`; its shots, steps, sources and receivers are ignored. */
int dgx_server_add_mesh(dgx_server *sv, const dgx_problem *mesh, int *mesh_id);
/* Queue a request (everything is copied).  Returns its id in *id. */
int dgx_server_submit(dgx_server *sv, const dgx_request *rq, long *id);
/* Admit what fits, run `blocks` blocks of steps over all running shots, retire finished requests.
 * Returns the number of requests completed by this call (>= 0) or a DGX_E code. */
int dgx_server_advance(dgx_server *sv, int blocks);
/* State of a request and its admission and completion times (seconds on the caller's clock passed in
 * dgx_server_now / the arrival field; completion is stamped with dgx_server_now()). */
int dgx_server_status(const dgx_server *sv, long id, int *state, double *t_admit, double *t_done);
/* Traces of a finished request: rec_x, rec_z [nshots][nsteps][nrec]. */
int dgx_server_result(dgx_server *sv, long id, double *rec_x, double *rec_z);
/* Forget a finished request's host-side results. */
int dgx_server_release(dgx_server *sv, long id);
int dgx_server_get_stats(const dgx_server *sv, dgx_server_stats *st);
/* Monotonic clock (seconds) the server stamps completions with. */
double dgx_server_now(void);

#ifdef __cplusplus
}
#endif
#endif
