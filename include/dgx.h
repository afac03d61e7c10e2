/*
 * dgx.h -- C-ABI of the MI355X time-stepping library (libdgx.so) for the 2-D nodal P2
 * velocity-stress discontinuous Galerkin solver.
 *
 * The library replaces the per-step calls of the reference's time loop
 * (DG_leap_Frog_OpenMP.c:405-425, prototypes OpenMP_Function.c:5-20) with device-resident HIP
 * kernels.  The caller keeps the reference's setup and file formats; it hands the mesh, materials,
 * sponge factors, source and receiver descriptors over once (dgx_create), then advances whole steps
 * (dgx_step) or calls the split entry points that mirror the reference's five step functions one to
 * one.  Plain C types only; every call returns 0 or a negative DGX_E* code, with a message from
 * dgx_last_error().  Caller-owned memory is never retained after a call returns; the solver handle
 * owns all device memory.  One host thread per handle; calls on a handle are stream-ordered and
 * only dgx_get_*, dgx_snapshot_wait and dgx_synchronize block.
 */
#ifndef DGX_H
#define DGX_H

#ifdef __cplusplus
extern "C" {
#endif

#define DGX_OK 0
#define DGX_EARG (-1)   /* invalid argument */
#define DGX_EMESH (-2)  /* inconsistent mesh (orientation, neighbours, sizes) */
#define DGX_EHIP (-3)   /* HIP runtime error or no usable GPU */
#define DGX_ENOMEM (-4) /* host or device allocation failed */
#define DGX_EIO (-5)    /* file could not be read or written */
#define DGX_ESTATE (-6) /* call not valid in the solver's current state */

/* DGX_FP32_COMPENSATED: FP32 arithmetic and FP32 storage with a second FP32 word per value that carries the
 * rounding error of every update (TwoSum accumulation), so that accumulated rounding does not grow with the
 * number of steps; snapshots and traces return the sum of the two words.  Runs on the blocked kernel only
 * (element layout) and moves twice the field bytes. */
/* DGX_FP64_REFERENCE: FP64 fields advanced with the reference's own arithmetic: its per-element Dx, Dz, Q, A, B
 * (Q, A, B formed on the host as its setup forms them, dgx_reference_operators; Dx, Dz in the kernels with its
 * operations), every sum in its order, no fused multiply-adds, sponge and sources applied before the velocity
 * update and receivers recorded after the stress update, so that fields and traces equal the reference's bit for
 * bit.  Element layout, two kernels per step (step_kernel AUTO or HALF): 2.4x the time per step of DGX_FP64
 * (640,000 elements 0.171 against 0.072 ms, 2M 0.532 against 0.208 ms). */
enum dgx_precision { DGX_FP64 = 0, DGX_FP32 = 1, DGX_FP32_COMPENSATED = 2, DGX_FP64_REFERENCE = 3 };
enum dgx_source_kind { DGX_SRC_FORCE_Z = 0, DGX_SRC_EXPLOSIVE = 1 };
/* Field layout on the device.  ELEMENT: 64-element blocks, one thread per element and shot (the
 * single-shot path).  SHOT_INNER: shots innermost, lanes over shots, for many shots on one mesh.
 * AUTO: ELEMENT (with the blocked kernel, faster at every shot count measured, in both precisions) unless a
 * shot tile or workgroup shape is given for 8 or more shots, which selects SHOT_INNER. */
enum dgx_layout { DGX_LAYOUT_AUTO = 0, DGX_LAYOUT_ELEMENT = 1, DGX_LAYOUT_SHOT_INNER = 2 };
/* How dgx_step runs a step.  HALF: two kernels per step (velocity, then stress), each updating the
 * fields in place.  FULL (element layout only): one kernel per step that reads the fields and writes the
 * next state to a second copy; workgroups recompute the velocities of the elements bordering their tile
 * instead of waiting for them, so each field is read and written once per step.  It gives the same
 * results; on MI355X it is slower than HALF (the kernels wait on memory latency, not bandwidth, and the
 * tile's two phases lengthen that wait).  BLOCKED (element layout): one kernel runs block_steps whole steps
 * per launch; each workgroup holds a tile of elements and the rings of elements within 2*block_steps - 1
 * face hops around it in registers and LDS, recomputes the rings instead of exchanging them, and reads and
 * writes each field once per launch; it keeps a second copy of the fields.  AUTO: BLOCKED in the element
 * layout, in every precision and at any shot count (measured faster than HALF and than the shot-innermost
 * kernels on every mesh and shot count tried); HALF in the shot-innermost layout, and HALF if no tile fits an
 * element and its rings (which only extreme vertex valences cause).  The automatic layout is the element
 * layout, unless a shot tile or workgroup shape is given for 8 or more shots. */
enum dgx_step_kernel { DGX_STEP_AUTO = 0, DGX_STEP_HALF = 1, DGX_STEP_FULL = 2, DGX_STEP_BLOCKED = 3 };

/* Reference-element operators as read from input_files/precomputed_matrix/ (row-major), the arrays
 * DG_leap_Frog_OpenMP.c:52-81 reads.  Vb/inverVb are read by the reference but never used. */
typedef struct {
    double V[36], inverV[36], Vr[36], Vs[36];
    double S1[18], S2[18], S3[18]; /* per-face lift, 6x3 each */
    double FeketeNew[12];          /* node (r,s), 6x2 */
} dgx_operators;

/* One point source.  fun is the per-node source vector ((M^-1 phi(x_s))_i / J, which is the
 * reference's fun at DG_leap_Frog_OpenMP.c:220-233 when x_s is a vertex).  Each step adds
 * fun*w[step] to Vz (force, OpenMP_Function.c:49-58) or to Txx and Tzz (explosive). */
typedef struct {
    long elem;    /* 1-based element in the caller's numbering */
    int kind;     /* enum dgx_source_kind */
    int shot;     /* 0-based shot this source belongs to */
    long wavelet; /* row of dgx_problem.wavelet to use */
    double fun[6];
} dgx_source;

/* One receiver: trace = sum_i weight[i] * V[elem][i] after each step (OpenMP_Function.c:61-72 is the
 * unit-weight case). */
typedef struct {
    long elem; /* 1-based element */
    double weight[6];
} dgx_receiver;

typedef struct {
    long tri;              /* Tri */
    long point_num;        /* Point_Num */
    const double *point;   /* [point_num*2] x,z */
    const long *triangle;  /* [tri*3] 1-based vertex ids, counter-clockwise */
    const long *neighbor;  /* [tri*3] Triangle .neigh columns (element opposite v0,v1,v2), -1 = none */
    const double *lambda;  /* [tri] */
    const double *miu;     /* [tri] */
    const double *den;     /* [tri] */
    const double *absorb_fun; /* [tri*6] sponge factors (Absorbing_Fun, fun_DG.c:906) or NULL */
    double dt;
    int precision;         /* enum dgx_precision: arithmetic and storage type of the fields */
    long nshots;           /* S >= 1 independent wavefields on this mesh */
    long nsteps;           /* steps of the run: wavelet samples and trace capacity; dgx_step refuses to go
                            * past it (DGX_ESTATE) */
    long nwavelets;
    const double *wavelet; /* [nwavelets*nsteps] FP64 source samples */
    long nsrc;
    const dgx_source *src;
    long nrec;
    const dgx_receiver *rec; /* one receiver spread shared by all shots */
    int layout;            /* enum dgx_layout */
    int shot_tile;         /* shots per lane group in the shot-innermost layout; 0 = autotune at create */
    int step_kernel;       /* enum dgx_step_kernel */
    int wg_waves;          /* wavefronts per workgroup of the shot-innermost kernels: 2, 4 or 8 with shot
                            * tiles 16-64, 4 with tiles 4 and 8.  0: autotuned with the shot tile when
                            * shot_tile is 0, else 4.  Given with shot_tile 0, only the tile is tuned. */
    int block_steps;       /* steps per launch of the blocked kernel (DGX_STEP_BLOCKED, or AUTO when it picks the
                            * blocked kernel): 1, 2 or 3; 0 = 2, and 3 in FP32 up to 300,000 elements (compensated
                            * FP32: 50,000) or where the two FP32 field copies of all shots exceed 256 MB */
} dgx_problem;

typedef struct dgx_solver dgx_solver;

/* Upload the problem and build the device-side mesh (replaces the arrays the reference builds at
 * DG_leap_Frog_OpenMP.c:190-367 for the kernels).  Fields start at zero
 * (DG_leap_Frog_OpenMP.c:399-403). */
int dgx_create(const dgx_problem *p, const dgx_operators *ops, dgx_solver **out);
int dgx_destroy(dgx_solver *s);

/* Advance n whole steps: absorb -> source -> velocity -> stress -> record, one step being the body of
 * DG_leap_Frog_OpenMP.c:405-425.  One or two kernels per step (dgx_problem.step_kernel). */
int dgx_step(dgx_solver *s, long n);

/* Split entry points, one per reference step function (same order, same semantics).  They are the
 * unfused form of dgx_step, for parity tests and for callers that interleave their own work. */
int dgx_absorb_boundary(dgx_solver *s);         /* Absorb_Boundary, OpenMP_Function.c:25-46 */
int dgx_source_kernel(dgx_solver *s, double w); /* source_kernel, OpenMP_Function.c:49-58 (all sources) */
int dgx_lu_velocity(dgx_solver *s);             /* LU_Velocity, OpenMP_Function.c:75-185 */
int dgx_lu_stress_elastic(dgx_solver *s);       /* LU_Stress_Elastic, OpenMP_Function.c:188-306 */
int dgx_record(dgx_solver *s, long t);          /* Record, OpenMP_Function.c:61-72 (trace slot t) */

/* Fields of one shot in the caller's element order, [tri*6] each (any pointer may be NULL): the host
 * arrays Save_Data writes at the snapshot cadence (DG_leap_Frog_OpenMP.c:433-444).  Blocks until the
 * steps enqueued so far are done. */
int dgx_get_fields(dgx_solver *s, long shot, double *vx, double *vz, double *txx, double *tzz, double *txz);
/* The same as two calls, so that a snapshot leaves the device while the next steps run.  begin takes the
 * fields in `mask` (bit c: Vx, Vz, Txx, Tzz, Txz for c = 0..4) of one shot at the current step, reorders
 * them on the device and starts their download into pinned memory without waiting; dgx_step may be called
 * right after it.  wait blocks until that download is done and copies the fields into the caller's arrays
 * (pointers of fields outside the mask are ignored).  One snapshot may be in flight per solver: begin
 * fails (DGX_ESTATE) until the previous one has been collected, and wait fails (DGX_ESTATE) with none.
 * (The reference writes snapshots synchronously inside its time loop, DG_leap_Frog_OpenMP.c:433-444.) */
int dgx_snapshot_begin(dgx_solver *s, long shot, int mask);
int dgx_snapshot_wait(dgx_solver *s, double *vx, double *vz, double *txx, double *tzz, double *txz);
int dgx_set_fields(dgx_solver *s, long shot, const double *vx, const double *vz, const double *txx,
                   const double *tzz, const double *txz);
/* Traces of one shot, [nsteps][nrec] like record_x/record_z (DG_leap_Frog_OpenMP.c:166-171). */
int dgx_get_traces(dgx_solver *s, long shot, double *rec_x, double *rec_z);
int dgx_synchronize(dgx_solver *s);
long dgx_current_step(const dgx_solver *s);
/* Average device time of the two fused kernels over n steps, measured with HIP events on the solver's
 * stream (the steps are real steps: the state advances). */
int dgx_kernel_times(dgx_solver *s, long n, double *ms_velocity, double *ms_stress);
/* Average duration of one fused kernel (which: 0 velocity, 1 stress, 2 full step, 3 blocked: one launch of
 * block_steps steps) over n back-to-back launches between two HIP events on the solver's stream.
 * Invalidates the fields (the launches repeat a step).  which = 2 and 3 need the full-step and blocked
 * kernels (element layout). */
int dgx_kernel_bench(dgx_solver *s, int which, long n, double *ms);
/* Streaming copy of `bytes` bytes (read + write) repeated `reps` times on the default stream: the
 * achievable-bandwidth reference and the calibration point of the HBM byte counters.  variant 0:
 * grid-stride loop; variant 1: 32 KiB per workgroup, non-temporal loads and stores; variant 2: read-only
 * stream of `bytes` (32 KiB summed per workgroup). */
int dgx_copy_probe(long bytes, long reps, int variant, double *ms);
/* Layout and shot tile in use (after autotuning), and the autotuner's timings: for each candidate
 * tile k, tiles[k] and ms_per_step[k]; returns the number of candidates (0 if none were timed). */
int dgx_layout_info(const dgx_solver *s, int *layout, int *shot_tile, int *tiles, double *ms_per_step, int cap);
/* The tile shape in use: shot tile and wavefronts per workgroup (elements per workgroup =
 * wg_waves * 64 / shot_tile), and the autotuner's table: candidate k is (tiles[k], waves[k]) at
 * ms_per_step[k].  Returns the number of candidates timed.  No counterpart in the reference, which has
 * one shot per run. */
int dgx_tile_info(const dgx_solver *s, int *shot_tile, int *wg_waves, int *tiles, int *waves, double *ms_per_step,
                  int cap);
/* Tiling of the blocked kernel (dgx_problem.step_kernel = DGX_STEP_BLOCKED): steps per launch, tiles,
 * elements owned by the tiles (the mesh) and lanes (threads holding an element, own or recomputed); lanes/own -
 * 1 is the fraction of the work that is recomputation.  All zero for the other step kernels.  Diagnostic with
 * no counterpart in the reference. */
int dgx_block_info(const dgx_solver *s, int *steps_per_launch, long *ntile, long *own, long *lanes);
/* 1 when the element-layout velocity kernels of this solver run with the P2 Fekete operators of the
 * shipped operator files compiled in (its operators equal them bit for bit and DGX_GENERIC_OPERATORS=1 is not
 * set), 0 when they read the operators from their kernel argument or the layout is shot-innermost.
 * Diagnostic with no counterpart in the reference, which reads its operators from files at start-up
 * (DG_leap_Frog_OpenMP.c:52-81); the two variants agree to rounding. */
int dgx_compiled_operators(const dgx_solver *s);
/* dgx_get_fields without allocating: the same fields, bit for bit, reordered on the device into a static
 * staging array of DGX_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it runs on (looked up at
 * dgx_create), then copied into the caller's arrays; no device or host memory is allocated, on any call.  For
 * callers that cannot allocate (tight serving loops, code that forbids device allocation); dgx_get_fields
 * allocates its snapshot buffers on first use.  The requested fields must fit the array: (fields requested) *
 * tri * 6 * 8 bytes <= DGX_NOALLOC_SCRATCH_BYTES, else DGX_EARG with nothing done (one field of a mesh of up to
 * 699,050 elements, all five up to 139,810).  The array is shared by every call on one GPU and keeps the last
 * call's contents: calls that overlap on one GPU (from several threads, or solvers on several streams) race;
 * calls made one after another are safe, as each waits for its copies before it returns.  Blocks until the
 * steps enqueued so far are done.  No counterpart in the reference. */
#define DGX_NOALLOC_SCRATCH_BYTES (32L << 20)
int dgx_get_fields_noalloc(dgx_solver *s, long shot, double *vx, double *vz, double *txx, double *tzz,
                           double *txz);
const char *dgx_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
