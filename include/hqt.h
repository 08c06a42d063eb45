/*
 * hqt.h -- C ABI of the MI355X-native AdS4 holographic-superfluid stepper.
 *
 * The reference (Eureka4399/AdS4-holographic-quantum-turbulence) is pure MATLAB; its hot
 * path has no function API: the right-hand sides are anonymous closures and the RK4
 * step is an inline loop body.  This header is the drop-in boundary a MATLAB MEX
 * gateway (csrc/hqt_mex.c) or any FFI binds instead of those closures.  Each entry
 * point below names the reference code it replaces.
 *
 * Conventions (docs/DATA_DICTIONARY.md:7-8 of the reference):
 *   - every field is a MATLAB column-major array [nx, ny, nz], linear index
 *     i + nx*(j + ny*k), x fastest; k = nz-1 is the AdS boundary z=0, k = 0 the horizon;
 *   - the complex scalar Psi (Z1) is interleaved (re, im) pairs (MATLAB R2018a layout);
 *   - Z2 = A_t, Z3 = A_x, Z4 = A_y are real; rhoi is [nx, ny, nz] (replicated along z
 *     by the reference, tevol_driven_cluster.m:98), stored in full here.
 * All host pointers are borrowed for the duration of the call.  Every function returns
 * an int status (HQT_OK = 0); on failure hqt_last_error() returns a thread-local
 * message.  No C++ exception crosses this boundary.
 */
#ifndef HQT_H
#define HQT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HQT_OK          0
#define HQT_ERR_ARG     1   /* bad shape / argument */
#define HQT_ERR_NAN     2   /* NaN in the state: the reference's guard, tevol_driven_cluster.m:182-185 */
#define HQT_ERR_HIP     3   /* HIP runtime error */
#define HQT_ERR_OOM     4   /* device allocation failed */
#define HQT_ERR_STATE   5   /* call out of order (e.g. step before set_state) */
#define HQT_ERR_NOCONV  6   /* Newton did not converge within it_max */
#define HQT_ERR_TIMEOUT 7   /* hqt_server_wait: the request is not done yet */

/* hqt_params.flags */
#define HQT_NO_GRAPH 1      /* launch every kernel of a step directly instead of replaying the step's graph */
#define HQT_LINES16  2      /* x/y line passes with 16 lines per work-group (no tuning) */
#define HQT_LINES32  4      /* x/y line passes with 32 lines per work-group where LDS allows (no tuning) */
#define HQT_RADIAL_W2 8     /* radial stage kernel compiled for 2 waves per SIMD (the default) */
#define HQT_RADIAL_W3 16    /* radial stage kernel compiled for 3 waves per SIMD (spills from nz = 13) */
#define HQT_DENSE_FILTER 32 /* filter's imaginary part as a dense Toeplitz product instead of its rank-4 form
                               (the default; ff.m's 2/3 mask leaves at most two unpaired modes) */
#define HQT_ORDER_XCD  128  /* with HQT_LINES16/32: folded line passes in the XCD-aware block order (default with
                               a forced line count: launch order; results are the same bitwise either way) */
#define HQT_DENSE_LINES 64  /* x/y operators and filter as dense Toeplitz products (func_Lx's own form)
                               instead of their even/odd folded form (the default: half the
                               multiply-adds; results agree to rounding) */
#define HQT_FOLDED_STAGES 256 /* the RK stages' x/y derivatives (and the filter) as even/odd folded matrix
                               products even where the line length has a factored-DFT form
                               (N = 32, 64, 96, 128, 192, 256, 301: each line's DFT once, the operators'
                               eigenvalues, one inverse per output; results agree to rounding) */
#define HQT_FACTORED_STAGES 512 /* the factored-DFT form wherever the line length has one (by default the
                               mixed form takes it at every such N, FP64 passes from N = 128) */
#define HQT_FOLDED_FILTER 1024 /* keep the filter passes as folded matrix products where the stages take the
                               factored-DFT form (by default the filter follows them, where the mask is
                               one range whose unpaired modes fall on complex columns) */
#define HQT_FP32_FILTER 2048 /* mixed form: the filter's passes in FP32 arithmetic with FP32 intermediates
                               (by default they compute in FP64, reading and writing the FP32 state) */

#define HQT_FP64   0        /* plain form: FP64 end to end */
#define HQT_MIXED  1        /* x/y operators and stage arithmetic FP32 (state stored FP32); z operators,
                               A_t, the constraint solve, boundary rows and the filter's arithmetic FP64 */

typedef struct hqt_ctx hqt_ctx;

/* Run parameters.  Replaces the hard-coded workspace of tevol_driven_cluster.m:11,23-24,79,86,110. */
typedef struct hqt_params {
    int32_t nx, ny, nz;     /* grid (nx=ny=301, nz=20 at :11) */
    double  Lx, Ly;         /* periodic box sides L0, L (:23-24,85) */
    double  mu;             /* chemical potential: A_t(z=0) (:79,131,133) */
    double  dt;             /* RK4 step h (:110) */
    int32_t winding;        /* 1: winding removal after each step (:136-139); 0: off (tevol_lattice.m:168-170) */
    int32_t filter;         /* 1: Fourier filter after each step (:142); 0: off */
    int32_t precision;      /* HQT_FP64 or HQT_MIXED */
    int32_t flags;          /* HQT_NO_GRAPH, HQT_LINES16/32 (default: lines per work-group measured per
                               pass at hqt_ctx_create and kept for the process; results are the same
                               bitwise), HQT_RADIAL_W2/W3 (radial kernel occupancy; results may differ
                               in the last bits between the two) */
    const double* ax_b;     /* [nx*ny] A_x(z=0) restored each step, or NULL for ax_b_const (:133: 0) */
    const double* ay_b;     /* [nx*ny] A_y(z=0) restored each step, or NULL for ay_b_const (:133: -vy) */
    double  ax_b_const, ay_b_const;
} hqt_params;

/* ---- context life cycle: tevol_driven_cluster.m:11-77 operator set-up ------------------ */
int hqt_ctx_create(const hqt_params* p, hqt_ctx** out);
int hqt_ctx_destroy(hqt_ctx* ctx);

/* Upload / download the state (Z1 interleaved complex, Z2, Z3, Z4, rhoi; each [nx,ny,nz]).
 * Replaces the gpuArray residency of :81-102 (upload) and the checkpoint save of :178-181
 * (download).  The state stays device-resident between calls.  hqt_get_state waits for
 * all queued work and returns HQT_ERR_NAN if the NaN guard tripped. */
int hqt_set_state(hqt_ctx* ctx, const double* Z1, const double* Z2, const double* Z3,
                  const double* Z4, const double* rhoi);
int hqt_get_state(hqt_ctx* ctx, double* Z1, double* Z2, double* Z3, double* Z4, double* rhoi);

/* Queue nsteps RK4 steps: the loop body tevol_driven_cluster.m:120-142 (+ the NaN guard of
 * :182-185) iterated.  Returns after queueing; hqt_sync waits.  A step's 20 kernel launches are
 * captured once into a hipGraph and replayed per step (unless HQT_NO_GRAPH, or while
 * per-kernel profiling is on). */
int hqt_step_n(hqt_ctx* ctx, int32_t nsteps);
/* Wait for queued work.  HQT_ERR_NAN if the guard tripped; *nan_step (nullable) is the
 * 1-based step (since hqt_set_state) whose state first held a NaN, else 0. */
int hqt_sync(hqt_ctx* ctx, int32_t* nan_step);
/* Winding number removed by the most recent step (:137); NaN when O had exact zeros. */
int hqt_last_winding(hqt_ctx* ctx, double* w);

/* Right-hand sides at a given state: the closures F1, F2, F3, F4 of
 * tevol_driven_cluster.m:72-76 evaluated over the full volume.  F1 interleaved complex.
 * Inputs and outputs are host arrays [nx,ny,nz]; overwrites the context's stage buffers. */
int hqt_rhs(hqt_ctx* ctx, const double* Z1, const double* Z2, const double* Z3, const double* Z4,
            double* F1, double* F2, double* F3, double* F4);

/* Boundary order parameter O = (d_z Psi)(z=0) of the current state, [nx,ny] interleaved
 * complex (tevol_driven_cluster.m:136,148; scripts/analysis/movie_show.m:91).  Computed on
 * the GPU (Lz(nz,:) applied to every column); only O is downloaded. */
int hqt_order_parameter(hqt_ctx* ctx, double* O);

/* Vortex census of the current state on the GPU: [n,np,nm] = countN(O)
 * (src/diagnostics/countN.m:7-21: periodic plaquette winding w, counted as +1 when
 * |w-1| <= 1e-4 and as -1 when |w+1| <= 1e-4).  counts[3] = {n, np, nm}.  charge (optional,
 * [nx,ny] int8, column-major) = +1 / -1 / 0 for the plaquette with corners (i,j), (i+1,j),
 * (i+1,j+1), (i,j+1), periodic (the +-1 entries of find_vortices.m:103-131's charge matrix).
 * Oabs (optional, [2]) = {max|O|, mean|O|}.  Only counts, charges and per-block partial sums
 * are downloaded. */
int hqt_vortex_census(hqt_ctx* ctx, int32_t* counts, int8_t* charge, double* Oabs);

/* ---- initial state on the device: tevol_driven_cluster.m:81-102, tevol_lattice.m:101-135 -- */
/* The stationary profile (Z1, Z2, Vy1, rhoi on the chebdif(nz0) nodes, as the reference's
 * data/reference .mat files hold it) is splined to the run's nz nodes (interp1 'spline', :84),
 * gauge-rotated by theta = -Lzr(A_t/f) (:90-93) and spread over the plane with the pattern
 *   amp(x,y)   = (1 + amp_noise g1) (1 - sum_v exp(-|r - r_v|^2/sigma^2))    (sigma <= 0: no dips)
 *   phase(x,y) = phase_noise g2 + sum_v q_v atan2(y - vy_v, x - vx_v)
 * (:101 is amp_noise = 1e-2; the lattice of tevol_lattice.m:104-135 is sigma = 3 with q = +-6).
 * g1, g2: standard normals from Philox4x32-10 keyed by seed and counted by the column index
 * i + nx*j (Box-Muller), independent of the device and launch shape.  A_t = Z2, A_x = 0,
 * A_y = -Vy1, rhoi = rhoi(z=0) replicated (:94-98).  Replaces hqt_set_state; the fields are
 * written in place on the device. */
typedef struct hqt_init {
    int32_t nz0;
    const double *Z1, *Z2, *Vy1, *rhoi;  /* [nz0] stationary profile */
    double amp_noise, phase_noise;
    uint64_t seed;
    int32_t nv;                          /* windings / dips */
    const double *vx, *vy, *q;           /* [nv] positions (box units) and charges */
    double sigma;
} hqt_init;
int hqt_init_state(hqt_ctx* ctx, const hqt_init* in);
/* yi = interp1(X, Y, Xi, 'spline'): not-a-knot cubic spline, X strictly monotone, n >= 4. */
int hqt_spline(int32_t n, const double* X, const double* Y, int32_t m, const double* Xi, double* out);

/* ---- checkpoints: tevol_driven_cluster.m:178-181, docs/DATA_DICTIONARY.md "Checkpoint files" -- */
/* File name of the checkpoint at time t: "check_point_t=" num2str(t) ".mat" (num2str of a
 * non-integer scalar prints %.<d>g with d = max(floor(log10|t|), 0) + 5). */
int hqt_checkpoint_name(double t, char* buf, int32_t len);
/* save(path, 'Z1','Z2','Z3','Z4','rhoi') of host arrays: a MAT level 5 file (uncompressed,
 * as save -v6) with Z1 complex and all five [nx,ny,nz] double. */
int hqt_checkpoint_write_host(const char* path, int32_t nx, int32_t ny, int32_t nz, const double* Z1,
                              const double* Z2, const double* Z3, const double* Z4, const double* rhoi);
/* The same for the context's current state: synchronous, or asynchronous (the state is
 * snapshotted on the device and written by a host thread while stepping goes on; one write in
 * flight per context, hqt_checkpoint_wait joins it and returns its status). */
int hqt_checkpoint_save(hqt_ctx* ctx, const char* path);
int hqt_checkpoint_save_async(hqt_ctx* ctx, const char* path);
int hqt_checkpoint_wait(hqt_ctx* ctx);

/* ---- batched ensemble ------------------------------------------------------------------ */
/* Many independent evolutions stepped together: the contexts may differ in nx, ny, nz, mu,
 * boundary drive, winding/filter switches and precision.  Every kernel of the step runs once
 * per kernel class (precision and padded line length for the x/y passes, nz for the radial
 * kernels) over all members of that class, on four lanes (streams) of about equal point
 * counts that meet only at the start and end of a replay; the steps are captured hipGraphs of
 * one step and of eight steps (unless flags has HQT_NO_GRAPH; the eight-step graph is captured
 * by the first hqt_ensemble_step_n of eight or more steps).  Each member's result equals
 * hqt_step_n on that context
 * alone.  A member that hits NaN stops (its later kernels exit at once) while the others go
 * on.  Members keep their state and accessors; while in an ensemble, their calls first wait
 * for the ensemble's queued steps.  Serves the batched evolutions of the reference's
 * parameter scans (scripts/simulation/tevol_driven_cluster.m run per (mu, kk) point); no
 * single reference function corresponds. */
typedef struct hqt_ensemble hqt_ensemble;
int hqt_ensemble_create(hqt_ctx* const* members, int32_t n, int32_t flags, hqt_ensemble** out);
/* Queue nsteps steps of every member (returns after queueing). */
int hqt_ensemble_step_n(hqt_ensemble* ens, int32_t nsteps);
/* Wait; nan_steps (nullable, [n]) = per member NaN step as hqt_sync.  HQT_ERR_NAN if any. */
int hqt_ensemble_sync(hqt_ensemble* ens, int32_t* nan_steps);
/* Number of kernel launches per step (the batched plan length). */
int hqt_ensemble_launches(hqt_ensemble* ens, int32_t* n);
/* Release the ensemble (members stay valid and are released from it). */
int hqt_ensemble_destroy(hqt_ensemble* ens);

/* ---- serving loop ------------------------------------------------------------------------ */
/* A host scheduler thread serving two request types with continuous batching:
 *   - evolution segments: a context's parameters, initial state and step count
 *     (tevol_driven_cluster.m:117-187 for nsteps steps, then the checkpoint download :178-181);
 *   - stationary Newton solves (src/solvers/NewtonRaphson_vy.m:1-105, it_max 1000, tol 1e-6).
 * Evolutions are admitted in arrival order while the sum of their device footprints
 * (hqt_ctx_footprint) stays within hbm_budget bytes; admitted ones stay device-resident and
 * advance together as one ensemble, `quantum` steps per scheduling round, and leave (state
 * downloaded, memory freed) when done or when their NaN guard trips.  Newton requests pending
 * at a round are solved as batches on a second stream while the round's steps run.  Every
 * result equals what the request gives alone (hqt_step_n / hqt_newton_vy_batch).  Requests
 * copy their inputs at submission.  No reference counterpart: the reference runs one script
 * per parameter point. */
typedef struct hqt_server hqt_server;
int hqt_server_create(int64_t hbm_budget, int32_t quantum, hqt_server** out);
/* Stops the scheduler; requests not finished end with HQT_ERR_STATE. */
int hqt_server_destroy(hqt_server* srv);
int hqt_server_submit_evolution(hqt_server* srv, const hqt_params* p, const double* Z1, const double* Z2,
                                const double* Z3, const double* Z4, const double* rhoi, int32_t nsteps,
                                int64_t* id);
int hqt_server_submit_newton(hqt_server* srv, int32_t nz, double zh, double small, double D, const double* Z1,
                             const double* Z2, const double* vy1, double vy, int64_t* id);
/* A Newton request solved by mu-continuation from a seed, as hqt_newton_vy_continue_batch with
 * mu_seed and dmu (a NaN mu_seed: as hqt_server_submit_newton).  The result's it and err are those
 * of the last solve. */
int hqt_server_submit_newton_continued(hqt_server* srv, int32_t nz, double zh, double small, double D,
                                       const double* Z1, const double* Z2, const double* vy1, double vy,
                                       double mu_seed, double dmu, int64_t* id);
/* Wait until request id is done (timeout_s < 0: no limit; HQT_ERR_TIMEOUT otherwise) and
 * return its status; *latency_s = completion - submission on the host clock. */
int hqt_server_wait(hqt_server* srv, int64_t id, double timeout_s, double* latency_s);
/* Results of a done request (then released).  Evolution: the state after its steps (or at
 * the NaN stop; *nan_step as hqt_sync).  Newton: z1, z2, VY1 ([nz]), it, err. */
int hqt_server_evolution_result(hqt_server* srv, int64_t id, double* Z1, double* Z2, double* Z3, double* Z4,
                                double* rhoi, int32_t* nan_step);
int hqt_server_newton_result(hqt_server* srv, int64_t id, double* z1, double* z2, double* VY1, int32_t* it,
                             double* err);
/* Latency percentiles (nearest rank, seconds) over completed requests of kind 0 (evolution),
 * 1 (Newton) or 2 (both). */
int hqt_server_stats(hqt_server* srv, int32_t kind, double* p50, double* p99, int64_t* completed);
/* out[HQT_SERVER_NCOUNTERS]: evolutions admitted, evolutions done, Newton done, ensemble
 * (re)builds, Newton batches, peak resident device bytes, peak resident evolutions, ensemble steps
 * run, Newton solves run (each step of a continuation counted), Newton requests continued. */
#define HQT_SERVER_NCOUNTERS 10
int hqt_server_counters(hqt_server* srv, int64_t* out);
/* Device bytes a context with parameters p allocates, and those a context holds. */
int hqt_ctx_footprint(const hqt_params* p, int64_t* bytes);
int hqt_ctx_device_bytes(hqt_ctx* ctx, int64_t* bytes);

/* ---- stand-alone operators ---------------------------------------------------------- */
/* w = ff(x, fx|fy): src/numerics/ff.m:7-11 with the mask of tevol_driven_cluster.m:27-31.
 * x is [nx, ny, nslices], real (x_im == NULL) or complex (x_im = imaginary part);
 * fx [nx], fy [ny] are 0/1 bytes; the 2-D mask is fx(i) | fy(j).  Computed on the GPU. */
int hqt_ff(int32_t nx, int32_t ny, int32_t nslices, const double* x_re, const double* x_im,
           const uint8_t* fx, const uint8_t* fy, double* w);
/* w = ff(x, y) for any logical mask y: src/numerics/ff.m:7-11 (b = fft2(x); b(y) = 0; w = real(ifft2(b))).
 * x is [nx, ny, nslices], real (x_im == NULL) or complex; mask is [nx, ny] 0/1 bytes in the field's order
 * (entry kx + nx ky masks the Fourier mode (kx, ky)), the same for every slice (mask_slices = 1) or one per slice
 * (mask_slices = nslices).  Computed on the GPU by direct DFTs in FP64, O(nx ny (nx + ny)) per slice; hqt_ff is
 * the fast form for the separable masks fx | fy the stepper uses. */
int hqt_ff_mask(int32_t nx, int32_t ny, int32_t nslices, const double* x_re, const double* x_im,
                const uint8_t* mask, int32_t mask_slices, double* w);

/* [x, DM] = fourdif(N, m): third_party/dmsuite/fourdif.m:30-71.  DM column-major N*N. */
int hqt_fourdif(int32_t N, int32_t m, double* x, double* DM);
/* [x, DM] = chebdif(N, M): third_party/dmsuite/chebdif.m:25-53.  DM column-major N*N*M. */
int hqt_chebdif(int32_t N, int32_t M, double* x, double* DM);

/* [z1,z2,VY1,it,error] = NewtonRaphson_vy(nz,zh,small,D,Z1,Z2,vy1,vy):
 * src/solvers/NewtonRaphson_vy.m:1-105 (plots and prints dropped), host FP64.
 * Inputs/outputs are nz-vectors; returns HQT_OK when the residual test (:92) passed,
 * HQT_ERR_NOCONV when it_max (:8) was reached or the update was NaN (:88). */
int hqt_newton_vy(int32_t nz, double zh, double small, double D, const double* Z1,
                  const double* Z2, const double* vy1, double vy, double* z1, double* z2,
                  double* VY1, int32_t* it, double* err);

/* NewtonRaphson_vy over a batch of nb independent problems of one radial size on the GPU,
 * one work-group per problem: the iteration of src/solvers/NewtonRaphson_vy.m:1-105 (as
 * hqt_newton_vy), e.g. a sweep over vy or over initial guesses in one launch.
 * Z1, Z2, vy1 (inputs) and z1, z2, VY1 (outputs) are [nz, nb] column-major (problem b in
 * column b); vy, it, err, conv are [nb].  err is norm(y, inf) of the last update (NaN when it
 * was NaN, :88); conv[b] = 1 when the residual test (:92) passed.  it_max <= 0 and tol <= 0
 * select the reference's 1000 and 1e-6 (:8-9).  3 <= nz <= 64.  Returns HQT_OK when every
 * problem converged, HQT_ERR_NOCONV otherwise (all outputs are written either way). */
int hqt_newton_vy_batch(int32_t nz, double zh, double small, double D, int32_t nb,
                        const double* Z1, const double* Z2, const double* vy1, const double* vy,
                        int32_t it_max, double tol, double* z1, double* z2, double* VY1,
                        int32_t* it, double* err, uint8_t* conv);

/* mu-continuation form of hqt_newton_vy_batch: the stationary-state notebook's continuation in mu
 * (scripts/legacy/Statinary_state_with_v.m:11-17 steps mu by hand, :50 rescales the seed profile's
 * A_t, :52 solves) run inside the GPU work-group of each problem.  Problem b's target is its guess's
 * own A_t(z=0) = Z2[nz-1 + nz*b]; the guess, rescaled so that A_t(z=0) = mu_seed[b], is solved at
 * the seed's mu, then mu is stepped to the target in steps of at most dmu, each from the last
 * solution rescaled the same way.  A step that does not converge is retried at half the length,
 * down to dmu/64.  Each solve stops at min(it_max, 100) iterations; at most 1024 solves.  it and
 * err are those of the last solve, steps[b] the number of solves; conv[b] = 1 when the target was
 * reached.  A NaN mu_seed[b] solves problem b once from its guess, as hqt_newton_vy_batch does.
 * mu_seed[b] and the target must be finite, nonzero and of one sign, dmu finite and positive
 * (HQT_ERR_ARG otherwise).  Returns as hqt_newton_vy_batch. */
int hqt_newton_vy_continue_batch(int32_t nz, double zh, double small, double D, int32_t nb,
                                 const double* Z1, const double* Z2, const double* vy1, const double* vy,
                                 const double* mu_seed, double dmu, int32_t it_max, double tol, double* z1,
                                 double* z2, double* VY1, int32_t* it, double* err, uint8_t* conv,
                                 int32_t* steps);

/* hqt_newton_vy_batch without any device allocation, for callers that cannot call hipMalloc
 * (code under HIP graph capture, tight serving loops, harnesses that forbid allocation).  No
 * counterpart in the reference.  Same arguments, results and status as hqt_newton_vy_batch (the same
 * kernel on the same inputs: bitwise the same outputs); the device workspace (operators, inputs,
 * results, job descriptors and, for nz > 41, the 3nz x 3nz Jacobians) lives in a fixed per-device
 * array of HQT_NEWTON_NOALLOC_BYTES bytes that each GPU the library runs on holds until the process
 * ends.  A batch needs 16 nz^2 + 48 nz nb + 168 nb bytes, plus 72 nz^2 nb for nz > 41 (about 7800
 * problems fit at nz <= 41, 100 at nz = 48, 56 at nz = 64); a larger one is rejected with
 * HQT_ERR_ARG before any GPU work.  Calls on one GPU share its array: the library
 * runs them one at a time (a per-device lock), so calls from several host threads wait for one
 * another.  Runs on the device's null stream and returns when the results are on the host. */
#define HQT_NEWTON_NOALLOC_BYTES (16u << 20)
int hqt_newton_vy_batch_noalloc(int32_t nz, double zh, double small, double D, int32_t nb,
                                const double* Z1, const double* Z2, const double* vy1, const double* vy,
                                int32_t it_max, double tol, double* z1, double* z2, double* VY1,
                                int32_t* it, double* err, uint8_t* conv);

/* Copy one internal device field of the last hqt_rhs / hqt_step_n into out ([nx,ny,nz], FP64
 * values).  Names: "dxR" "dxI" "dxxR" "dxxI" "dxAx" "dxAy" "dxxAy" (x-pass outputs), "dyR" "dyI"
 * "lapR" "lapI" "divA" "curl3" "curl4" (y-pass outputs), "dxAt" "dyAt" "dzAt" (frozen A_t
 * derivatives).  Diagnostic access for per-kernel parity tests; no counterpart in the reference. */
int hqt_debug_field(hqt_ctx* ctx, const char* name, double* out);

/* ---- measurement ---------------------------------------------------------------------- */
/* Enable (1) / disable (0) per-kernel HIP-event timing on the context's stream. */
int hqt_profile(hqt_ctx* ctx, int32_t enable);
/* Per-kernel totals since profiling was enabled: entry idx (0-based) -> name, launches,
 * summed duration in ms.  Returns HQT_ERR_ARG when idx is past the last entry. */
int hqt_kernel_stats(hqt_ctx* ctx, int32_t idx, char* name, int32_t name_len, int64_t* count,
                     double* total_ms);

/* Copy `bytes` between two fresh device buffers with a plain streaming kernel (k_copy_calib),
 * twice, and return the duration of the second copy.  Measurement helper with no counterpart in
 * the reference: its FETCH_SIZE / WRITE_SIZE counters calibrate the HBM traffic figures of the
 * other kernels, and its rate is the streaming bandwidth the HBM-bound kernels are compared with. */
int hqt_calibration_copy(int64_t bytes, double* ms);

/* Apply the Chebyshev derivative twice (u = Lz x, v = Lz u) to ncol columns of nz points (nz = 12 or 20) with
 * the radial matvec of the stage kernel in `form` 0 (FP64 MFMA on the register slot layout) or 1 (vector FMAs
 * fed through LDS, the LDS/wavefront alternative), `reps` timed launches after one warm-up; returns the average
 * duration per launch and the sum of squares of v (the two forms agree to rounding).  Measurement helper with
 * no counterpart in the reference: it compares the two ways of doing the radial contraction. */
int hqt_calibration_radial(int64_t ncol, int32_t nz, int32_t form, int32_t reps, double* ms, double* sumsq);

const char* hqt_last_error(void);
const char* hqt_version(void);

#ifdef __cplusplus
}
#endif
#endif /* HQT_H */
