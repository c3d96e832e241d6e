/* amdflow.h -- C ABI of the MI355X optical-flow library (libamdflow.so).
 *
 * Drop-in boundary for the reference's three per-algorithm entry points
 * (artecs-group/sycl-optic-flow).  Every function returns an of_status
 * (0 = OK) and never throws or exits.  Handles own all device memory; the
 * caller owns the host arrays it passes.  Calls on one handle must come from
 * one host thread at a time.  All *_run/compute calls taking host pointers
 * are synchronous: results are in host memory when they return, as in the
 * reference (src/tvl1/cuda/tvl1.cu:181-182, src/hs/cuda/flow.cu:206-207).
 * The *_device variants take device pointers and a HIP stream (passed as
 * void*; NULL is the HIP null stream) and return once the work is complete
 * on that stream.
 */
#ifndef AMDFLOW_H
#define AMDFLOW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    OF_OK = 0,
    OF_ERR_ARG = 1,          /* invalid size or parameter */
    OF_ERR_SIGMA = 2,        /* Gaussian window larger than an image level
                                (the reference's "GaussianSmooth: sigma too
                                large", mask.cpp:218-221) */
    OF_ERR_ALLOC = 3,        /* device or host allocation failed */
    OF_ERR_HIP = 4,          /* a HIP runtime call failed */
    OF_ERR_STATE = 5,        /* call not valid in the handle's state */
    OF_ERR_UNSUPPORTED = 6,  /* filter size or option not provided */
    OF_ERR_CAPACITY = 7      /* input larger than the fixed scratch of an
                                entry point that does not allocate */
} of_status;

const char *of_status_string(int status);
/* Library build identification (architecture, flags). */
const char *of_build_info(void);

/* ------------------------------------------------------------------------
 * TV-L1 (Zach-Pock-Bischof dual TV-L1, coarse to fine).
 * Replaces class TV_L1, src/tvl1/sequential/tvl1.hpp:10-41.
 * ---------------------------------------------------------------------- */
typedef struct of_tvl1 of_tvl1;

typedef struct {
    float tau;          /* time step                       (tvl1.hpp:12) */
    float lambda;       /* data-term weight                (tvl1.hpp:12) */
    float theta;        /* coupling weight                 (tvl1.hpp:13) */
    int nscales;        /* requested scales (upper bound)  (tvl1.hpp:13) */
    float zfactor;      /* pyramid zoom factor in (0,1)    (tvl1.hpp:13) */
    int warps;          /* warpings per scale              (tvl1.hpp:13) */
    float epsilon;      /* stopping threshold              (tvl1.hpp:14) */
    int max_iterations; /* per-warping cap, MAX_ITERATIONS (tvl1.hpp:6)  */
} of_tvl1_params;

/* Fills the reference defaults: tau 0.25, lambda 0.15, theta 0.3,
 * nscales 100, zfactor 0.5, warps 5, epsilon 0.01, max_iterations 300
 * (src/tvl1/sequential/tvl1.hpp:6,12-14). */
void of_tvl1_default_params(of_tvl1_params *p);

/* Constructor: TV_L1::TV_L1(int width, int height, float tau, ...)
 * (src/tvl1/sequential/tvl1.cpp:27-79).  The number of scales follows the
 * reference rule min(nscales, 1 + ln(hypot(W,H)/16)/ln(1/zfactor))
 * (tvl1.cpp:42-43), clamped to >= 1.  `params` may be NULL for defaults. */
int of_tvl1_create(const of_tvl1_params *params, int width, int height, of_tvl1 **out);
void of_tvl1_destroy(of_tvl1 *h);

/* Number of pyramid scales in use and the size of scale `s` (0 = finest). */
int of_tvl1_nscales(const of_tvl1 *h);
int of_tvl1_scale_size(const of_tvl1 *h, int s, int *nx, int *ny);

/* Pair solve on 8-bit frames: TV_L1::runDualTVL1Multiscale(uint8_t* I0,
 * uint8_t* I1) (src/tvl1/sequential/tvl1.hpp:17, tvl1.cpp:129-175).
 * Host pointers of width*height bytes each.  Flow available through
 * of_tvl1_flow. */
int of_tvl1_run_u8(of_tvl1 *h, const uint8_t *I0, const uint8_t *I1);

/* Same solve on float frames (host pointers, width*height floats each);
 * the u8->float cast of tvl1.cpp:118-123 is skipped. */
int of_tvl1_run_f32(of_tvl1 *h, const float *I0, const float *I1);

/* Streaming solve: TV_L1::runDualTVL1Multiscale(const float* I) of the GPU
 * variants (src/tvl1/cuda/tvl1.cuh:19, tvl1.cu:122-183): I0 is the frame
 * passed on the previous call, I1 the frame passed now.  The first call has
 * no previous frame: it stores I and leaves the flow at zero. */
int of_tvl1_push_f32(of_tvl1 *h, const float *I);

/* getU(): planar [u1 | u2], 2*width*height floats, owned by the handle and
 * valid until the next run or destroy (src/tvl1/sequential/tvl1.hpp:16). */
const float *of_tvl1_flow(of_tvl1 *h);

/* Device-resident variant used by throughput measurements and by callers
 * that keep frames in HBM: dI0/dI1 are device pointers to width*height
 * bytes, dU a device pointer to 2*width*height floats (planar [u1|u2]).
 * `stream` is a hipStream_t (NULL = the HIP null stream).  Returns after the
 * solve has completed on that stream. */
int of_tvl1_run_device_u8(of_tvl1 *h, const uint8_t *dI0, const uint8_t *dI1, float *dU,
                          void *stream);
int of_tvl1_run_device_f32(of_tvl1 *h, const float *dI0, const float *dI1, float *dU,
                           void *stream);

/* Per-(scale, warping) iteration counts of the last solve, finest scale
 * first: counts[s*warps + w].  `cap` is the array length. */
int of_tvl1_iterations(const of_tvl1 *h, int *counts, int cap);

typedef struct {
    int nscales;
    int warps;
    long long pixel_iterations;    /* sum over (scale,warp) of iters*N_s    */
    long long pixel_warps;         /* warps * sum of N_s                     */
    long long pixels_all_scales;   /* sum of N_s                             */
    long long launches;            /* iteration-kernel launches issued       */
    double algorithmic_bytes;      /* 60*PI + 36*PW + 64*sum N_s + 48*N_0    */
    double host_launch_seconds;    /* host time inside iteration launches    */
    double host_wait_seconds;      /* host time waiting for device progress  */
} of_tvl1_stats;
int of_tvl1_get_stats(const of_tvl1 *h, of_tvl1_stats *st);

/* Iteration-kernel launch shape for pyramid scales whose pixel count is at
 * least `min_pixels`: `iters_per_launch` primal-dual iterations run in one
 * launch on tiles of 64 x (4*rows_per_wave) pixels with a halo of
 * iters_per_launch pixels.  Supported: iters_per_launch in {1,2,4,8},
 * rows_per_wave in {2,4,8}; rows_per_wave 0 selects square 16 x 16 tiles,
 * one pixel per lane (iters_per_launch in {1,2,4}; the built-in choice,
 * host-paced, below 200000 pixels).  `persistent` = 0 runs host-paced launches;
 * 1 runs each warping of those scales as one launch of a co-resident grid,
 * chunks of iters_per_launch iterations separated by grid barriers; 2 runs
 * the same single launch with neighbour flags instead of grid barriers (a
 * tile starts a chunk when its 8 neighbours have finished the previous one,
 * and one extra workgroup applies the stopping rule while the tiles go on;
 * see of_set_persist_launch_mode); other values return OF_ERR_ARG.  Rules are
 * matched largest min_pixels first; calling with min_pixels < 0 clears all
 * rules (back to the built-in table). */
int of_tvl1_set_launch_rule(of_tvl1 *h, long long min_pixels, int iters_per_launch,
                            int rows_per_wave, int persistent);

/* Launch mode of the persistent warping kernel, process-wide (no counterpart
 * in the reference, whose solver is sequential).  OF_PERSIST_PLAIN (default):
 * warpings whose launch rule is persistent run as one launch of a grid that
 * fits the device at once; its first grid barrier checks within 2 ms that
 * every workgroup started.  If one did not (another persistent kernel on
 * another stream or process held the CUs, a CU-masked stream), the solve is
 * repeated and the process stays in OF_PERSIST_HOST_PACED: those warpings
 * then run host-paced with the same launch shape, which needs no
 * co-residency.  AMDFLOW_HOST_PACED=1 in the environment starts a process in
 * host-paced mode.  Results are the same bits in both modes. */
enum { OF_PERSIST_PLAIN = 0, OF_PERSIST_HOST_PACED = 1 };
int of_persist_launch_mode(void);
int of_set_persist_launch_mode(int mode);

/* Iteration launches kept queued ahead of the device within a warping
 * (1..64, default 3).  Launches queued past convergence return at once. */
int of_tvl1_set_inflight(of_tvl1 *h, int launches);

/* Autotuner of the launch shapes (no reference counterpart): tries every
 * supported (iters_per_launch, rows_per_wave) in the three launch modes
 * (host-paced, persistent with grid barriers, persistent with neighbour
 * flags) on all scales, times each scale with HIP events over `reps` solves
 * of the device pair dI0/dI1 (width*height bytes each) and installs the
 * fastest variant per scale as launch rules.  `chosen` (optional) receives
 * 3*nscales ints: K, R, persistent for each scale, finest first; `ms_table`
 * (optional) receives OF_TVL1_TUNE_VARIANTS (36) x nscales mean times in ms:
 * rows 0-8 host-paced, 9-17 persistent with grid barriers, shapes in the
 * order (1,2) (2,2) (1,4) (2,4) (4,4) (1,8) (2,8) (4,8) (8,8); rows 18-20
 * square 16 x 16 tiles (rows_per_wave 0) host-paced with 1, 2 and 4
 * iterations per launch, rows 21-23 the same persistent with grid barriers;
 * rows 24-32 the nine shapes persistent with neighbour flags, rows 33-35
 * the square tiles the same; -1 marks a variant the device cannot run.
 * Results are bit-identical for every variant; only the time changes. */
#define OF_TVL1_TUNE_VARIANTS 36
int of_tvl1_autotune(of_tvl1 *h, const uint8_t *dI0, const uint8_t *dI1, int reps, int *chosen,
                     double *ms_table);

/* Per-launch timing of the primal-dual iteration kernel.  When enabled,
 * every iteration launch is bracketed by two HIP events recorded on the
 * stream it runs on; after each solve the elapsed times are accumulated per
 * launch shape together with the algorithmic bytes of the iterations those
 * launches completed (60 bytes per pixel-iteration: read u1,u2,p11,p12,p21,
 * p22,I1wx,I1wy,rho_c, write u1,u2,p11,p12,p21,p22).  Enabling resets the
 * accumulators.  Adds host overhead: use it outside timed regions. */
typedef struct {
    int iters_per_launch;
    int rows_per_wave;
    int persistent;                /* 1, 2: whole warping in one launch      */
    long long launches;            /* launches timed (including no-op ones) */
    long long pixel_iterations;    /* iterations completed x scale pixels    */
    double milliseconds;           /* summed launch durations                */
    double algorithmic_bytes;      /* 60 * pixel_iterations                  */
} of_tvl1_kernel_profile;
int of_tvl1_profile_enable(of_tvl1 *h, int enable);
/* Writes up to `cap` entries, one per launch shape used; *count receives
 * the number of entries. */
int of_tvl1_profile_get(const of_tvl1 *h, of_tvl1_kernel_profile *out, int cap, int *count);

/* Per-scale device time (no reference counterpart): when enabled, HIP
 * events recorded on the solve's stream before each scale's gradient and
 * after its prolongation time every scale of the solves that follow.
 * Enabling (or disabling) resets the sums.  of_tvl1_scale_times writes up
 * to `cap` summed milliseconds, finest scale first, and the number of
 * solves timed. */
int of_tvl1_scale_timing_enable(of_tvl1 *h, int enable);
int of_tvl1_scale_times(const of_tvl1 *h, double *ms, int cap, int *solves);

/* ------------------------------------------------------------------------
 * Batched TV-L1 (no reference counterpart): many frame pairs of mixed sizes
 * solved together.  Results are identical to solving each pair alone.
 *   OF_BATCH_LOCKSTEP (default): one solver handle per pair (pooled by frame
 *     size); every launch covers all pairs of a lane: the preparation step
 *     by step, then the solve scale by scale and warping by warping; each
 *     pair stops on its own rule.  Up to 512 pairs per lane at a time.
 *   OF_BATCH_WORKERS: `streams` host workers, each with its own HIP stream
 *     and one handle per frame size, solve pairs one by one, longest first.
 * The batch's streams do not wait for other streams: the frames must be
 * complete on the device when of_tvl1_batch_run_u8 is called.
 * ---------------------------------------------------------------------- */
typedef struct of_tvl1_batch of_tvl1_batch;
enum { OF_BATCH_LOCKSTEP = 0, OF_BATCH_WORKERS = 1 };
int of_tvl1_batch_create(const of_tvl1_params *params, int streams, of_tvl1_batch **out);
void of_tvl1_batch_destroy(of_tvl1_batch *b);
/* OF_BATCH_LOCKSTEP or OF_BATCH_WORKERS.  In lockstep mode `streams` is the
 * number of lanes: groups of pairs (balanced by pixel count) solved
 * concurrently, each on its own stream, all driven by the calling thread. */
int of_tvl1_batch_set_mode(of_tvl1_batch *b, int mode);
/* Launch shape (iterations per launch, rows per wave) for scales of at
 * least `min_pixels` pixels: in lockstep mode chosen by the largest pair
 * at each scale, in worker mode applied to every handle. */
int of_tvl1_batch_set_shape(of_tvl1_batch *b, long long min_pixels, int iters_per_launch,
                            int rows_per_wave);
/* Launch-shape autotuner of the batch (no reference counterpart).  Solves
 * the given batch with each of OF_BATCH_TUNE_SHAPES shapes (iterations per
 * launch, rows per wave) = (4,8), (2,8), (8,8), (4,4), (2,4) at every scale,
 * then with the three fastest paired (one at scales of at least 200000
 * pixels, another below); each variant once untimed and `reps` times timed.
 * Installs the fastest (as of_tvl1_batch_set_shape would) and returns it in
 * chosen[4] = {K large, R large, K small, R small}; ms_table (optional,
 * OF_BATCH_TUNE_SHAPES entries) receives the uniform shapes' ms per batch,
 * best_ms (optional) the chosen variant's.  The flows in dU are those of a
 * solve (every shape gives the same bits). */
#define OF_BATCH_TUNE_SHAPES 5
int of_tvl1_batch_autotune(of_tvl1_batch *b, int npairs, const int *widths, const int *heights,
                           const uint8_t *const *dI0, const uint8_t *const *dI1, float *const *dU,
                           int reps, int *chosen, double *ms_table, double *best_ms);
/* Solves pair i: device u8 frames dI0[i], dI1[i] of widths[i] x heights[i],
 * flow (planar [u1|u2]) into device buffer dU[i].  `iterations_total`
 * (optional) receives each pair's total primal-dual iteration count.
 * Returns when all pairs are done. */
int of_tvl1_batch_run_u8(of_tvl1_batch *b, int npairs, const int *widths, const int *heights,
                         const uint8_t *const *dI0, const uint8_t *const *dI1, float *const *dU,
                         int *iterations_total);

/* One call for a batch (no reference counterpart): of_tvl1_batch_create
 * with 3 lanes, of_tvl1_batch_run_u8, of_tvl1_batch_destroy.  Callers that
 * solve batches repeatedly keep a batch object instead (it pools handles). */
int of_tvl1_run_batch(const of_tvl1_params *params, int npairs, const int *widths,
                      const int *heights, const uint8_t *const *dI0, const uint8_t *const *dI1,
                      float *const *dU, int *iterations_total);

/* ------------------------------------------------------------------------
 * Two-phase TV-L1 solve (no reference counterpart; used by warm starts).
 * of_tvl1_run_device_u8 = prepare + solve_prepared(NULL, 0, 0).
 * ---------------------------------------------------------------------- */
/* Initial flow for the next solve of this handle (any of_tvl1_run_* or
 * of_tvl1_push_f32 call; no reference counterpart: the reference starts the
 * coarsest scale from zero, tvl1.cpp:157-158).  u and v (host or device
 * memory) hold scale `level`'s flow components, of_tvl1_scale_size(level)
 * floats each, in that scale's pixel units.  The next solve starts at
 * `level` from that flow with the handle's warpings and iteration cap and
 * skips the coarser scales; the flow is used once.  u = v = NULL clears it. */
int of_tvl1_set_init_flow(of_tvl1 *h, int level, const float *u, const float *v);
/* Phase 1 of runDualTVL1Multiscale (src/tvl1/sequential/tvl1.cpp:137-154):
 * frames to float, joint normalisation, pre-smoothing, both pyramids. */
int of_tvl1_prepare_device_u8(of_tvl1 *h, const uint8_t *dI0, const uint8_t *dI1, void *stream);
/* Device pointers to pyramid scale s of the prepared I0 and I1 (row-major,
 * of_tvl1_scale_size), valid until the next prepare or run.  OF_ERR_STATE
 * before the first prepare. */
int of_tvl1_scale_images(const of_tvl1 *h, int s, const float **dI0s, const float **dI1s);
/* Phase 2 (tvl1.cpp:157-172): coarse-to-fine solve of scales first_scale
 * .. 0 (first_scale = -1: the coarsest, as the reference) from `dUinit`, a
 * device planar [u1|u2] flow at scale first_scale in that scale's pixel
 * units, or zero flow when NULL (as the reference).  At most `warps`
 * warpings per scale (<= the handle's) and `max_iterations` per warping; 0
 * selects the handle's value.  Flow into dU (device, 2*width*height
 * floats).  May be called again on the same prepared pair. */
int of_tvl1_solve_prepared(of_tvl1 *h, const float *dUinit, int first_scale, int warps,
                           int max_iterations, float *dU, void *stream);
/* Phase 2 with limits per scale (no reference counterpart; warm-start
 * schedules): warps[s] warpings (<= the handle's) and max_iterations[s]
 * iterations per warping at scale s, arrays of of_tvl1_nscales entries,
 * finest scale first; 0 selects the handle's value.  Otherwise as
 * of_tvl1_solve_prepared. */
int of_tvl1_solve_prepared_schedule(of_tvl1 *h, const float *dUinit, int first_scale,
                                    const int *warps, const int *max_iterations, float *dU,
                                    void *stream);
/* Carries a planar flow from scale from_scale to the finer scale to_scale
 * the way the solve does between scales (zoom_in, zoom.cpp:61-85, times
 * 1/zfactor per scale, tvl1.cpp:167-172). */
int of_tvl1_lift_flow(of_tvl1 *h, int from_scale, const float *dUfrom, int to_scale, float *dUto,
                      void *stream);
/* Resamples a full-resolution planar flow to scale s: zoom_out
 * (zoom.cpp:20-53) of each component s times, times zfactor^s (training
 * labels of the warm start). */
int of_tvl1_flow_to_scale(of_tvl1 *h, const float *dU, int s, float *dUs, void *stream);
/* Copies the flow that scale s ended with in the last solve (planar, that
 * scale's size and units) into dUs. */
int of_tvl1_scale_flow(of_tvl1 *h, int s, float *dUs, void *stream);
/* TV-L1 energy of the device flow dUs at scale s of the prepared pair:
 * E = lambda * sum |I1(x + u) - I0(x)| + sum (|grad u1| + |grad u2|), with
 * I1 sampled by the bicubic of the warps (clamped border) and forward
 * differences.  energy[0] = E, energy[1] = data sum, energy[2] = TV sum. */
int of_tvl1_energy(of_tvl1 *h, int s, const float *dUs, double *energy, void *stream);

/* ------------------------------------------------------------------------
 * Learned warm start (no reference counterpart).  A conv3x3 network,
 * 5 -> 32 channels, `hidden_layers` x (32 -> 32), 32 -> 2, ReLU between,
 * predicts the coarsest-scale flow from that scale's I0, I1, Ix, Iy, It.
 * Convolutions run as implicit GEMMs on bf16 MFMA with fp32 accumulation;
 * tensors are NHWC fp32 in device memory.  Parameters (of_cnn_get_params
 * order): for each layer W[out][tap 0..8][in] then bias[out].
 * ---------------------------------------------------------------------- */
typedef struct of_cnn of_cnn;
/* channels must be 32; He-normal initial weights from `seed`. */
int of_cnn_create(int hidden_layers, int channels, unsigned long long seed, of_cnn **out);
void of_cnn_destroy(of_cnn *h);
long long of_cnn_num_params(const of_cnn *h);
int of_cnn_get_params(const of_cnn *h, float *host);
/* Loads fp32 parameters and resets the optimiser state. */
int of_cnn_set_params(of_cnn *h, const float *host);
/* Network input of one scale: dX[H*W][5] = (I0, I1, Ix, Iy, It) / 255 with
 * centred differences of I1 (clamped at the border). */
int of_cnn_features(const float *dI0s, const float *dI1s, int H, int W, float *dX, void *stream);
/* dX: [B][H][W][5] -> dY: [B][H][W][2].  Asynchronous on `stream`. */
int of_cnn_forward(of_cnn *h, const float *dX, int B, int H, int W, float *dY, void *stream);
/* Features + forward of one scale pair; dU receives planar [u1|u2]. */
int of_cnn_predict(of_cnn *h, const float *dI0s, const float *dI1s, int H, int W, float *dU,
                   void *stream);
/* Loss L = sum (Y - T)^2 / (2*B*H*W) against planar targets dT[B][2][H*W];
 * host_grads (optional) receives dL/dparams, loss (optional) L. */
int of_cnn_gradients(of_cnn *h, const float *dX, const float *dT, int B, int H, int W,
                     float *host_grads, double *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on that loss. */
int of_cnn_train_step(of_cnn *h, const float *dX, const float *dT, int B, int H, int W, float lr,
                      double *loss);
/* Warm-started TV-L1 of one device u8 pair: predicts the coarsest-scale
 * flow, carries it to scale first_scale (-1: the coarsest) with
 * of_tvl1_lift_flow, and keeps it if its energy there (of_tvl1_energy) is
 * not above zero flow's: then scales first_scale .. 0 are solved with at
 * most `warps` warpings and `max_iterations` per warping.  Otherwise
 * (*fell_back = 1) the plain solve runs (zero flow at the coarsest scale,
 * the handle's parameters).  energies (optional): [0] predicted, [1] zero
 * flow, at first_scale. */
int of_tvl1_warm_run_device_u8(of_tvl1 *tv, of_cnn *net, const uint8_t *dI0, const uint8_t *dI1,
                               int first_scale, int warps, int max_iterations, float *dU,
                               int *fell_back, double *energies, void *stream);

/* ------------------------------------------------------------------------
 * Synthetic frame pairs on the device (no reference counterpart; the
 * measurement and training inputs).  frame0 = U(0,1) noise, Gaussian blur
 * sigma 2, rescaled to [0,255]; flow = sum of 4 sinusoids per component
 * plus `nbumps` Gaussian bumps, scaled to max |flow| = cap; frame1 =
 * frame0 sampled bilinearly (clamped) at x - flow(x) plus N(0, noise^2);
 * both rounded and clamped to u8.  params: for each component k and
 * sinusoid j, (period, angle, phase, amplitude) at params[(4k + j)*4], then
 * nbumps x (cx, cy, radius, du, dv) from params[32].  dflow receives the
 * scaled flow (planar), dscratch needs 2*width*height + 4 floats.
 * Asynchronous on `stream`.
 * ---------------------------------------------------------------------- */
int of_synth_pair(int width, int height, float cap, float noise, unsigned long long seed,
                  const float *params, int nbumps, uint8_t *dI0, uint8_t *dI1, float *dflow,
                  float *dscratch, void *stream);

/* ------------------------------------------------------------------------
 * Output formats of a planar flow [u | v] (width*height floats each).
 * ---------------------------------------------------------------------- */
/* App::flowToColor (src/tvl1/sequential/main.cpp:93-126): hue = atan2(v,u)
 * in degrees / 2, value = min(10*|flow|, 255), each rounded to 8 bits with
 * saturation, saturation 255, HSV -> BGR; 3 bytes per pixel, row-major.
 * Device variant asynchronous on `stream`; host variant synchronous. */
int of_flow_to_color_device(const float *dU, int width, int height, uint8_t *dBGR, void *stream);
int of_flow_to_color(const float *U, int width, int height, uint8_t *bgr);
/* Middlebury .flo (no reference counterpart): float tag 202021.25, int32
 * width, int32 height, then (u, v) float pairs row by row, little-endian.
 * of_flo_read stores the size and, when U is not NULL, the planar flow
 * (cap = capacity of U in floats). */
int of_flo_write(const char *path, const float *U, int width, int height);
int of_flo_read(const char *path, float *U, long long cap, int *width, int *height);

/* ------------------------------------------------------------------------
 * Horn-Schunck.  Replaces initFlow / ComputeFlow / deleteFlow_mem,
 * src/hs/cuda/flow.cuh:19-33 (global state there; a handle here).
 * ---------------------------------------------------------------------- */
typedef struct of_hs of_hs;

/* initFlow(int nLevels, int stride, int width, int height)
 * (src/hs/cuda/flow.cu:46-72). */
int of_hs_create(int nLevels, int stride, int width, int height, of_hs **out);
/* deleteFlow_mem(int nLevels) (src/hs/cuda/flow.cu:74-98). */
void of_hs_destroy(of_hs *h);
/* ComputeFlow(I0, I1, width, height, stride, alpha, nLevels, nWarpIters,
 * nSolverIters, u, v) (src/hs/cuda/flow.cu:112-208).  Host pointers with
 * row stride `stride` (floats); u, v receive stride*height floats. */
int of_hs_compute(of_hs *h, const float *I0, const float *I1, int width, int height, int stride,
                  float alpha, int nLevels, int nWarpIters, int nSolverIters, float *u, float *v);
/* Jacobi launch shape: `sweeps_per_launch` sweeps per launch on tiles of
 * 64 x (8*rows_per_wave) pixels (8 waves per workgroup); supported:
 * sweeps_per_launch in {1,2,4,8}, rows_per_wave in {2,4,8}. */
int of_hs_set_jacobi_shape(of_hs *h, int sweeps_per_launch, int rows_per_wave);
/* Device-pointer variant (same layout), synchronous on `stream`. */
int of_hs_compute_device(of_hs *h, const float *dI0, const float *dI1, int width, int height,
                         int stride, float alpha, int nLevels, int nWarpIters, int nSolverIters,
                         float *du, float *dv, void *stream);

/* ------------------------------------------------------------------------
 * Lucas-Kanade.  Replaces class LucasKanade, src/lk/cuda/lucaskanade.cuh:6-35.
 * ---------------------------------------------------------------------- */
typedef struct of_lk of_lk;

/* LucasKanade(int spac_filt_size, int temp_filt_size, int window_size,
 * int nx, int ny) (src/lk/cuda/lucaskanade.cu:5-32).  Spatial filter sizes
 * 1/3/5/7/9, temporal 1/2/3/5/7/9 (lucaskanade.cu:79-161). */
int of_lk_create(int spac_filt_size, int temp_filt_size, int window_size, int nx, int ny,
                 of_lk **out);
void of_lk_destroy(of_lk *h);
/* copy_frames_circularbuffer_GPU_wrapper(frame, temp_size, fr, frame_size)
 * (src/lk/cuda/lucaskanade.cu:74-76, kernels.cu:6-10): frame `fr` goes into
 * slot fr % temp_size of the device ring. */
int of_lk_push_u8(of_lk *h, const uint8_t *frame, int temp_size, int fr, int frame_size);
/* lucas_kanade(float* Vx, float* Vy, int iframe) (lucaskanade.cu:54-71):
 * host outputs of nx*ny floats. */
int of_lk_compute(of_lk *h, float *Vx, float *Vy, int iframe);

/* Pyramidal iterative LK on a float pair (no reference counterpart: the
 * reference's LK is single-scale).  `levels` Gaussian-pyramid levels
 * (blur `sigma`, decimation by `zf`), `window` x `window` sums, `iters`
 * iterations per level, update only where det(G) > det_eps.  Host pointers;
 * u, v receive width*height floats. */
int of_lk_pyr_compute(const float *I0, const float *I1, int width, int height, int levels,
                      float zf, float sigma, int window, int iters, float det_eps, float *u,
                      float *v);
/* The same on device pointers, on `stream` (a hipStream_t; NULL = the null
 * stream); returns when du, dv hold the flow.  The frames are read in place;
 * du and dv may overlap the frames (then the result is copied out last). */
int of_lk_pyr_compute_device(const float *dI0, const float *dI1, int width, int height,
                             int levels, float zf, float sigma, int window, int iters,
                             float det_eps, float *du, float *dv, void *stream);

/* No reference counterpart.  The same solve as of_lk_pyr_compute_device,
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): it calls
 * no hipMalloc/hipFree or allocating routine and keeps its workspace in a
 * fixed 128 MiB block of device memory per GPU, held from the library's
 * first use on that GPU until the process ends.  The workspace needs
 * 4 bytes x (2 x (pixels of levels 1..) + 15 x width*height), each part
 * rounded up to 256 bytes: up to 1920x1080 at any level count.  Larger
 * inputs return OF_ERR_CAPACITY.  Every call on one GPU shares that block:
 * calls that overlap in time on one GPU (other host threads, other streams)
 * race and must be serialised by the caller.  Results are bit-identical to
 * of_lk_pyr_compute_device. */
int of_lk_pyr_compute_device_noalloc(const float *dI0, const float *dI1, int width, int height,
                                     int levels, float zf, float sigma, int window, int iters,
                                     float det_eps, float *du, float *dv, void *stream);

#ifdef __cplusplus
}
#endif
#endif
