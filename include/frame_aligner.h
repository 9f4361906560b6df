/*
 * frame_aligner.h - C ABI of the MI355X (gfx950) warp-registration library.
 *
 * Drop-in boundary for the non-rigid warp path of flimfit/frame-aligner.
 * The entry points are what a host written against the reference's
 * AbstractFrameWarper (FrameAligner/FrameWarper.h:95-143) and its
 * trust-region caller (FrameAligner/FrameWarpAligner.cpp:117-223,
 * FrameAligner/OptimisationModel.cpp:45-63) binds: plain pointers and sizes,
 * no C++ or torch types, status codes instead of exceptions.
 *
 * Conventions
 *  - Volumes are (nz, ny, nx) with x fastest (FrameAligner/Cv3dUtils.cpp:6-8).
 *    A 2-D image is nz = 1; n_dim = 3 when nz > 1, else 2
 *    (FrameAligner/FrameWarper.cpp:198).
 *  - D, the node displacements, is nD x 3 doubles (x, y, z) per node, as the
 *    reference's std::vector<cv::Point3d>; z is ignored in 2-D.
 *  - Parameter vectors (trust region x, Jacobian) use the reference's D2col
 *    layout: [x0, y0, (z0), x1, y1, ...] (FrameAligner/OptimisationModel.cpp:14-27).
 *  - Every call returns FA_OK (0) or a negative FA_E_* code; fa_last_error()
 *    returns the calling thread's last message. Calls on distinct frames may
 *    run from different threads; a reference is read-only after set-up.
 *  - Per-frame calls (fa_get_error, fa_get_jacobian, fa_warp_*, fa_align) lock
 *    only their frame and run on the frame's own stream, so calls on different
 *    frames from different threads run at the same time on the device; calls on
 *    one frame take turns. Batch runs and evaluations lock only their batch.
 *    fa_align_noalloc, servers and set-up calls (references, frame registration,
 *    batch creation) take the context's lock.
 */
#ifndef FRAME_ALIGNER_H
#define FRAME_ALIGNER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FA_OK 0
#define FA_E_ARG (-1)           /* invalid argument / size ("Frame is incorrect size") */
#define FA_E_OOM (-2)           /* device allocation failed ("Not enough memory to allocate"), or host memory or threads exhausted */
#define FA_E_UNKNOWN_FRAME (-3) /* handle not registered ("Unrecognised frame") */
#define FA_E_DEVICE (-4)        /* HIP runtime error or no supported GPU ("No supported GPU") */
#define FA_E_NODE_MAP (-5)      /* node map inconsistent ("Something went wrong") */
#define FA_E_STATE (-6)         /* call out of order */

typedef struct fa_context fa_context;
typedef struct fa_reference fa_reference;
typedef struct fa_frame fa_frame;
typedef struct fa_batch fa_batch;
typedef struct fa_server fa_server;

/* Raw constructor arguments of ImageScanParameters (FrameAligner/ImageScanParameters.h:7-20);
   the library applies the constructor's fix-ups and processScanParameters
   (FrameAligner/FrameWarper.cpp:223-236). bidirectional must be 0: the reference never
   initialises it (ImageScanParameters.h:31) and its bidirectional node map is inconsistent. */
typedef struct fa_scan_params {
   double line_duration;
   double interline_duration;
   double interframe_duration;
   int32_t n_x, n_y, n_z;
   int32_t bidirectional;
} fa_scan_params;

enum fa_dtype { FA_F32 = 0, FA_U16 = 1 };

/* find_min_trust_region(objective_delta_stop_strategy(min_delta), model, x, radius)
   (FrameAligner/FrameWarpAligner.cpp:205-209), plus the relative stop and caps the
   benchmark configurations ask for. */
typedef struct fa_tr_params {
   double radius;        /* initial radius (reference: 40) */
   double min_delta;     /* stop when |f - f_prev| < min_delta (reference: 1e-8) */
   double rel_delta;     /* stop when |f - f_prev| < rel_delta * |f_prev|; 0 = off */
   int32_t max_iter;     /* stop-strategy iterations; 0 = unlimited (reference) */
   int32_t max_evals;    /* cap on objective evaluations; 0 = unlimited */
   double sub_eps;       /* subproblem boundary tolerance (0.1) */
   int32_t sub_max_iter; /* subproblem iterations (20) */
   int32_t skip_repeat;  /* 1: skip re-evaluating a rejected step proposed unchanged */
} fa_tr_params;

enum fa_stop {
   FA_STOP_ABS_DELTA = 1,
   FA_STOP_REL_DELTA = 2,
   FA_STOP_MAX_ITER = 3,
   FA_STOP_NO_PROGRESS = 4,
   FA_STOP_NONFINITE = 5,
   FA_STOP_RADIUS = 6,
   FA_STOP_MAX_EVALS = 7
};

typedef struct fa_align_stats {
   int32_t iterations;   /* stop-strategy calls */
   int32_t evaluations;  /* objective evaluations (fused passes) */
   int32_t rejected;     /* rejected steps */
   int32_t status;       /* FA_STOP_* */
   double f_final;       /* SSD at the returned x */
   double radius_final;
} fa_align_stats;

/* Device-side timing of a batch run, filled when fa_batch_set_timing(b, 1). */
typedef struct fa_batch_timing {
   double wall_ms;       /* host wall time of the run */
   double pass_ms;       /* sum of fused-pass kernel durations (HIP events on the batch stream) */
   int64_t pass_launches;
   double pass_bytes;    /* algorithmic bytes of those launches: active pixels x (s_f + 4 + 4 n_dim),
                            s_f = 2 (uint16 frames) or 4 (float): frame, reference, gradient */
   int64_t tr_rounds;    /* launch rounds (pass + update) issued */
   double pass_px;       /* active pixels of those launches */
} fa_batch_timing;

/* Launch shape of the fused pass, chosen per batch by the autotuner (fa_amd.tune):
   lds_floats - floats of LDS per workgroup for staged frame rows (0 disables staging;
                at most 15360); work items whose reachable rows need more use global loads;
   blocks_per_cu - persistent workgroups per compute unit;
   use_lds - 0 forces global loads; check_every - trust-region rounds between host polls;
   use_graph - 1 replays the check_every rounds as one captured hipGraph (runs with
               fa_batch_set_timing on launch eagerly, to time each pass with HIP events);
   tr_lds - 1 runs the trust-region update with its factor blocks in LDS when the largest
            frame's fit one workgroup's LDS, 0 keeps them in global memory; reading the
            tuning back gives the effective choice. */
typedef struct fa_tuning {
   int32_t lds_floats;
   int32_t blocks_per_cu;
   int32_t use_lds;
   int32_t check_every;
   int32_t use_graph;
   int32_t tr_lds;
} fa_tuning;

const char* fa_last_error(void);
void fa_tr_default(fa_tr_params* p);

/* Replaces GpuFrameWarper() / getGpuSupportInformation (FrameAligner/GpuFrameWarper.cpp:10-44). */
int fa_create(int device, fa_context** out);
int fa_destroy(fa_context* ctx);

/* AbstractFrameWarper::setReference(reference, nD, image_params)
   (FrameAligner/FrameWarper.cpp:186-205; GPU set-up GpuFrameWarper.cpp:75-152).
   ref: nz*ny*nx floats, already smoothed and offset by the caller, as the reference's
   aligner passes it (FrameWarpAligner.cpp:86-102). Builds the node map, Sobel gradient,
   block-tridiagonal JtJ and its Cholesky factor on the device. Work items (pixels per
   workgroup task) cut the part of each node interval in one plane into the fewest pieces of
   at most chunk_px pixels, of equal length; chunk_px = 0 picks the default. */
int fa_set_reference(fa_context* ctx, const float* ref, int32_t nz, int32_t ny, int32_t nx,
                     int32_t nD, const fa_scan_params* scan, int32_t chunk_px,
                     fa_reference** out);
int fa_reference_release(fa_reference* ref);
/* out[0..5] = nz, ny, nx, nD, n_dim, number of work chunks. */
int fa_reference_info(const fa_reference* ref, int32_t* out6);
/* AbstractFrameWarper::H (FrameWarper.h:134), dense (nD*n_dim)^2 row-major. */
int fa_get_hessian(const fa_reference* ref, double* H);
/* Node map Di/Df and D_range (FrameWarper.h:132-136); any pointer may be NULL.
   D_range is nD x (begin, end). */
int fa_get_node_map(const fa_reference* ref, uint16_t* Di, float* Df, int32_t* D_range);
/* Sobel gradient of the reference (Sobel3d.cpp:60-84); gz may be NULL in 2-D. */
int fa_get_gradient(const fa_reference* ref, float* gx, float* gy, float* gz);

/* GpuFrameWarper::registerFrame / deregisterFrame (FrameAligner/GpuFrameWarper.cpp:46-73).
   data: nz*ny*nx values of dtype in host memory, copied to a resident device slot.
   Samples are data + add_offset (the aligner's +1.0f, FrameWarpAligner.cpp:142, can be
   applied here for raw uint16 counts; pass 0 for frames already offset). */
int fa_register_frame(fa_reference* ref, const void* data, int32_t dtype, float add_offset,
                      fa_frame** out);
int fa_deregister_frame(fa_frame* frame);
/* GpuFrameWarper::clearTemp (FrameAligner/GpuFrameWarper.cpp:154-158; FrameWarper.h:118): drops
   the device working spaces the library caches for the reference's registered frames (the
   one-frame batch behind fa_align and the single evaluations); they are rebuilt on next use,
   and the frames stay registered. */
int fa_clear_temp(fa_reference* ref);

/* getError (FrameWarper.h:103; CPU FrameWarper.cpp:169-176, GPU GpuFrameWarper.cpp:182-190):
   SSD of the warped frame against the reference over valid pixels. */
int fa_get_error(fa_frame* frame, const double* D, double* err);
/* getJacobian (FrameWarper.h:104; FrameWarper.cpp:147-167,178-182; GpuFrameWarper.cpp:192-235):
   Jtr at D (nD*n_dim, D2col layout). err (may be NULL) receives the SSD at D. */
int fa_get_jacobian(fa_frame* frame, const double* D, double* jac, double* err);
/* warpImageInterpolated (FrameWarper.h:106; FrameWarper.cpp:31-69): out gets nz*ny*nx floats;
   samples that are exactly zero become invalid_value. */
int fa_warp_interpolated(fa_frame* frame, const double* D, float* out, float invalid_value);
/* warpImage (FrameWarper.h:107; FrameWarper.cpp:71-136): forward splat of every source pixel
   to p - d(p) with trilinear weights; wimg gets the weighted values, coverage the weights
   (nz*ny*nx floats each). apply_offset = 0 splats the raw registered values (the aligner
   splats the raw frame, FrameWarpAligner.cpp:317-346); 1 adds the registration offset.
   warpImageNormalised (FrameWarper.cpp:138-144) is this followed by a type conversion. */
int fa_warp_image(fa_frame* frame, const double* D, float* wimg, float* coverage,
                  int32_t apply_offset);
/* warpPoint (FrameWarper.h:110; FrameWarper.cpp:397-419). */
int fa_warp_point(const fa_reference* ref, const double* D, int32_t x, int32_t y, int32_t z,
                  int32_t spatial_binning, double* out3);

/* The trust-region loop of FrameWarpAligner::addFrame (FrameWarpAligner.cpp:200-215) run
   on the device for one frame: x_inout (nD*n_dim) is the start and receives the result. */
int fa_align(fa_frame* frame, const fa_tr_params* prm, double* x_inout, fa_align_stats* stats);

/* fa_align without allocating: no counterpart in the reference; for callers that cannot
   allocate device memory (code under graph capture, serving loops, harnesses that forbid it).
   Same kernels and results as fa_align, with the frame's trust-region buffers in a 4 MiB
   device array the library holds per GPU for the life of the process (references whose
   buffers exceed it are rejected with FA_E_ARG). It launches eagerly on the context's stream
   and waits for the result. Calls on one GPU share the array: calls from different contexts
   that overlap on one GPU race. */
int fa_align_noalloc(fa_frame* frame, const fa_tr_params* prm, double* x_inout, fa_align_stats* stats);

/* Batched form: many frames (any references, sizes and dtypes) aligned together, each
   trust-region round one fused pass over every active frame plus one device update. */
int fa_batch_create(fa_frame* const* frames, int32_t n, const fa_tr_params* prm,
                    fa_batch** out);
int fa_batch_destroy(fa_batch* b);
/* enable = 1: the following runs launch eagerly and time every fused pass with HIP events
   (fa_batch_timing); a batch large enough to run as two interleaved parts then runs its parts
   one after the other, so each pass is timed alone. Results are the same either way. */
int fa_batch_set_timing(fa_batch* b, int32_t enable);
/* Diagnostics, no counterpart in the reference: records every objective evaluation of the
   trust region of each frame in the following runs (at most max_records per frame; 0 turns
   tracing off). A record is FA_TRACE_FIELDS doubles - evaluations so far, f of this
   evaluation, f at x after the decision, predicted and measured decrease of the step tried,
   rho, accepted, radius after the decision, stop-strategy iterations, stop status (0 while
   running), lambda and iterations of the next subproblem, |p| of the next step, |g| at x, g'p,
   rejected steps so far - followed by x, g and p (nD*n_dim each) after the record's update,
   padded to FA_TRACE_FIELDS + 3 * n_max doubles (n_max: the batch's largest nD*n_dim). */
#define FA_TRACE_FIELDS 16
int fa_batch_set_trace(fa_batch* b, int32_t max_records);
/* out: n_frames x max_records records (caller's frame order) of the last run; records past a
   frame's last evaluation are zero. */
int fa_batch_get_trace(fa_batch* b, double* out);
int fa_batch_get_tuning(const fa_batch* b, fa_tuning* t);
int fa_batch_set_tuning(fa_batch* b, const fa_tuning* t);
/* x_inout: concatenated per-frame parameter vectors; stats: n entries (may be NULL). */
int fa_batch_run(fa_batch* b, double* x_inout, fa_align_stats* stats, fa_batch_timing* timing);
/* One fused pass over every frame of the batch at the given D (concatenated nD x 3 per
   frame): err[n] SSDs and jac (concatenated nD*n_dim per frame); either may be NULL. */
int fa_batch_eval(fa_batch* b, const double* D, double* err, double* jac);
int fa_align_batch(fa_frame* const* frames, int32_t n, const fa_tr_params* prm,
                   double* x_inout, fa_align_stats* stats);

/* Serving (continuous batching). The references are grouped into lanes by kind (2-D, 3-D)
   and size (a lane's largest reference has at most 16 times the pixels of its smallest; at
   most 3 lanes). Each lane runs its own slot batch on one of the context's three lane streams
   (a server with one lane and 16 or more slots runs it as two that interleave), with `slots`
   slots each sized for the lane's largest reference; at most `slots` frames are admitted at
   once over all lanes. The lane streams are created by fa_create with the context's own stream,
   so they keep hardware queues of their own; host frames are uploaded on the context's stream.
   Servers of one context share its lane streams (their rounds then run one after another). Host
   frames go into a device frame arena of arena_bytes (0: slots times the largest float frame,
   within the device's free memory less 32 MiB; FA_E_OOM "Not enough video memory to use GPU"
   when even one frame does not fit, as GpuFrameWarper::setupReferenceInformation,
   GpuFrameWarper.cpp:113-130). Submitted frames wait in arrival order.
   fa_server_step: (1) takes the results of every lane whose launched rounds have finished:
   x_inout (caller memory, valid until the result is returned) receives the result and
   fa_server_results lists the finished tags; (2) admits queued frames in arrival order while
   fewer than `slots` are admitted, the frame's lane has a free slot and (host frames) the
   arena has room for it (admission by free memory: the oldest frame that does not fit waits,
   and those behind it with it), uploading host frames on a copy stream that overlaps the
   running rounds; (3) for every lane with no rounds in flight, starts its frames whose data is
   on the device and launches `rounds` trust-region rounds over its running slots (frames
   admitted at different times share the launches) without waiting for them. When no lane had
   finished at (1), the step waits up to 0.5 ms for one to finish and repeats (1)-(3), so a
   lane of small frames is harvested and relaunched on its own cadence, not on that of a lane
   of large frames. Each result is the same as fa_align's for that frame. Host frames should
   be pinned (fa_host_alloc) for the upload to be asynchronous, and stay untouched until their
   result is returned; registered frames stay registered until then. Use one thread per
   server. There is no counterpart in the reference beyond its thread pool of addFrame tasks
   (FrameAlignerMex/FrameAlignerMex.cpp:110-130). */
typedef struct fa_server_result {
   int64_t tag;
   fa_align_stats stats;
} fa_server_result;

int fa_server_create(fa_context* ctx, fa_reference* const* refs, int32_t n_refs, int32_t slots,
                     uint64_t arena_bytes, const fa_tr_params* prm, fa_server** out);
int fa_server_destroy(fa_server* server);
/* A registered frame (fa_register_frame). */
int fa_server_submit(fa_server* server, fa_frame* frame, double* x_inout, int64_t tag);
/* A frame in host memory, nz*ny*nx values of dtype, sampled as value + add_offset (as
   fa_register_frame), uploaded into the arena when admitted. */
int fa_server_submit_host(fa_server* server, fa_reference* ref, const void* data, int32_t dtype,
                          float add_offset, double* x_inout, int64_t tag);
int fa_server_step(fa_server* server, int32_t rounds, int32_t* in_flight, int32_t* queued);
int fa_server_results(fa_server* server, fa_server_result* out, int32_t max, int32_t* n_out);
/* A request loop in the library: arrival i, arrival_s[i] seconds after the call starts (in
   non-decreasing order), submits host frame data[i] of refs[i] (dtype[i], add_offset[i], start
   and result vector x_inout + x_offset[i]) with tag i, as fa_server_submit_host; the loop steps
   the server with `rounds` while frames are in flight or queued, sleeps until the next arrival
   otherwise, and returns when every frame's result is in: latency_s[i] = time of the step that
   returned result i - arrival_s[i], stats[i] its trust-region statistics. The frames must stay
   untouched until the call returns. A server that holds other frames (queued, running or with
   results not yet taken) is refused with FA_E_ARG, since results are matched by tag. A frame that
   cannot be submitted (fa_server_submit_host's errors) ends the stream: the frames submitted
   before it are run to their results and its error is returned; after a device error the
   server must not be stepped again. It replaces a
   host-language loop of fa_server_submit_host / fa_server_step / fa_server_results with the
   same calls made natively (no counterpart in the reference beyond its MEX thread pool,
   FrameAlignerMex/FrameAlignerMex.cpp:110-130). */
typedef struct fa_stream_counts {
   int64_t steps;          /* fa_server_step calls */
   int64_t max_in_flight;  /* most frames admitted at once */
   int64_t max_queued;     /* most frames waiting for a slot */
   double wall_s;          /* seconds from the call's start to the last result */
} fa_stream_counts;
int fa_server_run_stream(fa_server* server, int64_t n, const double* arrival_s, fa_reference* const* refs,
                         const void* const* data, const int32_t* dtype, const float* add_offset,
                         double* x_inout, const int64_t* x_offset, int32_t rounds, double* latency_s,
                         fa_align_stats* stats, fa_stream_counts* counts);
/* Bytes of the frame arena in use and in total. */
int fa_server_memory(const fa_server* server, uint64_t* used, uint64_t* total);
/* Pinned host memory for frames handed to fa_server_submit_host. */
int fa_host_alloc(uint64_t bytes, void** out);
int fa_host_free(void* p);
/* Free and total device memory (hipMemGetInfo), for admission by free memory. */
int fa_context_mem_info(fa_context* ctx, uint64_t* free_bytes, uint64_t* total_bytes);

/* ---- aligner front end: FrameWarpAligner::setReference / addFrame ------------------------
   RealignmentParameters (FrameAligner/AbstractFrameAligner.h:31-48), the fields the warp
   aligner reads. */
enum fa_realign_type { FA_REALIGN_NONE = 0, FA_REALIGN_TRANSLATION = 1, FA_REALIGN_WARP = 2 };

typedef struct fa_realign_params {
   int32_t type;                /* fa_realign_type (reference default None; Warp runs the TR) */
   int32_t spatial_binning;     /* phase-correlation downsampling in x and y (1) */
   int32_t n_resampling_points; /* nodes per plane (30): nD = n * nz + 1 (FrameWarpAligner.cpp:100) */
   int32_t pad;
   double smoothing;            /* Gaussian sigma in x of the per-slice blur, 0 = none; sigma in y
                                   is 1 (FrameWarpAligner.cpp:46-64) */
} fa_realign_params;

typedef struct fa_aligner fa_aligner;

/* Per frame result of fa_aligner_add_frames: the three starts of addFrame
   (FrameWarpAligner.cpp:150-198) are {0, 0, rigid}; start is the index of the first with the
   smallest SSD. */
typedef struct fa_add_result {
   int32_t start;
   int32_t pad;
   double rigid_shift[3];  /* phase-correlation shift: x, y (scaled by the binning), z */
   double start_error[3];  /* SSD at the three starts */
   fa_align_stats stats;   /* trust region (Warp only) */
   /* RealignmentResult metrics (FrameWarpAligner.cpp:238-240; Cv3dUtils.cpp:64-88): Pearson
      correlation with the blurred reference + 1 over the pixels the raw frame's splat covers,
      of the warped and of the unwarped blurred frame + 1, and the covered fraction. */
   double correlation;
   double unaligned_correlation;
   double coverage;
} fa_add_result;

/* Output images of addFrame (each n x nz*ny*nx, any pointer may be NULL), from the raw frame
   at the final D: realigned - getInterpolatedWarpedFrame (warpImageInterpolated, invalid 0, to
   uint16 as cv::Mat::convertTo rounds and saturates); preserving - getWarpedFrame (forward
   splat divided by its coverage, to uint16); mask - getMask (the splat's coverage); warped -
   the blurred frame + 1 warped by warpImageInterpolated (FrameWarpAligner.cpp:218-226,317-346). */
typedef struct fa_add_outputs {
   uint16_t* realigned;
   uint16_t* preserving;
   float* mask;
   float* warped;
} fa_add_outputs;

/* FrameWarpAligner(params) + setReference(0, reference) (FrameWarpAligner.cpp:68-105):
   reference is the raw nz*ny*nx volume of dtype (FA_U16 counts or FA_F32). Builds the blurred
   reference + 1 for the warper (as fa_set_reference) and the phase correlator's reference
   spectrum (windowed, binned, FP64 3-D FFT). */
int fa_aligner_create(fa_context* ctx, const fa_realign_params* rp, const fa_scan_params* scan,
                      const void* reference, int32_t dtype, fa_aligner** out);
int fa_aligner_destroy(fa_aligner* a);
/* The warper's reference (read-only; owned by the aligner) and nD. */
int fa_aligner_reference(fa_aligner* a, fa_reference** ref, int32_t* nD);
/* smoothStack + 1 of one raw frame (FrameWarpAligner.cpp:46-64,141-142): out gets the
   nz*ny*nx floats the warper aligns. */
int fa_aligner_preprocess(fa_aligner* a, const void* raw, int32_t dtype, float* out);
/* Rigid start of one raw frame (VolumePhaseCorrelator::computeShift of the binned frame, x and
   y scaled by the binning, FrameWarpAligner.cpp:186-189). */
int fa_aligner_rigid_shift(fa_aligner* a, const void* raw, int32_t dtype, double* shift3);
/* FrameWarpAligner::addFrame (FrameWarpAligner.cpp:117-250) for n raw frames of dtype: blur
   and +1, the rigid start, the start with the smallest SSD, then (type Warp) the trust region
   of all n frames as one batch, then the output images and metrics. D_out: n x nD x 3;
   outputs and results (n entries) may be NULL. prm NULL = the reference's stop rule. The
   frames run through batched phases (uploads, then rigid shifts and blur up to 32 frames at a
   time, the batch's evaluations and trust region, the output warps), and the aligner keeps
   its device arrays, and the frames and batch of the last call, for the next call with the
   same frame count (fa_aligner_clear_temp frees them). */
int fa_aligner_add_frames(fa_aligner* a, const void* const* frames, int32_t n, int32_t dtype,
                          const fa_tr_params* prm, double* D_out, const fa_add_outputs* outputs,
                          fa_add_result* results);
/* FrameWarpAligner::clearTemp (FrameWarpAligner.cpp:72-75, which calls the warper's
   clearTemp): frees what fa_aligner_add_frames keeps between calls and the warper's cached
   working spaces (fa_clear_temp). */
int fa_aligner_clear_temp(fa_aligner* a);
/* FrameWarpAligner::realignAsFrame (FrameWarpAligner.cpp:270-285): frame (nz*ny*nx values of
   dtype, warped as given: no blur, no offset) realigned with the displacements D (nD x 3) that
   were found for some frame. interpolate_missing = 1: warpImageInterpolated with invalid value
   0 (FrameWarper.cpp:31-69), out gets float32. interpolate_missing = 0: warpImageNormalised
   (FrameWarper.cpp:138-144), the forward splat's weighted values (not divided by the coverage)
   converted as cv::Mat::convertTo does to cv_type: 0 = CV_8U (the reference's default), 2 =
   CV_16U, 5 = CV_32F (round half to even, saturate). The reference's check of the frame's
   usability (FrameWarpAligner.cpp:277-278) is followed by the warp, which replaces its result,
   so it has no effect and is not reproduced. */
int fa_aligner_realign_as_frame(fa_aligner* a, const double* D, const void* frame, int32_t dtype,
                                int32_t interpolate_missing, int32_t cv_type, void* out);
/* FrameWarpAligner::shiftPixel (FrameWarpAligner.cpp:253-266): the point (x, y, z), truncated
   to whole pixels, minus warpPoint(D) there; xyz is updated in place. */
int fa_shift_pixel(const fa_reference* ref, const double* D, double* xyz);
/* FrameWarpAligner::writeRealignmentInfo (FrameWarpAligner.cpp:288-314): the CSV (version 2)
   of the scan timing (ImageScanParameters after its constructor's fix-ups) and, per frame,
   the unaligned correlation, correlation, coverage and the nD displacements. D: n x nD x 3. */
int fa_write_realignment_info(const char* filename, const fa_scan_params* scan, int32_t n, int32_t nD,
                              const double* D, const double* unaligned_correlation,
                              const double* correlation, const double* coverage);

/* out[0..4] = device ordinal, compute units, bytes of LDS one workgroup may allocate,
   wavefront size, hipError_t of raising the kernels' dynamic-LDS limit (0 = success). */
int fa_context_info(const fa_context* ctx, int32_t* out5);

/* The HIP stream the context launches on (hipStream_t), for callers that time it. */
void* fa_context_stream(fa_context* ctx);

#ifdef __cplusplus
}
#endif

#endif
