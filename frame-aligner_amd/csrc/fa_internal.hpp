// Internal structures of the gfx950 warp-registration library: device views of references
// and frames, the per-frame trust-region state, and the host objects behind the C ABI
// handles of include/frame_aligner.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "frame_aligner.h"

namespace fa {

// Derived scan timing in double (ImageScanParameters.h:7-20 after
// processScanParameters, FrameWarper.cpp:223-236): t(z,y,x) = z*interframe +
// y*interline + x*pixel and u = t / stack * (nD - 1).
struct ScanDev {
   double interframe, interline, pixel, stack, nDm1;
};

// One work item of the fused passes: the pixels [p0, p1) of node interval k inside one
// plane. end_p is the last pixel of interval k, which the reference leaves out of the
// steepest-descent images, JtJ and Jtr (its consumers iterate [begin, end),
// FrameWarper.cpp:309-344) but keeps in the SSD.
//
// A work item is either a run of whole-width rows (x0 = 0, w = nx: the pixels p0 .. p1-1) or,
// in frames whose width is a multiple of kStripWidth, a strip of kStripWidth columns
// [x0, x0 + w) over a run of rows: the pixels of those rows and columns between p0 (its
// first) and p1 - 1 (its last) in scan order.
struct Chunk {
   int32_t p0, p1, k, end_p, x0, w;
};
// Column width of strip work items (one column per pass thread).
constexpr int kStripWidth = 256;

// Doubles written per work item by the fused pass: SSD, Jtr of node k (n_dim values,
// weight 1-f), Jtr of node k+1 (n_dim values, weight f); padded to 8.
constexpr int kPassStride = 8;
// Per work item of the JtJ pass: aa (upper, n_dim(n_dim+1)/2), bb (same), ba (n_dim^2);
// 21 in 3-D, padded to 24.
constexpr int kHessStride = 24;
constexpr int kBlock = 256;  // threads per fused-pass workgroup (4 wave64)
static_assert(kStripWidth == kBlock, "a strip item has one column per fused-pass thread");
// Zero columns staged on each side of a frame row in LDS: taps need no bounds checks while
// |dx| stays below kPad - 2 over a work item.
constexpr int kPad = 16;
constexpr int kMaxRefs = 4096;  // distinct references in one batch
constexpr int kEigMax = 90;     // largest nD * n_dim whose subproblem uses the eigendecomposition of H
// nD * n_dim range whose subproblem runs in the batched eigen-space kernels (Q, Q^T: up to
// 144 MiB). They cost four short launches per round, which the cyclic reduction's update beats
// on shorter chains (serving references of 231 to 843 parameters ran slower, C3's 1539 faster).
constexpr int kEigBatchMin = 1000, kEigBatchMax = 3072;

// Per-pixel reference record read by every pass. 2-D: (R, gx, gy, Df) in one float4;
// 3-D: (R, gx, gy, gz) plus Df in a separate float array. Df is the node fraction of
// precomputeInterp (FrameWarper.cpp:258-267), computed once in FP64 and stored as float
// exactly as the reference stores it.
struct RefDev {
   const float4* rg4;
   const float* df;        // 3-D only (2-D keeps Df in rg4.w)
   const Chunk* chunks;    // work items, ordered by interval then pixel
   const int32_t* ichunk;  // first chunk of interval k (k = 0..nD-2), ichunk[nD-1] = n_chunks
   const double* Hdiag;    // nD blocks n_dim x n_dim
   const double* Hsub;     // nD-1 blocks, block (k+1, k)
   const double* L0;       // cyclic-reduction Cholesky factor of H (A, C, W blocks; valid when h_pd)
   const double* Q;        // eigenvectors of H, n x n row-major, column i for ev[i], then Q^T (valid when eig)
   const double* ev;       // eigenvalues of H
   ScanDev scan;
   int32_t nx, ny, nz, nD, n_dim, n_chunks, h_pd;
   int32_t eig;            // 0: none; 1: Q, Q^T (n <= 90, solved in the update); 2: Q, Q^T (batched kernels)
   int32_t q_lds;          // 1: an update that stages its working set in LDS also stages Q
   double eig_lo;          // Gershgorin lower bound of the spectrum of H
   double ev_min;          // smallest eigenvalue (valid when eig)
   int64_t n_px;
};

struct FrameDev {
   const void* data;
   int32_t dtype;  // FA_F32 / FA_U16
   float offset;
};

// Per frame of a batch.
struct BatchFrame {
   FrameDev f;
   int32_t ref;        // index into the batch's reference table
   int32_t n;          // nD * n_dim
   int64_t work_base;  // first work item of this frame (all frames active)
   int64_t x_base;     // offset of this frame's parameter vector x
   int64_t vec_base;   // offset of this frame's work vectors (p, g, t1, t2) in the pool
   int64_t fac_base;   // offset of this frame's factor scratch
   int64_t d_base;     // offset of this frame's nD x 3 trial displacement
   int32_t trc;        // launch class of its trust-region update (LDS size)
   int32_t pad;
   RefDev r;           // its reference (a copy of refs[ref]: one load chain fewer per use)
};

struct TrParamsDev {
   double radius, min_delta, rel_delta, sub_eps;
   int32_t max_iter, max_evals, sub_max_iter, skip_repeat;
};

struct TrState {
   double f, radius, radius0, prev_f, lambda, pnorm;
   int32_t phase, stale, been_used, iterations, evaluations, rejected, status, done, sub_iters,
       pad;
};

static_assert(sizeof(TrState) % 8 == 0, "trust-region states are copied as 8-byte words");
static_assert(sizeof(BatchFrame) % 8 == 0, "frame records are copied as 8-byte words");

// Counters the schedule kernel maintains for the host and the pass kernel.
struct Counters {
   int32_t n_active_frames;
   int32_t upd_done;          // update workgroups finished in this round (folded schedule)
   int64_t n_active_work;
   int64_t n_active_work2;    // work items of active 2-D frames (their references come first)
   unsigned long long bytes;  // algorithmic bytes of the passes scheduled so far
   unsigned long long px;     // active pixels of those passes
   unsigned long long rounds;
};

// Algorithmic bytes of one fused pass per pixel: the stored frame sample, the reference
// intensity and its n_dim gradient components (s_f + 4 + 4 n_dim). The node fraction and the
// 2-D record's padding, which the pass also reads, are not counted.
__host__ __device__ constexpr int pass_bytes_per_px(int dtype, int n_dim)
{
   return (dtype == FA_U16 ? 2 : 4) + 4 + 4 * n_dim;
}

struct PassArgs {
   const RefDev* refs;
   const BatchFrame* frames;
   const double* Dtrial;
   double* partials;
   // Active frames grouped by reference: reference r's active frames are
   // active_frames[run_first[r] .. + run_count[r]) and its work items
   // [run_base[r], run_base[r + 1]), chunk-major (item = chunk * run_count + frame).
   const int32_t* active_frames;
   const int32_t* run_first;
   const int32_t* run_count;
   const int64_t* run_base;
   int32_t n_refs;
   const Counters* counters;
   int32_t lds_cap;   // floats of dynamic LDS available for staging
   int32_t use_lds;
   int32_t ref_lo, ref_hi;  // the launch takes the work items of references [ref_lo, ref_hi)
};

// Fields of one trust-region trace record (one per objective evaluation of a frame), in this
// order: evaluations so far, f of this evaluation, f at x after the decision, predicted and
// measured decrease of the step tried, rho, accepted (1/0), radius after the decision,
// stop-strategy iterations, stop status (0 while running), lambda and iterations of the next
// subproblem, |p| of the next step, |g| at x, g'p, rejected steps so far. x, g and p follow.
constexpr int kTraceFields = FA_TRACE_FIELDS;

struct BatchDev {
   const RefDev* refs;
   BatchFrame* frames;
   TrState* state;
   double* x;         // parameter vectors, concatenated in batch order
   double* vec;       // per frame: p, g, t1, t2 (4 n doubles)
   double* fac;       // per frame: factor blocks A, C, W of H + lambda I (3 nD n_dim^2 doubles)
   double* Dtrial;
   const double* partials;
   int32_t* active_frames;
   int32_t* run_first;
   int32_t* run_count;
   int64_t* run_base;
   Counters* counters;
   TrParamsDev prm;
   int32_t n_frames;
   int32_t n_refs;    // references: 2-D before 3-D, by LDS need within a kind
   double* trace;     // null, or trace_cap records of trace_rec doubles per frame
   int32_t trace_cap;
   int32_t trace_rec; // kTraceFields + 3 * (largest nD n_dim of the batch)
   int32_t fold;      // 1: the update's last workgroup schedules the next round (one reference)
   // 1: frames of references with a large eigendecomposition (RefDev::eig == 2) leave their
   // subproblem to the batched eigen-space kernels launched after the update (k_eig_gemm,
   // k_eig_secular); 0: they solve it by cyclic reduction in the update
   int32_t split_eig;
   // The launches of those kernels: eig_tiles holds n_eig_tiles product tiles of 4 ints
   // (reference, first row, first entry of its active list, k group), eig_frames the
   // n_eig_frames frames (or slots) that may hold such a reference's frame; null tables: the
   // tiles of reference 0 for one frame, and frames 0 .. n_eig_frames - 1.
   int32_t n_eig_tiles, n_eig_frames;
   const int32_t* eig_tiles;
   const int32_t* eig_frames;
};

// Trust-region states of a frame whose subproblem runs in the batched eigen-space kernels:
// left by the update (qv = Q^T g to be formed), by the secular iteration with w = qv / (ev + lam)
// to be turned into p = -Q w, or with p already zero (no solution). The next update takes the
// last two as 1 (a step was proposed) and sets 1 when the frame stops.
constexpr int kPhaseEigPending = 2, kPhaseEigW = 3, kPhaseEigZero = 4;

}  // namespace fa

// ---- host objects behind the C handles ----------------------------------------------------

struct fa_context {
   int device = 0;
   int n_cu = 256;
   int lds_per_block = 65536;  // bytes of LDS one workgroup may allocate
   int lds_config = 0;         // hipError_t of raising the kernels' dynamic-LDS limit
   hipStream_t stream = nullptr;
   // Streams of the serving lanes, created with the context's own stream, before any other
   // stream of the context: the runtime maps streams onto a few hardware queues (4 by default)
   // in creation order, so these get queues of their own, which streams created and destroyed
   // later (batches, aligners) could otherwise share with them.
   static constexpr int kLaneStreams = 3;
   hipStream_t lane_stream[kLaneStreams] = {nullptr, nullptr, nullptr};
   std::mutex mu;
   int live_refs = 0;  // references not yet released; fa_destroy refuses while > 0
   int live_servers = 0;  // servers using the lane streams
};

struct fa_reference {
   fa_context* ctx = nullptr;
   int nz = 0, ny = 0, nx = 0, nD = 0, n_dim = 0, chunk_px = 0;
   int item_rows = 1;  // most frame rows one work item spans
   int item_w = 0;     // columns of a work item: nx, or kStripWidth for strip items
   int64_t n_px = 0;
   fa::ScanDev scan{};
   double scan_host[6] = {0};
   std::vector<int32_t> D_range;  // nD x 2
   std::vector<fa::Chunk> chunks;
   std::vector<int32_t> ichunk;
   std::vector<double> Hdiag, Hsub;
   int h_pd = 0;
   double eig_lo = 0;
   // device
   float4* d_rg4 = nullptr;
   float* d_df = nullptr;
   fa::Chunk* d_chunks = nullptr;
   int32_t* d_ichunk = nullptr;
   double* d_Hdiag = nullptr;
   double* d_Hsub = nullptr;
   double* d_L0 = nullptr;  // 3 nD n_dim^2: factor blocks A, C, W of H
   double* d_Q = nullptr;   // eigenvectors and eigenvalues of H (nD n_dim <= kEigBatchMax)
   double* d_ev = nullptr;
   fa::RefDev dev{};
   int live_frames = 0;
   std::vector<fa_frame*> frames;  // registered frames (fa_clear_temp drops their cached batches)
   bool counted = false;  // included in ctx->live_refs
};

struct fa_frame {
   fa_reference* ref = nullptr;
   int dtype = FA_F32;
   float offset = 0;
   void* d_data = nullptr;
   bool own_data = true;        // false: d_data belongs to the caller of adopt_device_frame
   fa_batch* single = nullptr;  // cached one-frame batch for fa_align and single evaluations
   // Serialises the calls on this frame (fa_align, the single evaluations and warps, which use
   // the frame's own batch and stream); calls on different frames run concurrently.
   std::mutex mu;
};

namespace fa {
// Sets the calling thread's fa_last_error() message; returns code.
int set_error(int code, const std::string& msg);
// The C entry points' handler for a C++ exception (host memory or threads exhausted), which
// must not leave the C interface: sets the message without throwing and returns FA_E_OOM.
int fa_caught() noexcept;
int adopt_device_frame(fa_reference* r, void* d_data, int32_t dtype, float offset, fa_frame** out, bool own = true);
// Trust-region parameters of the next runs of batch b (and of its parts).
void batch_set_params(fa_batch* b, const fa_tr_params& p);
}  // namespace fa
