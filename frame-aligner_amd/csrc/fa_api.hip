// C ABI of the gfx950 warp-registration library (include/frame_aligner.h).
//
// Host side of the drop-in boundary: validates arguments, owns device memory behind opaque
// handles, builds the per-reference tables (node intervals, work chunks) and drives the
// kernels of fa_kernels.hip on the context's stream. Exceptions never cross the ABI.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include <dlfcn.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "fa_kernels.hpp"

using namespace fa;

namespace {

thread_local std::string g_last_error;

int set_err(int code, const std::string& msg)
{
   g_last_error = msg;
   return code;
}

#define HIPCHK(expr)                                                                              \
   do {                                                                                           \
      hipError_t e_ = (expr);                                                                     \
      if (e_ != hipSuccess)                                                                       \
         return set_err(FA_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));          \
   } while (0)

#define CHK(expr)                                                                                 \
   do {                                                                                           \
      int r_ = (expr);                                                                            \
      if (r_ != FA_OK) return r_;                                                                 \
   } while (0)

template <typename T>
int dalloc(T** p, size_t count)
{
   *p = nullptr;
   if (count == 0) count = 1;
   if (hipMalloc((void**)p, count * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      *p = nullptr;
      return set_err(FA_E_OOM, "Not enough memory to allocate");
   }
   return FA_OK;
}

template <typename T>
void dfree(T*& p)
{
   if (p) (void)hipFree((void*)p);
   p = nullptr;
}

constexpr int kLdsCapFloats = 15360;  // 60 KB of staged frame rows per workgroup at most
constexpr int kBlocksPerCu = 64;  // pass grid cap per compute unit: about one workgroup per work item

// ImageScanParameters ctor (ImageScanParameters.h:7-20) then processScanParameters
// (FrameWarper.cpp:223-236).
void derive_scan(const fa_scan_params& in, int nz, int ny, int nx, int nD, double* o6, ScanDev* sd)
{
   int sx = in.n_x > 0 ? in.n_x : nx;
   int sy = in.n_y > 0 ? in.n_y : ny;
   int sz = in.n_z > 0 ? in.n_z : nz;
   sx = std::max(sx, 1);
   sy = std::max(sy, 1);
   double line = in.line_duration, interline = in.interline_duration, interframe = in.interframe_duration;
   if (interline < line) interline = 1.01 * line;
   double pixel = line / sx;
   double frame = sy * interline;
   double stack = (sz - 1) * interframe + frame;
   double cap = nz * frame / (nD - nz);
   if ((interframe - frame) > cap) {
      interframe = frame + cap;
      stack = (nz - 1) * interframe + frame;
   }
   o6[0] = line;
   o6[1] = interline;
   o6[2] = pixel;
   o6[3] = frame;
   o6[4] = interframe;
   o6[5] = stack;
   sd->interframe = interframe;
   sd->interline = interline;
   sd->pixel = pixel;
   sd->stack = stack;
   sd->nDm1 = (double)(nD - 1);
}

struct DeviceGuard {
   explicit DeviceGuard(int dev) { (void)hipSetDevice(dev); }
};

}  // namespace

// ---- batch ----------------------------------------------------------------------------------

struct fa_batch {
   fa_context* ctx = nullptr;
   std::mutex mu;                  // serialises runs, evaluations and tuning of this batch
   hipStream_t stream = nullptr;   // the batch's own stream (default priority)
   bool own_stream = true;         // false: a stream of the context (serving lanes)
   // A batch of many frames runs as parts on their own streams, whose trust-region rounds
   // interleave so that one part's update overlaps another part's pass. Leaf batches have no
   // parts. part_first[i] is the caller's index of part i's first frame.
   std::vector<fa_batch*> parts;
   std::vector<int> part_first;
   bool finished = false;
   size_t timed = 0;
   // hipGraph of check_every trust-region rounds (schedule, pass, update), replayed per poll
   hipGraphExec_t gexec = nullptr;
   int graph_rounds = 0;
   fa_tr_params graph_prm{};
   std::vector<fa_frame*> frames;
   fa_tr_params prm{};
   std::vector<BatchFrame> hframes;
   std::vector<int64_t> x_off;  // offset of each frame's parameter vector in x_inout
   int64_t total_work = 0, total_x = 0, total_vec = 0, total_fac = 0, total_D = 0;
   // Launch classes of the fused pass: references of one kind with a similar LDS need (the
   // batch's references are ordered by kind, then need), each launched with its own LDS size.
   struct PassClass {
      int n_dim = 2, ref_lo = 0, ref_hi = 0;
      int64_t lds_first = 0, lds = 0;  // LDS floats of the class's first and largest reference
      int64_t work = 0;                // work items with every frame active
      int grid = 0;
   };
   std::vector<PassClass> pcls;
   // Launch classes of the trust-region update: LDS doubles of each (0: global memory) and
   // the class of each reference.
   std::vector<int> trc_lds;
   std::vector<int> trc_full;  // 1: the class stages everything in LDS
   std::vector<int> trc_threads;  // workgroup size of the class's update (64 or 256)
   std::vector<int> ref_trc;
   fa_tuning tune{};
   int timing = 0;
   std::vector<int> user;  // internal index -> caller's index
   int n_slots = 0;        // serving batches: number of slots (frames are admitted into them)
   RefDev* d_refs = nullptr;
   BatchFrame* d_frames = nullptr;
   TrState* d_state = nullptr;
   double* d_x = nullptr;
   double* d_vec = nullptr;
   double* d_fac = nullptr;
   double* d_D = nullptr;
   double* d_partials = nullptr;
   int32_t* d_active_frames = nullptr;
   int32_t* d_run_first = nullptr;
   int32_t* d_run_count = nullptr;
   int64_t* d_run_base = nullptr;
   int n_refs = 0;
   Counters* d_counters = nullptr;
   Counters* h_counters = nullptr;  // pinned
   std::vector<hipEvent_t> ev;      // start/end pairs
   float* d_warp = nullptr;         // output buffers of warps (one-frame batches)
   float* d_cov = nullptr;
   int64_t warp_px = 0;
   double* d_trace = nullptr;       // trust-region trace (fa_batch_set_trace), or null
   int trace_cap = 0, trace_rec = 0;
   // 1: one reference, one update launch: the update's last workgroup compacts the active
   // frames for the next round, so a round is a pass and an update (no schedule launch)
   int fold = 0;
   // tables of the batched eigen-space kernels (RefDev::eig == 2): n_eig_tiles product tiles
   // of 4 ints, then the n_eig_frames frames or slots that may hold such a reference's frame
   int32_t* d_eig = nullptr;
   int n_eig_tiles = 0, n_eig_frames = 0;

   BatchDev dev() const
   {
      BatchDev B;
      B.refs = d_refs;
      B.frames = d_frames;
      B.state = d_state;
      B.x = d_x;
      B.vec = d_vec;
      B.fac = d_fac;
      B.Dtrial = d_D;
      B.partials = d_partials;
      B.active_frames = d_active_frames;
      B.run_first = d_run_first;
      B.run_count = d_run_count;
      B.run_base = d_run_base;
      B.n_refs = n_refs;
      B.counters = d_counters;
      B.prm.radius = prm.radius;
      B.prm.min_delta = prm.min_delta;
      B.prm.rel_delta = prm.rel_delta;
      B.prm.sub_eps = prm.sub_eps;
      B.prm.max_iter = prm.max_iter;
      B.prm.max_evals = prm.max_evals;
      B.prm.sub_max_iter = prm.sub_max_iter;
      B.prm.skip_repeat = prm.skip_repeat;
      B.n_frames = n_slots > 0 ? n_slots : (int32_t)frames.size();
      B.trace = d_trace;
      B.trace_cap = trace_cap;
      B.trace_rec = trace_rec;
      B.fold = fold;
      B.split_eig = n_eig_tiles > 0 ? 1 : 0;
      B.n_eig_tiles = n_eig_tiles;
      B.n_eig_frames = n_eig_frames;
      B.eig_tiles = d_eig;
      B.eig_frames = d_eig ? d_eig + 4 * (int64_t)n_eig_tiles : nullptr;
      return B;
   }

   PassArgs pass_args() const
   {
      PassArgs a;
      a.refs = d_refs;
      a.frames = d_frames;
      a.Dtrial = d_D;
      a.partials = d_partials;
      a.active_frames = d_active_frames;
      a.run_first = d_run_first;
      a.run_count = d_run_count;
      a.run_base = d_run_base;
      a.n_refs = n_refs;
      a.counters = d_counters;
      a.lds_cap = tune.lds_floats;
      a.use_lds = tune.use_lds;
      a.ref_lo = 0;
      a.ref_hi = n_refs;
      return a;
   }

   void drop_graph()
   {
      if (gexec) (void)hipGraphExecDestroy(gexec);
      gexec = nullptr;
   }

   ~fa_batch()
   {
      for (fa_batch* q : parts) delete q;
      if (stream) (void)hipStreamSynchronize(stream);
      drop_graph();
      for (auto e : ev) (void)hipEventDestroy(e);
      dfree(d_refs);
      dfree(d_frames);
      dfree(d_state);
      dfree(d_x);
      dfree(d_vec);
      dfree(d_fac);
      dfree(d_D);
      dfree(d_partials);
      dfree(d_active_frames);
      dfree(d_run_first);
      dfree(d_run_count);
      dfree(d_run_base);
      dfree(d_counters);
      dfree(d_warp);
      dfree(d_cov);
      dfree(d_trace);
      dfree(d_eig);
      if (h_counters) (void)hipHostFree(h_counters);
      if (stream && own_stream) (void)hipStreamDestroy(stream);
   }
};

namespace {

// LDS a work item of reference r stages: the rows the longest item spans plus the row below
// for the taps and one for a displacement range that crosses a row boundary; in 3-D two
// planes (work items whose z range crosses a plane boundary need three and sample from
// global memory instead).
int64_t lds_floats_for(const fa_reference* r)
{
   const int64_t stride = ((r->item_w + 2 * kPad) + 3) & ~3;
   const int64_t rows = r->item_rows + 2;
   return rows * stride * (r->n_dim == 3 ? 2 : 1);
}

// Doubles of LDS the trust-region update of a frame of reference r uses: the cyclic-reduction
// factor blocks (3 nD n_dim^2) and the solve vector (nD n_dim); staging everything
// (tr_lds_full) adds the pass partials, H, x, p, g and the interval table.
int tr_lds_need(const fa_reference* r)
{
   return 3 * r->nD * r->n_dim * r->n_dim + r->nD * r->n_dim;
}

// Updates whose full staging fits this many doubles stage everything (several such waves
// still share a compute unit's LDS).
constexpr int kTrFullDoubles = 8192;

int tr_lds_full_base(const fa_reference* r)
{
   const int n = r->nD * r->n_dim;
   return tr_lds_need(r) + (int)r->chunks.size() * kPassStride + 2 * r->nD * r->n_dim * r->n_dim + 3 * n +
          (r->nD + 1) / 2;
}

// Q and the eigenvalues of the eigendecomposition path in LDS: n^2 and n doubles, each rounded
// up to 16 bytes, one double of alignment slack before them.
int tr_lds_q(const fa_reference* r)
{
   const int n = r->nD * r->n_dim;
   return ((n * n + 1) & ~1) + 1 + ((n + 1) & ~1);  // Q, then the eigenvalues
}

// Full staging of the update: tr_lds_full_base, plus Q when the eigendecomposition path's Q
// fits beside it within kTrFullDoubles (q_lds).
int tr_lds_full(const fa_reference* r)
{
   return tr_lds_full_base(r) + (r->dev.q_lds ? tr_lds_q(r) : 0);
}

// Frames of references with at least this many nodes run their update with 256 threads (the
// cyclic reduction's first levels then have a block per thread), others with one wave. The
// width depends on the reference alone, so a frame's result is the same in any batch.
constexpr int kWideUpdateNodes = 128;

int tr_threads(const fa_reference* r) { return r->nD >= kWideUpdateNodes ? 256 : 64; }

// LDS the update kernels declare statically (their workgroup reductions), kept free beside the
// dynamic LDS.
constexpr int kTrStaticLds = 1024;

// The update runs from LDS when the largest frame's need fits one workgroup's LDS, else from
// global memory (0).
int tr_lds_for(const fa_context* ctx, int need)
{
   return (int64_t)need * (int64_t)sizeof(double) + kTrStaticLds <= ctx->lds_per_block ? need : 0;
}

// Default LDS budget: the largest over the batch's frames, capped.
int default_lds_floats(const std::vector<fa_frame*>& frames)
{
   int64_t best = 0;
   for (fa_frame* f : frames) best = std::max(best, lds_floats_for(f->ref));
   return (int)std::min<int64_t>(best, kLdsCapFloats);
}

void default_tuning(fa_tuning* t, const std::vector<fa_frame*>& frames)
{
   t->lds_floats = default_lds_floats(frames);
   t->blocks_per_cu = kBlocksPerCu;
   t->use_lds = 1;
   t->check_every = 8;
   t->use_graph = 1;
   t->tr_lds = 1;
}

// Launch order of references: 2-D before 3-D, then by the LDS their work items stage.
void order_refs(std::vector<fa_reference*>& rs)
{
   std::stable_sort(rs.begin(), rs.end(), [](const fa_reference* a, const fa_reference* c) {
      if (a->n_dim != c->n_dim) return a->n_dim < c->n_dim;
      return lds_floats_for(a) < lds_floats_for(c);
   });
}

// Fused-pass workgroups whose LDS tiles fit one compute unit's LDS at once (at most 8).
// Pass workgroups of a class resident per compute unit by LDS: the staging tile, at least the
// work item's reduction buffer (1 + 2 n_dim doubles per thread, aliasing the tile).
int64_t resident_blocks(const fa_context* ctx, int64_t lds_floats, int n_dim)
{
   const int64_t bytes = std::max<int64_t>(lds_floats * (int64_t)sizeof(float),
                                           (int64_t)(1 + 2 * n_dim) * kBlock * (int64_t)sizeof(double));
   return std::min<int64_t>(8, (int64_t)ctx->lds_per_block / (bytes + 1024));
}

// Launch classes over the ordered references rs. Pass: a class ends at a change of kind or
// before a reference whose tile lets fewer workgroups share a compute unit than the class's
// first (the class launches with its largest tile, so a larger one would cost the others
// occupancy). Update: the references' factor sizes, a class starting above 2x its first.
// ref_items: each reference's work items with every frame of the batch active.
void make_classes(fa_batch* b, const std::vector<fa_reference*>& rs, const std::vector<int64_t>& ref_items,
                  int n_frames)
{
   b->pcls.clear();
   for (int i = 0; i < (int)rs.size(); i++) {
      const int64_t need = lds_floats_for(rs[i]);
      if (b->pcls.empty() || b->pcls.back().n_dim != rs[i]->n_dim ||
          resident_blocks(b->ctx, need, rs[i]->n_dim) <
              resident_blocks(b->ctx, b->pcls.back().lds_first, rs[i]->n_dim)) {
         fa_batch::PassClass c;
         c.n_dim = rs[i]->n_dim;
         c.ref_lo = i;
         c.lds_first = need;
         b->pcls.push_back(c);
      }
      fa_batch::PassClass& c = b->pcls.back();
      c.ref_hi = i + 1;
      c.lds = std::max(c.lds, need);
      c.work += ref_items[i];
   }
   // Update classes: one per workgroup width, split further by factor size (a class starting
   // above 2x its first) unless every frame's workgroup is resident at once even with the
   // largest LDS size; launch_updates then runs the classes in one launch.
   int most = 0, most_full = 0;
   for (fa_reference* r : rs) {
      most = std::max(most, tr_lds_need(r));
      most_full = std::max(most_full, tr_lds_full(r));
   }
   const int64_t per_cu = most ? b->ctx->lds_per_block / ((int64_t)most * (int64_t)sizeof(double)) : 1;
   const bool by_width_only =
      (int64_t)n_frames <= (int64_t)b->ctx->n_cu * std::max<int64_t>(per_cu, 1) || most_full <= kTrFullDoubles;
   std::vector<int> order(rs.size());
   for (size_t i = 0; i < rs.size(); i++) order[i] = (int)i;
   std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
      if (tr_threads(rs[a]) != tr_threads(rs[c])) return tr_threads(rs[a]) < tr_threads(rs[c]);
      return tr_lds_need(rs[a]) < tr_lds_need(rs[c]);
   });
   b->ref_trc.assign(rs.size(), 0);
   b->trc_lds.clear();
   b->trc_full.clear();
   b->trc_threads.clear();
   int first = 0, largest = 0, largest_full = 0;
   for (int i : order) {
      const int need = tr_lds_need(rs[i]);
      const int nt = tr_threads(rs[i]);
      if (b->trc_lds.empty() || nt != b->trc_threads.back() || (!by_width_only && need > 2 * first)) {
         b->trc_lds.push_back(0);
         b->trc_full.push_back(0);
         b->trc_threads.push_back(nt);
         first = need;
         largest = largest_full = 0;
      }
      largest = std::max(largest, need);
      largest_full = std::max(largest_full, tr_lds_full(rs[i]));
      const bool full = largest_full <= kTrFullDoubles;
      b->trc_lds.back() = tr_lds_for(b->ctx, full ? largest_full : largest);
      b->trc_full.back() = full && b->trc_lds.back() > 0 ? 1 : 0;
      b->ref_trc[i] = (int)b->trc_lds.size() - 1;
   }
}

// Grids of the pass classes: about one workgroup per work item, at most blocks_per_cu per
// compute unit.
void set_grids(fa_batch* b)
{
   const int64_t cap = (int64_t)b->ctx->n_cu * std::max(1, b->tune.blocks_per_cu);
   for (auto& c : b->pcls) c.grid = c.work ? (int)std::max<int64_t>(1, std::min(c.work, cap)) : 0;
}

// stream: a stream of the context to run on (the parts of a split batch), or null for a
// stream of the batch's own.
// Tables of the batched eigen-space kernels: the product tiles of each reference with
// RefDev::eig == 2 for frames_of[r] frames of it at most, and the frames (or slots) that may
// hold such a reference's frame; uploaded on the batch's stream. Nothing without such a
// reference.
int set_eig_tables(fa_batch* b, const std::vector<RefDev>& refs, const std::vector<int>& frames_of,
                   const std::vector<int32_t>& eig_frames)
{
   std::vector<int32_t> t;
   for (size_t r = 0; r < refs.size(); r++)
      if (refs[r].eig == 2 && frames_of[r] > 0) eig_tiles_of((int)r, refs[r].nD * refs[r].n_dim, frames_of[r], t);
   if (t.empty() || eig_frames.empty()) return FA_OK;
   const int tiles = (int)(t.size() / 4);
   t.insert(t.end(), eig_frames.begin(), eig_frames.end());
   int rc = dalloc(&b->d_eig, t.size());
   if (rc) return rc;
   if (hipMemcpyAsync(b->d_eig, t.data(), t.size() * sizeof(int32_t), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
       hipStreamSynchronize(b->stream) != hipSuccess) {
      (void)hipGetLastError();
      return set_err(FA_E_DEVICE, "batch upload failed");
   }
   b->n_eig_tiles = tiles;
   b->n_eig_frames = (int)eig_frames.size();
   return FA_OK;
}

int batch_build(fa_frame* const* frames, int32_t n, const fa_tr_params* prm, fa_batch** out,
                hipStream_t stream = nullptr)
{
   *out = nullptr;
   if (n <= 0 || !frames) return set_err(FA_E_ARG, "empty batch");
   fa_context* ctx = nullptr;
   for (int i = 0; i < n; i++) {
      if (!frames[i] || !frames[i]->ref) return set_err(FA_E_UNKNOWN_FRAME, "Unrecognised frame");
      if (!ctx) ctx = frames[i]->ref->ctx;
      if (frames[i]->ref->ctx != ctx) return set_err(FA_E_ARG, "frames of different contexts");
   }
   auto* b = new fa_batch();
   b->ctx = ctx;
   b->frames.assign(frames, frames + n);
   if (prm) b->prm = *prm;
   else fa_tr_default(&b->prm);
   default_tuning(&b->tune, b->frames);

   // caller's layout of x and D
   b->x_off.resize(n);
   std::vector<int64_t> d_off(n);
   for (int i = 0; i < n; i++) {
      fa_reference* r = frames[i]->ref;
      b->x_off[i] = b->total_x;
      d_off[i] = b->total_D;
      b->total_x += (int64_t)r->nD * r->n_dim;
      b->total_D += 3 * (int64_t)r->nD;
   }
   // internal order: 2-D frames first (k_pass<2> takes them), caller's order within a kind
   for (int i = 0; i < n; i++)
      if (frames[i]->ref->n_dim == 2) b->user.push_back(i);
   for (int i = 0; i < n; i++)
      if (frames[i]->ref->n_dim == 3) b->user.push_back(i);

   std::vector<fa_reference*> rs;
   for (int j = 0; j < n; j++) {
      fa_reference* r = frames[b->user[j]]->ref;
      if (std::find(rs.begin(), rs.end(), r) == rs.end()) rs.push_back(r);
   }
   if ((int)rs.size() > kMaxRefs) {
      delete b;
      return set_err(FA_E_ARG, "too many distinct references in one batch (at most 4096)");
   }
   order_refs(rs);
   std::map<fa_reference*, int> ref_idx;
   std::vector<RefDev> refs;
   for (size_t ri = 0; ri < rs.size(); ri++) {
      ref_idx[rs[ri]] = (int)ri;
      refs.push_back(rs[ri]->dev);
   }
   std::vector<int64_t> ref_items(rs.size(), 0);
   for (int j = 0; j < n; j++) ref_items[ref_idx[frames[b->user[j]]->ref]] += (int64_t)frames[b->user[j]]->ref->chunks.size();
   make_classes(b, rs, ref_items, n);
   b->hframes.resize(n);
   for (int j = 0; j < n; j++) {
      const int i = b->user[j];
      fa_reference* r = frames[i]->ref;
      const int ri = ref_idx[r];
      BatchFrame& bf = b->hframes[j];
      bf.f.data = frames[i]->d_data;
      bf.f.dtype = frames[i]->dtype;
      bf.f.offset = frames[i]->offset;
      bf.ref = ri;
      bf.r = r->dev;
      bf.n = r->nD * r->n_dim;
      bf.work_base = b->total_work;
      bf.x_base = b->x_off[i];
      bf.vec_base = b->total_vec;
      bf.fac_base = b->total_fac;
      bf.d_base = d_off[i];
      bf.trc = b->ref_trc[ri];
      b->total_work += (int64_t)r->chunks.size();
      b->total_vec += 3 * (int64_t)bf.n;
      b->total_fac += 3 * (int64_t)r->nD * r->n_dim * r->n_dim;
   }
   b->n_refs = (int)refs.size();
   int rc = FA_OK;
   auto fail = [&](int code) {
      delete b;
      return code;
   };
   if ((rc = dalloc(&b->d_refs, refs.size())) || (rc = dalloc(&b->d_frames, n)) ||
       (rc = dalloc(&b->d_state, n)) || (rc = dalloc(&b->d_x, b->total_x)) ||
       (rc = dalloc(&b->d_vec, b->total_vec)) || (rc = dalloc(&b->d_fac, b->total_fac)) ||
       (rc = dalloc(&b->d_D, b->total_D)) ||
       (rc = dalloc(&b->d_partials, (size_t)b->total_work * kPassStride)) ||
       (rc = dalloc(&b->d_active_frames, n)) || (rc = dalloc(&b->d_run_first, refs.size())) ||
       (rc = dalloc(&b->d_run_count, refs.size())) || (rc = dalloc(&b->d_run_base, refs.size() + 1)) ||
       (rc = dalloc(&b->d_counters, 1)))
      return fail(rc);
   if (hipHostMalloc((void**)&b->h_counters, sizeof(Counters), hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      b->h_counters = nullptr;
      return fail(set_err(FA_E_OOM, "Not enough memory to allocate"));
   }
   if (stream) {
      b->stream = stream;
      b->own_stream = false;
   } else if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      b->stream = nullptr;
      return fail(set_err(FA_E_DEVICE, "stream creation failed"));
   }
   hipStream_t s = b->stream;
   if (hipMemcpyAsync(b->d_refs, refs.data(), refs.size() * sizeof(RefDev), hipMemcpyHostToDevice, s) !=
           hipSuccess ||
       hipMemcpyAsync(b->d_frames, b->hframes.data(), n * sizeof(BatchFrame), hipMemcpyHostToDevice, s) !=
           hipSuccess ||
       hipMemsetAsync(b->d_state, 0, n * sizeof(TrState), s) != hipSuccess ||
       hipMemsetAsync(b->d_counters, 0, sizeof(Counters), s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess) {
      (void)hipGetLastError();
      return fail(set_err(FA_E_DEVICE, "batch upload failed"));
   }
   set_grids(b);
   b->fold = (rs.size() == 1 && b->trc_lds.size() == 1) ? 1 : 0;
   {
      std::vector<int> frames_of(refs.size(), 0);
      std::vector<int32_t> eig_frames;
      for (int i = 0; i < n; i++) {
         frames_of[b->hframes[i].ref]++;
         if (refs[b->hframes[i].ref].eig == 2) eig_frames.push_back(i);
      }
      if ((rc = set_eig_tables(b, refs, frames_of, eig_frames))) return fail(rc);
   }
   *out = b;
   return FA_OK;
}

// Enqueues the fused pass over the active frames of every kind.
// Each kind's pass reserves the LDS its own work items stage (within the tuned cap), so
// 2-D frames keep their occupancy beside 3-D frames.
int launch_passes(fa_batch* b, hipStream_t s)
{
   PassArgs pa = b->pass_args();
   for (const auto& c : b->pcls) {
      if (!c.grid) continue;
      pa.ref_lo = c.ref_lo;
      pa.ref_hi = c.ref_hi;
      pa.lds_cap = (int)std::min<int64_t>(b->tune.lds_floats, c.lds);
      HIPCHK(launch_pass(pa, c.n_dim, c.grid, s));
   }
   return FA_OK;
}

// Enqueues the trust-region update: one launch per update class or, when there are several
// and every frame's workgroup is resident at once even with the largest class's LDS size, one
// launch of all of them (the update is latency-bound, so class launches one after another
// would add their latencies).
int launch_updates(fa_batch* b, hipStream_t s)
{
   const BatchDev B = b->dev();
   const size_t nc = b->trc_lds.size();
   if (nc > 1 && nc <= 32) {
      int most = 0;
      bool fit = true;
      uint32_t wide = 0, full = 0;
      for (size_t c = 0; c < nc; c++) {
         most = std::max(most, b->trc_lds[c]);
         fit = fit && b->trc_lds[c] > 0;
         if (b->trc_threads[c] == 256) wide |= 1u << c;
         if (b->trc_full[c]) full |= 1u << c;
      }
      const int lds = b->tune.tr_lds ? most : 0;
      if ((fit || !b->tune.tr_lds) && (int64_t)B.n_frames <= (int64_t)b->ctx->n_cu * tr_update_all_per_cu(lds)) {
         HIPCHK(launch_tr_update_all(B, lds, wide, lds ? full : 0u, s));
         HIPCHK(launch_eig_subproblems(B, s));
         return FA_OK;
      }
   }
   for (size_t c = 0; c < nc; c++)
      HIPCHK(launch_tr_update(B, (int)c, b->tune.tr_lds ? b->trc_lds[c] : 0, b->trc_full[c], b->trc_threads[c], s));
   HIPCHK(launch_eig_subproblems(B, s));
   return FA_OK;
}

int event_pair(fa_batch* b, size_t i, hipEvent_t* s, hipEvent_t* e)
{
   while (b->ev.size() < 2 * (i + 1)) {
      hipEvent_t x;
      HIPCHK(hipEventCreate(&x));
      b->ev.push_back(x);
   }
   *s = b->ev[2 * i];
   *e = b->ev[2 * i + 1];
   return FA_OK;
}

constexpr size_t kMaxTimedRounds = 4096;
constexpr int kServerSplitSlots = 16;  // a one-lane server with at least this many slots runs kServerParts
constexpr int kServerParts = 2;        // slot batches (at least 8 slots each)
constexpr int kMaxLanes = fa_context::kLaneStreams;  // server lanes (each on a lane stream of the context)
constexpr int64_t kLaneSpan = 16;      // a lane's references span at most this ratio of pixels
constexpr size_t kArenaMargin = (size_t)32 << 20;  // device memory a default arena leaves free
// Batches of at least kSplitFrames frames run as two parts, of at least kSplit3Frames as three
// (one per lane stream of the context): 200 frames of 512^2 run 7 % faster in three parts than
// in two, a fourth part (sharing a hardware queue) 28 % slower.
constexpr int kSplitFrames = 64, kSplit3Frames = 144;

int64_t max_rounds(const fa_batch* b) { return b->prm.max_evals > 0 ? (int64_t)b->prm.max_evals + 2 : 100000; }

// Phases of a run on a leaf batch, all enqueued on its stream.
int run_begin(fa_batch* b, const double* x)
{
   hipStream_t s = b->stream;
   HIPCHK(hipMemcpyAsync(b->d_x, x, b->total_x * sizeof(double), hipMemcpyHostToDevice, s));
   HIPCHK(hipMemsetAsync(b->d_counters, 0, sizeof(Counters), s));
   if (b->d_trace)
      HIPCHK(hipMemsetAsync(b->d_trace, 0, (size_t)b->frames.size() * b->trace_cap * b->trace_rec * sizeof(double), s));
   HIPCHK(launch_tr_init(b->dev(), 0, s));
   // a folded schedule runs at the end of each update: the first round's runs here
   if (b->fold) HIPCHK(launch_schedule(b->dev(), 1, s));
   b->finished = false;
   b->timed = 0;
   return FA_OK;
}

int run_round(fa_batch* b)
{
   hipStream_t s = b->stream;
   const BatchDev B = b->dev();
   if (!b->fold) HIPCHK(launch_schedule(B, 1, s));
   hipEvent_t es = nullptr, ee = nullptr;
   if (b->timing && b->timed < kMaxTimedRounds) {
      CHK(event_pair(b, b->timed, &es, &ee));
      HIPCHK(hipEventRecord(es, s));
   }
   CHK(launch_passes(b, s));
   if (es) {
      HIPCHK(hipEventRecord(ee, s));
      b->timed++;
   }
   CHK(launch_updates(b, s));
   return FA_OK;
}

// Refreshes the counters the host polls and copies them back (asynchronously).
int run_poll(fa_batch* b)
{
   // with the folded schedule the counters are current after every update
   if (!b->fold) HIPCHK(launch_schedule(b->dev(), 0, b->stream));
   HIPCHK(hipMemcpyAsync(b->h_counters, b->d_counters, sizeof(Counters), hipMemcpyDeviceToHost, b->stream));
   return FA_OK;
}

int run_finish(fa_batch* b, double* x, fa_align_stats* stats, int first_user)
{
   const int n = (int)b->frames.size();
   std::vector<TrState> st(n);
   HIPCHK(hipMemcpyAsync(x, b->d_x, b->total_x * sizeof(double), hipMemcpyDeviceToHost, b->stream));
   HIPCHK(hipMemcpyAsync(st.data(), b->d_state, n * sizeof(TrState), hipMemcpyDeviceToHost, b->stream));
   HIPCHK(hipStreamSynchronize(b->stream));
   if (stats)
      for (int j = 0; j < n; j++) {
         fa_align_stats& o = stats[first_user + b->user[j]];
         o.iterations = st[j].iterations;
         o.evaluations = st[j].evaluations;
         o.rejected = st[j].rejected;
         o.status = st[j].done ? st[j].status : FA_STOP_MAX_EVALS;
         o.f_final = st[j].f;
         o.radius_final = st[j].radius;
      }
   return FA_OK;
}

int add_timing(fa_batch* b, fa_batch_timing* t)
{
   double pass_ms = 0;
   for (size_t i = 0; i < b->timed; i++) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, b->ev[2 * i], b->ev[2 * i + 1]));
      pass_ms += ms;
   }
   t->pass_ms += pass_ms;
   t->pass_launches += (int64_t)b->timed;
   t->pass_bytes += (double)b->h_counters->bytes;
   t->pass_px += (double)b->h_counters->px;
   t->tr_rounds = std::max<int64_t>(t->tr_rounds, (int64_t)b->h_counters->rounds);
   return FA_OK;
}

// Captures check_every rounds into a graph (once per tuning / parameter set).
int ensure_graph(fa_batch* b, int K)
{
   if (b->gexec && b->graph_rounds == K && std::memcmp(&b->graph_prm, &b->prm, sizeof(fa_tr_params)) == 0)
      return FA_OK;
   b->drop_graph();
   // captured on a stream of its own: the batch's stream may be a lane stream of the context,
   // on which another thread could enqueue work during the capture
   hipStream_t s = nullptr;
   HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
   struct Drop {
      hipStream_t s;
      ~Drop() { (void)hipStreamDestroy(s); }
   } drop{s};
   const BatchDev B = b->dev();
   HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
   int rc = FA_OK;
   for (int r = 0; r < K && rc == FA_OK; r++) {
      if (!b->fold && launch_schedule(B, 1, s) != hipSuccess) rc = set_err(FA_E_DEVICE, "graph capture failed");
      if (rc == FA_OK) rc = launch_passes(b, s);
      if (rc == FA_OK) rc = launch_updates(b, s);
   }
   hipGraph_t graph = nullptr;
   hipError_t e = hipStreamEndCapture(s, &graph);
   if (rc != FA_OK || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      (void)hipGetLastError();
      return rc != FA_OK ? rc : set_err(FA_E_DEVICE, "graph capture failed");
   }
   e = hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0);
   (void)hipGraphDestroy(graph);
   if (e != hipSuccess) {
      b->gexec = nullptr;
      (void)hipGetLastError();
      return set_err(FA_E_DEVICE, "graph instantiation failed");
   }
   b->graph_rounds = K;
   b->graph_prm = b->prm;
   return FA_OK;
}

// Runs every leaf (a leaf batch alone, or the parts of a split batch) to completion,
// interleaving their rounds: rounds are enqueued part by part on their own streams and the
// host polls every check_every rounds.
int batch_run(fa_batch* b, double* x_inout, fa_align_stats* stats, fa_batch_timing* timing)
{
   auto t0 = std::chrono::steady_clock::now();
   std::vector<fa_batch*> leaves;
   std::vector<int64_t> xoff;
   std::vector<int> first;
   if (b->parts.empty()) {
      leaves.push_back(b);
      xoff.push_back(0);
      first.push_back(0);
   } else {
      int64_t o = 0;
      for (size_t i = 0; i < b->parts.size(); i++) {
         leaves.push_back(b->parts[i]);
         xoff.push_back(o);
         first.push_back(b->part_first[i]);
         o += b->parts[i]->total_x;
      }
   }
   const int K = std::max(1, b->tune.check_every);
   // per-pass HIP event timing needs eager launches: events recorded inside a captured graph
   // cannot be read back with hipEventElapsedTime
   const bool graphs = b->tune.use_graph != 0 && !b->timing;
   if (graphs)
      for (fa_batch* q : leaves) CHK(ensure_graph(q, K));
   for (size_t i = 0; i < leaves.size(); i++) CHK(run_begin(leaves[i], x_inout + xoff[i]));
   const int64_t limit = max_rounds(b);
   // runs the leaves of `set` to completion, their rounds interleaved
   auto run_set = [&](const std::vector<fa_batch*>& set) -> int {
      int64_t round = 0;
      for (;;) {
         bool any = false;
         if (graphs) {
            for (fa_batch* q : set)
               if (!q->finished) {
                  HIPCHK(hipGraphLaunch(q->gexec, q->stream));
                  any = true;
               }
            round += K;
         } else {
            for (int r = 0; r < K && round < limit; r++, round++)
               for (fa_batch* q : set)
                  if (!q->finished) {
                     CHK(run_round(q));
                     any = true;
                  }
         }
         if (!any) break;
         for (fa_batch* q : set)
            if (!q->finished) CHK(run_poll(q));
         bool all = true;
         for (fa_batch* q : set)
            if (!q->finished) {
               HIPCHK(hipStreamSynchronize(q->stream));
               q->finished = q->h_counters->n_active_frames == 0;
               all = all && q->finished;
            }
         if (all || round >= limit) break;
      }
      return FA_OK;
   };
   if (b->timing) {
      // timed runs take the parts one after the other, so each pass's events bracket that
      // pass alone rather than time it shares with the other part's passes
      for (fa_batch* q : leaves) CHK(run_set(std::vector<fa_batch*>{q}));
   } else {
      CHK(run_set(leaves));
   }
   for (size_t i = 0; i < leaves.size(); i++) CHK(run_finish(leaves[i], x_inout + xoff[i], stats, first[i]));
   if (timing) {
      *timing = fa_batch_timing{};
      for (fa_batch* q : leaves) CHK(add_timing(q, timing));
      timing->wall_ms =
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
   }
   return FA_OK;
}

// One fused pass over every frame of a leaf batch at D (concatenated nD x 3 per frame).
// One evaluation of a leaf at D, in two halves: the launches (nothing waits), then the
// copies of the states (and Jtr) and their delivery, which wait for that leaf's stream.
int leaf_eval_launch(fa_batch* b, const double* D)
{
   hipStream_t s = b->stream;
   BatchDev B = b->dev();
   HIPCHK(hipMemcpyAsync(b->d_D, D, b->total_D * sizeof(double), hipMemcpyHostToDevice, s));
   HIPCHK(launch_tr_init(B, 1, s));
   HIPCHK(launch_schedule(B, 1, s));
   CHK(launch_passes(b, s));
   HIPCHK(launch_reduce_eval(B, s));
   return FA_OK;
}

int leaf_eval_collect(fa_batch* b, double* err, double* jac)
{
   hipStream_t s = b->stream;
   const int n = (int)b->frames.size();
   std::vector<TrState> st(n);
   std::vector<double> vec;
   HIPCHK(hipMemcpyAsync(st.data(), b->d_state, n * sizeof(TrState), hipMemcpyDeviceToHost, s));
   if (jac) {
      vec.resize(b->total_vec);
      HIPCHK(hipMemcpyAsync(vec.data(), b->d_vec, b->total_vec * sizeof(double), hipMemcpyDeviceToHost, s));
   }
   HIPCHK(hipStreamSynchronize(s));
   for (int j = 0; j < n; j++) {
      const int i = b->user[j];
      if (err) err[i] = st[j].f;
      if (jac) {
         const BatchFrame& bf = b->hframes[j];
         std::memcpy(jac + b->x_off[i], vec.data() + bf.vec_base + bf.n, bf.n * sizeof(double));
      }
   }
   return FA_OK;
}

int leaf_eval(fa_batch* b, const double* D, double* err, double* jac)
{
   CHK(leaf_eval_launch(b, D));
   return leaf_eval_collect(b, err, jac);
}

// A split batch launches every part's evaluation before collecting any, so the parts' passes
// run together on their streams.
int batch_eval(fa_batch* b, const double* D, double* err, double* jac)
{
   if (b->parts.empty()) return leaf_eval(b, D, err, jac);
   int64_t dof = 0;
   for (fa_batch* q : b->parts) {
      CHK(leaf_eval_launch(q, D + dof));
      dof += q->total_D;
   }
   int64_t xo = 0;
   for (size_t i = 0; i < b->parts.size(); i++) {
      fa_batch* q = b->parts[i];
      CHK(leaf_eval_collect(q, err ? err + b->part_first[i] : nullptr, jac ? jac + xo : nullptr));
      xo += q->total_x;
   }
   return FA_OK;
}

// A batch for the caller: a leaf, or two or three leaves over consecutive runs of the frames
// (balanced by pixel count) when it is large enough to overlap their rounds. Each part runs on
// its own lane stream of the context, which has a hardware queue of its own, so the parts'
// rounds interleave on the device; a frame's result does not depend on the partition.
int batch_create_split(fa_frame* const* frames, int32_t n, const fa_tr_params* prm, fa_batch** out)
{
   if (n < kSplitFrames) return batch_build(frames, n, prm, out);
   static_assert(fa_context::kLaneStreams >= 3, "a part per lane stream");
   const int P = n >= kSplit3Frames ? 3 : 2;
   int64_t total = 0;
   for (int i = 0; i < n; i++) total += frames[i]->ref->n_px;
   // part p ends after the frames whose pixels reach p / P of the total (at least one frame
   // per part)
   std::vector<int> cut = {0};
   int64_t acc = 0;
   int i = 0;
   for (int p = 1; p < P; p++) {
      while (i < n - (P - p) && P * (acc + frames[i]->ref->n_px) <= p * total) acc += frames[i++]->ref->n_px;
      if (i <= cut.back()) acc += frames[i++]->ref->n_px;
      cut.push_back(i);
   }
   cut.push_back(n);
   fa_context* ctx = frames[0]->ref->ctx;
   auto* b = new fa_batch();
   b->ctx = ctx;
   b->frames.assign(frames, frames + n);
   for (int p = 0; p < P; p++) {
      fa_batch* q = nullptr;
      const int rc = batch_build(frames + cut[p], cut[p + 1] - cut[p], prm, &q, ctx->lane_stream[p]);
      if (rc != FA_OK) {
         delete b;
         return rc;
      }
      b->parts.push_back(q);
      b->part_first.push_back(cut[p]);
      b->total_x += q->total_x;
      b->total_D += q->total_D;
   }
   b->prm = b->parts[0]->prm;
   b->tune = b->parts[0]->tune;
   *out = b;
   return FA_OK;
}

// Largest parameter count nD * n_dim of the batch's frames.
int max_params(const fa_batch* b)
{
   int m = 0;
   for (fa_frame* f : b->frames) m = std::max(m, f->ref->nD * f->ref->n_dim);
   return m;
}

int leaf_set_trace(fa_batch* b, int cap, int rec)
{
   HIPCHK(hipStreamSynchronize(b->stream));
   dfree(b->d_trace);
   b->trace_cap = 0;
   b->trace_rec = 0;
   b->drop_graph();  // graphs carry the trace pointer
   if (cap <= 0) return FA_OK;
   CHK(dalloc(&b->d_trace, b->frames.size() * (size_t)cap * rec));
   HIPCHK(hipMemset(b->d_trace, 0, b->frames.size() * (size_t)cap * rec * sizeof(double)));
   b->trace_cap = cap;
   b->trace_rec = rec;
   return FA_OK;
}

int leaf_get_trace(fa_batch* b, double* out, int first_user, int cap, int rec)
{
   const size_t per = (size_t)cap * rec;
   if (!b->d_trace || b->trace_cap != cap || b->trace_rec != rec) return set_err(FA_E_STATE, "tracing is off");
   std::vector<double> h(b->frames.size() * per);
   HIPCHK(hipMemcpyAsync(h.data(), b->d_trace, h.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
   HIPCHK(hipStreamSynchronize(b->stream));
   for (size_t j = 0; j < b->frames.size(); j++)
      std::memcpy(out + (size_t)(first_user + b->user[j]) * per, h.data() + j * per, per * sizeof(double));
   return FA_OK;
}

int single_batch(fa_frame* f, fa_batch** out)
{
   if (!f->single) CHK(batch_build(&f, 1, nullptr, &f->single));
   *out = f->single;
   return FA_OK;
}

}  // namespace

// ---- serving: continuous batching over fixed slots ----------------------------------------

struct fa_server {
   // A request: a registered frame (f), or host data uploaded into the server's frame arena.
   struct Req {
      fa_frame* f = nullptr;
      fa_reference* ref = nullptr;
      const void* host = nullptr;
      int32_t dtype = FA_F32;
      float offset = 0.f;
      double* x = nullptr;
      int64_t tag = 0;
      int64_t seq = 0;  // arrival order
   };
   enum : char { kFree = 0, kUploading = 1, kRunning = 2 };
   fa_context* ctx = nullptr;
   // Lanes: the references grouped by kind and size. Each lane has its own slot batches
   // (parts) on their own streams, launched and harvested on their own cadence, so a lane of
   // small frames runs its rounds without waiting for the rounds of a lane of large ones. A
   // server with one lane and 16 or more slots runs it as two parts whose rounds interleave
   // (one part's update overlaps the other's pass). Global slot j is local slot j - first[p]
   // of part p, whose frames use the staging entries (h_bf, h_x, h_slots) from
   // first[p] on.
   struct Lane {
      std::vector<fa_reference*> refs;  // in the slot batches' launch order
      int part_lo = 0, part_hi = 0;
      std::deque<Req> pending;          // arrival order
   };
   std::vector<Lane> lanes;
   std::vector<fa_reference*> refs;  // every served reference
   std::vector<int> ref_lane;        // lane of refs[i]
   std::vector<fa_batch*> parts;
   std::vector<int> first;
   std::vector<int> part_lane;
   std::vector<char> launched;       // per part: rounds launched, results not yet harvested
   std::vector<hipEvent_t> done_ev;  // per part: recorded after its launched rounds and copies
   int S = 0;         // slots over all parts
   int cap = 0;       // frames admitted at once (uploading or running), over all lanes
   int admitted = 0;
   int nmax = 0;
   int64_t next_seq = 0;
   int part_of(int j) const
   {
      int p = 0;
      while (p + 1 < (int)first.size() && j >= first[p + 1]) p++;
      return p;
   }
   int lane_of(const fa_reference* r) const
   {
      return ref_lane[std::find(refs.begin(), refs.end(), r) - refs.begin()];
   }
   std::vector<std::vector<int>> free_slots;  // per part
   // the lane's part with the most free slots (frames spread evenly over a lane's parts, so
   // both streams carry work), or -1
   int free_part(const Lane& L) const
   {
      int best = -1;
      for (int p = L.part_lo; p < L.part_hi; p++)
         if (!free_slots[p].empty() && (best < 0 || free_slots[p].size() > free_slots[best].size())) best = p;
      return best;
   }
   size_t queued() const
   {
      size_t n = 0;
      for (const Lane& L : lanes) n += L.pending.size();
      return n;
   }
   std::vector<Req> slot_req;
   std::vector<char> sstate;
   std::vector<size_t> blk_off, blk_size;  // arena block of each slot (size 0: none)
   std::vector<hipEvent_t> up_ev;          // upload of each slot's frame done
   std::deque<fa_server_result> results;
   // frame arena: first-fit over free blocks (offset -> size), neighbours merged on release
   uint8_t* d_arena = nullptr;
   size_t arena_bytes = 0, arena_used = 0;
   std::map<size_t, size_t> free_blocks;
   hipStream_t cstream = nullptr;  // uploads (default priority)
   // pinned host staging, and the device's addresses of the same memory (the slot-start and
   // publish kernels read and write it over the host link)
   BatchFrame* h_bf = nullptr;
   double* h_x = nullptr;
   int32_t* h_slots = nullptr;
   TrState* h_state = nullptr;
   double* h_xall = nullptr;
   BatchFrame* dh_bf = nullptr;
   double* dh_x = nullptr;
   int32_t* dh_slots = nullptr;
   TrState* dh_state = nullptr;
   double* dh_xall = nullptr;

   static size_t round_up(size_t n) { return (n + 255) & ~(size_t)255; }

   bool arena_alloc(size_t n, size_t* off)
   {
      n = round_up(n);
      for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
         if (it->second < n) continue;
         *off = it->first;
         const size_t rest = it->second - n;
         free_blocks.erase(it);
         if (rest) free_blocks[*off + n] = rest;
         arena_used += n;
         return true;
      }
      return false;
   }

   void arena_free(size_t off, size_t n)
   {
      n = round_up(n);
      arena_used -= n;
      auto it = free_blocks.emplace(off, n).first;
      auto nx = std::next(it);
      if (nx != free_blocks.end() && it->first + it->second == nx->first) {
         it->second += nx->second;
         free_blocks.erase(nx);
      }
      if (it != free_blocks.begin()) {
         auto pv = std::prev(it);
         if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_blocks.erase(it);
         }
      }
   }

   ~fa_server()
   {
      if (cstream) (void)hipStreamSynchronize(cstream);
      for (fa_batch* q : parts) {
         if (q->stream) (void)hipStreamSynchronize(q->stream);
         delete q;
      }
      for (auto e : up_ev) (void)hipEventDestroy(e);
      for (auto e : done_ev) (void)hipEventDestroy(e);
      dfree(d_arena);
      if (h_bf) (void)hipHostFree(h_bf);
      if (h_x) (void)hipHostFree(h_x);
      if (h_slots) (void)hipHostFree(h_slots);
      if (h_state) (void)hipHostFree(h_state);
      if (h_xall) (void)hipHostFree(h_xall);
   }
};

namespace {

template <typename T>
int halloc(T** p, size_t count)
{
   if (hipHostMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      *p = nullptr;
      return set_err(FA_E_OOM, "Not enough memory to allocate");
   }
   return FA_OK;
}

size_t frame_bytes(const fa_reference* r, int dtype) { return (size_t)r->n_px * (dtype == FA_U16 ? 2 : 4); }

// A batch whose frames are slots: every slot is sized for the largest of the references and
// starts finished (out of the schedule) until a request is admitted into it.
int slot_batch_build(fa_context* ctx, const std::vector<fa_reference*>& refs, int S,
                     const fa_tr_params* prm, hipStream_t stream, fa_batch** out)
{
   auto* b = new fa_batch();
   b->ctx = ctx;
   b->n_slots = S;
   if (prm) b->prm = *prm;
   else fa_tr_default(&b->prm);
   int nmax = 0, ndmax = 0, fmax = 0;
   int64_t cmax = 0;
   if ((int)refs.size() > kMaxRefs) {
      delete b;
      return set_err(FA_E_ARG, "too many distinct references in one server (at most 4096)");
   }
   std::vector<RefDev> rdev;
   int64_t widest = 0;
   b->n_refs = (int)refs.size();
   for (fa_reference* r : refs) {
      rdev.push_back(r->dev);
      nmax = std::max(nmax, r->nD * r->n_dim);
      ndmax = std::max(ndmax, r->nD);
      fmax = std::max(fmax, r->nD * r->n_dim * r->n_dim);
      cmax = std::max<int64_t>(cmax, (int64_t)r->chunks.size());
      widest = std::max(widest, lds_floats_for(r));
   }
   // a slot may hold a frame of any reference: a class's work is bounded by every slot
   // holding its largest reference
   std::vector<int64_t> ref_items;
   for (fa_reference* r : refs) ref_items.push_back((int64_t)S * (int64_t)r->chunks.size());
   make_classes(b, refs, ref_items, S);
   for (auto& c : b->pcls) {
      int64_t m = 0;
      for (int i = c.ref_lo; i < c.ref_hi; i++) m = std::max(m, ref_items[i]);
      c.work = m;
   }
   b->tune.lds_floats = (int)std::min<int64_t>(widest, kLdsCapFloats);
   b->tune.blocks_per_cu = kBlocksPerCu;
   b->tune.use_lds = 1;
   b->tune.check_every = 8;
   b->tune.use_graph = 1;
   b->tune.tr_lds = 1;
   b->hframes.resize(S);
   for (int j = 0; j < S; j++) {
      BatchFrame& bf = b->hframes[j];
      bf = BatchFrame{};
      bf.work_base = (int64_t)j * cmax;
      bf.x_base = (int64_t)j * nmax;
      bf.vec_base = (int64_t)j * 3 * nmax;
      bf.fac_base = (int64_t)j * 3 * fmax;
      bf.d_base = (int64_t)j * 3 * ndmax;
      b->user.push_back(j);
   }
   b->total_work = S * cmax;
   b->total_x = (int64_t)S * nmax;
   b->total_vec = 3 * (int64_t)S * nmax;
   b->total_fac = 3 * (int64_t)S * fmax;
   b->total_D = 3 * (int64_t)S * ndmax;
   int rc;
   auto fail = [&](int code) {
      delete b;
      return code;
   };
   if ((rc = dalloc(&b->d_refs, rdev.size())) || (rc = dalloc(&b->d_frames, S)) ||
       (rc = dalloc(&b->d_state, S)) || (rc = dalloc(&b->d_x, b->total_x)) ||
       (rc = dalloc(&b->d_vec, b->total_vec)) || (rc = dalloc(&b->d_fac, b->total_fac)) ||
       (rc = dalloc(&b->d_D, b->total_D)) ||
       (rc = dalloc(&b->d_partials, (size_t)b->total_work * kPassStride)) ||
       (rc = dalloc(&b->d_active_frames, S)) || (rc = dalloc(&b->d_run_first, rdev.size())) ||
       (rc = dalloc(&b->d_run_count, rdev.size())) || (rc = dalloc(&b->d_run_base, rdev.size() + 1)) ||
       (rc = dalloc(&b->d_counters, 1)) || (rc = halloc(&b->h_counters, 1)))
      return fail(rc);
   b->stream = stream;  // one of the context's lane streams
   b->own_stream = false;
   std::vector<TrState> st(S);
   for (auto& t : st) {
      t = TrState{};
      t.done = 1;
   }
   hipStream_t s = b->stream;
   if (hipMemcpyAsync(b->d_refs, rdev.data(), rdev.size() * sizeof(RefDev), hipMemcpyHostToDevice, s) != hipSuccess ||
       hipMemcpyAsync(b->d_frames, b->hframes.data(), S * sizeof(BatchFrame), hipMemcpyHostToDevice, s) != hipSuccess ||
       hipMemcpyAsync(b->d_state, st.data(), S * sizeof(TrState), hipMemcpyHostToDevice, s) != hipSuccess ||
       hipMemsetAsync(b->d_counters, 0, sizeof(Counters), s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
      (void)hipGetLastError();
      return fail(set_err(FA_E_DEVICE, "server upload failed"));
   }
   set_grids(b);
   {
      // any slot may hold a frame of any of the lane's references
      std::vector<int32_t> slots(S);
      for (int j = 0; j < S; j++) slots[j] = j;
      if ((rc = set_eig_tables(b, rdev, std::vector<int>(rdev.size(), S), slots))) return fail(rc);
   }
   *out = b;
   return FA_OK;
}

// Takes the results of the parts whose launched rounds have finished (their done events
// have completed; nothing waits). *any: some part was harvested.
int server_harvest(fa_server* sv, bool* any)
{
   *any = false;
   for (int p = 0; p < (int)sv->parts.size(); p++) {
      if (!sv->launched[p]) continue;
      const hipError_t q = hipEventQuery(sv->done_ev[p]);
      if (q == hipErrorNotReady) continue;
      HIPCHK(q);
      sv->launched[p] = 0;
      *any = true;
      const int j0 = sv->first[p], j1 = j0 + sv->parts[p]->n_slots;
      for (int j = j0; j < j1; j++) {
         if (sv->sstate[j] != fa_server::kRunning || !sv->h_state[j].done) continue;
         const fa_server::Req& r = sv->slot_req[j];
         const fa_reference* ref = r.f ? r.f->ref : r.ref;
         const int n = ref->nD * ref->n_dim;
         // the part's parameter vectors as laid out on the device (slot strides of its lane)
         const fa_batch* b = sv->parts[p];
         std::memcpy(r.x, sv->h_xall + (size_t)j0 * sv->nmax + b->hframes[j - j0].x_base, n * sizeof(double));
         fa_server_result res{};
         res.tag = r.tag;
         const TrState& t = sv->h_state[j];
         res.stats.iterations = t.iterations;
         res.stats.evaluations = t.evaluations;
         res.stats.rejected = t.rejected;
         res.stats.status = t.status;
         res.stats.f_final = t.f;
         res.stats.radius_final = t.radius;
         sv->results.push_back(res);
         if (sv->blk_size[j]) sv->arena_free(sv->blk_off[j], sv->blk_size[j]);
         sv->blk_size[j] = 0;
         sv->sstate[j] = fa_server::kFree;
         sv->free_slots[p].push_back(j);
         sv->admitted--;
      }
   }
   return FA_OK;
}

// Takes queued requests in arrival order (over all lanes) while the server admits fewer than
// cap frames, the request's lane has a free slot and, for host data, the frame arena has room
// (admission by free memory: the oldest request that does not fit waits, and the ones behind
// it with it). Host data is uploaded on the copy stream while running frames compute.
int server_admit(fa_server* sv)
{
   bool uploading = false;
   while (sv->admitted < sv->cap) {
      int li = -1;
      for (int l = 0; l < (int)sv->lanes.size(); l++) {
         const fa_server::Lane& L = sv->lanes[l];
         if (L.pending.empty() || sv->free_part(L) < 0) continue;
         if (li < 0 || L.pending.front().seq < sv->lanes[li].pending.front().seq) li = l;
      }
      if (li < 0) break;
      fa_server::Lane& L = sv->lanes[li];
      fa_server::Req& r = L.pending.front();
      const int fp = sv->free_part(L);
      const int slot = sv->free_slots[fp].back();
      if (r.host) {
         const size_t nb = frame_bytes(r.ref, r.dtype);
         size_t off;
         if (!sv->arena_alloc(nb, &off)) break;
         sv->blk_off[slot] = off;
         sv->blk_size[slot] = nb;
         HIPCHK(hipMemcpyAsync(sv->d_arena + off, r.host, nb, hipMemcpyHostToDevice, sv->cstream));
      } else {
         sv->blk_size[slot] = 0;
      }
      HIPCHK(hipEventRecord(sv->up_ev[slot], sv->cstream));
      sv->free_slots[fp].pop_back();
      sv->slot_req[slot] = r;
      sv->sstate[slot] = fa_server::kUploading;
      L.pending.pop_front();
      sv->admitted++;
      uploading = true;
   }
   bool busy = false;
   for (char c : sv->launched) busy = busy || c;
   if (uploading && !busy) HIPCHK(hipStreamSynchronize(sv->cstream));  // nothing to overlap with
   return FA_OK;
}

// For every part without launched rounds: starts its frames whose data is on the device, then
// launches `rounds` trust-region rounds over its running frames (one graph replay), the copies
// of their states and parameter vectors to the host, and its done event.
int server_launch(fa_server* sv, int rounds)
{
   const int P = (int)sv->parts.size();
   for (int pi = 0; pi < P; pi++) {
      if (sv->launched[pi]) continue;
      fa_batch* b = sv->parts[pi];
      const fa_server::Lane& L = sv->lanes[sv->part_lane[pi]];
      const int j0 = sv->first[pi], nj = b->n_slots;
      const int base = j0;  // the part's region of the staging arrays
      int na = 0;
      for (int slot = j0; slot < j0 + nj; slot++) {
         if (sv->sstate[slot] != fa_server::kUploading) continue;
         const hipError_t q = hipEventQuery(sv->up_ev[slot]);
         if (q == hipErrorNotReady) continue;
         HIPCHK(q);
         const int l = slot - j0;
         const fa_server::Req& r = sv->slot_req[slot];
         fa_reference* ref = r.f ? r.f->ref : r.ref;
         BatchFrame bf = b->hframes[l];
         if (r.f) {
            bf.f.data = r.f->d_data;
            bf.f.dtype = r.f->dtype;
            bf.f.offset = r.f->offset;
         } else {
            bf.f.data = sv->d_arena + sv->blk_off[slot];
            bf.f.dtype = r.dtype;
            bf.f.offset = r.offset;
         }
         bf.ref = (int)(std::find(L.refs.begin(), L.refs.end(), ref) - L.refs.begin());
         bf.r = ref->dev;
         bf.trc = b->ref_trc[bf.ref];
         bf.n = ref->nD * ref->n_dim;
         const int i = base + na;
         sv->h_bf[i] = bf;
         std::memcpy(sv->h_x + (size_t)i * sv->nmax, r.x, bf.n * sizeof(double));
         sv->h_slots[i] = l;
         sv->sstate[slot] = fa_server::kRunning;
         na++;
      }
      hipStream_t s = b->stream;
      // the part's stream touches no copy engine: slots start from the pinned staging and the
      // results are published into pinned memory by kernels, so one lane's copies never wait
      // behind another's in a shared copy-engine queue
      if (na)
         HIPCHK(launch_tr_start_slots(b->dev(), sv->dh_bf + base, sv->dh_x + (size_t)base * sv->nmax,
                                      sv->dh_slots + base, na, sv->nmax, s));
      bool busy = false;
      for (int j = j0; j < j0 + nj; j++) busy = busy || sv->sstate[j] == fa_server::kRunning;
      if (!busy) continue;  // its states on the host are current (nothing ran)
      if (b->tune.use_graph) {
         CHK(ensure_graph(b, rounds));
         HIPCHK(hipGraphLaunch(b->gexec, s));
      } else {
         for (int r = 0; r < rounds; r++) CHK(run_round(b));
      }
      HIPCHK(launch_publish(b->dev(), sv->dh_state + j0, sv->dh_xall + (size_t)j0 * sv->nmax, b->total_x, s));
      HIPCHK(hipEventRecord(sv->done_ev[pi], s));
      sv->launched[pi] = 1;
   }
   return FA_OK;
}

// How long a step waits for some part's rounds to finish when none had finished at its start.
constexpr int kStepWaitUs = 500;

// One server step: harvest the parts whose rounds finished, admit, start and launch the idle
// parts; if no part had finished, wait up to kStepWaitUs for one to finish and do the same
// again, so a step normally returns results and relaunches that part at once.
int server_step(fa_server* sv, int rounds)
{
   bool any = false;
   CHK(server_harvest(sv, &any));
   CHK(server_admit(sv));
   CHK(server_launch(sv, rounds));
   if (any) return FA_OK;
   const auto t0 = std::chrono::steady_clock::now();
   for (;;) {
      bool busy = false, ready = false;
      for (int p = 0; p < (int)sv->parts.size(); p++) {
         if (!sv->launched[p]) continue;
         busy = true;
         const hipError_t q = hipEventQuery(sv->done_ev[p]);
         if (q == hipSuccess) ready = true;
         else if (q != hipErrorNotReady) HIPCHK(q);
      }
      if (!busy) return FA_OK;
      if (ready) break;
      if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(kStepWaitUs)) return FA_OK;
      std::this_thread::yield();
   }
   CHK(server_harvest(sv, &any));
   CHK(server_admit(sv));
   CHK(server_launch(sv, rounds));
   return FA_OK;
}
}  // namespace

namespace fa {

int set_error(int code, const std::string& msg) { return set_err(code, msg); }

int fa_caught() noexcept
{
   try {
      g_last_error = "host memory or threads exhausted";
   } catch (...) {
      g_last_error.clear();
   }
   return FA_E_OOM;
}

void batch_set_params(fa_batch* b, const fa_tr_params& p)
{
   std::lock_guard<std::mutex> lb(b->mu);
   b->prm = p;
   for (fa_batch* q : b->parts) q->prm = p;  // a changed set re-captures the round graphs
}

// A frame whose samples are already in device memory (d_data, nz*ny*nx values of dtype), owned
// by the frame from now on (freed on deregistration).
int adopt_device_frame(fa_reference* r, void* d_data, int32_t dtype, float offset, fa_frame** out, bool own)
{
   std::lock_guard<std::mutex> lk(r->ctx->mu);
   auto* f = new fa_frame();
   f->ref = r;
   f->dtype = dtype;
   f->offset = offset;
   f->d_data = d_data;
   f->own_data = own;
   r->live_frames++;
   r->frames.push_back(f);
   *out = f;
   return FA_OK;
}

}  // namespace fa

// ---- eigendecomposition of long node chains -------------------------------------------------
namespace {

// rocSOLVER's symmetric divide-and-conquer eigensolver, opened on first use: only references
// whose nD n_dim lies in [kEigBatchMin, kEigBatchMax] need it, and the library is large.
// Without it those references keep the cyclic-reduction subproblem.
struct EigLib {
   decltype(&rocblas_create_handle) create = nullptr;
   decltype(&rocblas_destroy_handle) destroy = nullptr;
   decltype(&rocblas_set_stream) set_stream = nullptr;
   decltype(&rocsolver_dsyevd) syevd = nullptr;
};

const EigLib* eig_lib()
{
   static EigLib lib;
   static std::once_flag once;
   std::call_once(once, [] {
      void* h = dlopen("librocsolver.so.0", RTLD_NOW | RTLD_LOCAL);
      if (!h) return;
      EigLib l;
      l.create = (decltype(l.create))dlsym(h, "rocblas_create_handle");
      l.destroy = (decltype(l.destroy))dlsym(h, "rocblas_destroy_handle");
      l.set_stream = (decltype(l.set_stream))dlsym(h, "rocblas_set_stream");
      l.syevd = (decltype(l.syevd))dlsym(h, "rocsolver_dsyevd");
      if (l.create && l.destroy && l.set_stream && l.syevd) lib = l;
   });
   return lib.syevd ? &lib : nullptr;
}

// H = Q diag(ev) Q^T for the batched subproblem kernels (RefDev::eig == 2): the dense matrix
// from the blocks H holds on the host, rocsolver_dsyevd on the stream. r->d_Q gets Q
// row-major (column i: the eigenvector of ev[i]), then Q^T, which is the solver's column-major
// output read row-major; ev ascending. False (and nothing kept) when the solver is absent,
// fails or does not converge.
bool large_eigen(fa_reference* r, hipStream_t s)
{
   const EigLib* L = eig_lib();
   if (!L) return false;
   const int m = r->n_dim, mm = m * m, nb = r->nD, n = nb * m;
   std::vector<double> H((size_t)n * n, 0.0);
   for (int k = 0; k < nb; k++)
      for (int a = 0; a < m; a++)
         for (int b = 0; b < m; b++) {
            H[(size_t)(k * m + a) * n + k * m + b] = r->Hdiag[(size_t)k * mm + a * m + b];
            if (k + 1 < nb) {
               const double v = r->Hsub[(size_t)k * mm + a * m + b];  // H((k+1) m + a, k m + b)
               H[(size_t)((k + 1) * m + a) * n + k * m + b] = v;
               H[(size_t)(k * m + b) * n + (k + 1) * m + a] = v;
            }
         }
   double* d_Q = nullptr;
   double* d_ev = nullptr;
   double* d_E = nullptr;
   rocblas_int* d_info = nullptr;
   rocblas_handle h = nullptr;
   bool ok = dalloc(&d_Q, 2 * (size_t)n * n) == FA_OK && dalloc(&d_ev, n) == FA_OK && dalloc(&d_E, n) == FA_OK &&
             dalloc(&d_info, 1) == FA_OK;
   double* d_Qt = d_Q ? d_Q + (size_t)n * n : nullptr;
   rocblas_int info = -1;
   double ev0 = 0;
   ok = ok && L->create(&h) == rocblas_status_success;
   ok = ok && L->set_stream(h, s) == rocblas_status_success &&
        hipMemcpyAsync(d_Qt, H.data(), H.size() * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess &&
        L->syevd(h, rocblas_evect_original, rocblas_fill_upper, n, d_Qt, n, d_ev, d_E, d_info) ==
            rocblas_status_success &&
        launch_transpose(d_Qt, d_Q, n, s) == hipSuccess &&
        hipMemcpyAsync(&info, d_info, sizeof info, hipMemcpyDeviceToHost, s) == hipSuccess &&
        hipMemcpyAsync(&ev0, d_ev, sizeof ev0, hipMemcpyDeviceToHost, s) == hipSuccess &&
        hipStreamSynchronize(s) == hipSuccess && info == 0;
   if (h) L->destroy(h);
   (void)hipGetLastError();
   dfree(d_E);
   dfree(d_info);
   if (!ok) {
      dfree(d_Q);
      dfree(d_ev);
      return false;
   }
   r->d_Q = d_Q;
   r->d_ev = d_ev;
   r->dev.Q = d_Q;
   r->dev.ev = d_ev;
   r->dev.ev_min = ev0;
   r->dev.eig = 2;
   return true;
}

}  // namespace

// ---- allocation-free alignment --------------------------------------------------------------
// fa_align_noalloc runs fa_align's kernels with every device buffer it needs (the frame's
// trust-region vectors, factor scratch, trial displacement, pass partials, state and counters)
// carved from one file-scope device array, so it calls no allocator. Each GPU the library runs
// on holds its own copy of the array until the process ends; its device address is looked up
// once per device when a context is created on it. Calls on one GPU share the array: calls
// from different contexts that overlap on one GPU race (calls through one context are
// serialised by its mutex).
namespace {

constexpr size_t kNoAllocScratchBytes = 4u << 20;  // 4 MiB per GPU
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

// With the device current: the device address of this GPU's scratch array (once per device).
void scratch_init(int device)
{
   if (device < 0 || device >= kMaxDevices) return;
   std::lock_guard<std::mutex> lk(g_scratch_mu);
   if (g_scratch[device]) return;
   void* ptr = nullptr;
   const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
   if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      g_scratch_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
      return;
   }
   g_scratch[device] = ptr;
}

// Carves aligned pieces off the scratch array; overflow leaves ok false.
struct Carve {
   unsigned char* base;
   size_t used = 0;
   bool ok = true;
   template <typename T>
   T* take(size_t count)
   {
      const size_t at = (used + 255) & ~(size_t)255;
      const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
      if (at + bytes > kNoAllocScratchBytes) {
         ok = false;
         return nullptr;
      }
      used = at + bytes;
      return (T*)(base + at);
   }
};

int align_noalloc(fa_frame* f, const fa_tr_params* prm, double* x_inout, fa_align_stats* stats)
{
   fa_reference* r = f->ref;
   fa_context* ctx = r->ctx;
   const int dev = ctx->device;
   void* base = (dev >= 0 && dev < kMaxDevices) ? g_scratch[dev] : nullptr;
   if (!base)
      return set_err(FA_E_DEVICE, std::string("allocation-free scratch unavailable on this device: ") +
                                      hipGetErrorString(g_scratch_err[dev >= 0 && dev < kMaxDevices ? dev : 0]));
   const int n = r->nD * r->n_dim;
   const int64_t n_chunks = (int64_t)r->chunks.size();
   Carve cv{(unsigned char*)base};
   RefDev* d_ref = cv.take<RefDev>(1);
   BatchFrame* d_frame = cv.take<BatchFrame>(1);
   TrState* d_state = cv.take<TrState>(1);
   Counters* d_counters = cv.take<Counters>(1);
   int32_t* d_active = cv.take<int32_t>(1);
   int32_t* d_run_first = cv.take<int32_t>(1);
   int32_t* d_run_count = cv.take<int32_t>(1);
   int64_t* d_run_base = cv.take<int64_t>(2);
   double* d_x = cv.take<double>(n);
   double* d_vec = cv.take<double>(3 * (size_t)n);
   double* d_fac = cv.take<double>(3 * (size_t)r->nD * r->n_dim * r->n_dim);
   double* d_D = cv.take<double>(3 * (size_t)r->nD);
   double* d_partials = cv.take<double>((size_t)n_chunks * kPassStride);
   if (!cv.ok)
      return set_err(FA_E_ARG, "problem too large for the allocation-free entry point (its scratch is " +
                                   std::to_string(kNoAllocScratchBytes >> 20) + " MiB per GPU)");
   fa_tr_params P;
   if (prm) P = *prm;
   else fa_tr_default(&P);

   BatchFrame bf{};
   bf.f.data = f->d_data;
   bf.f.dtype = f->dtype;
   bf.f.offset = f->offset;
   bf.ref = 0;
   bf.r = f->ref->dev;
   bf.n = n;
   bf.trc = 0;
   BatchDev B{};
   B.refs = d_ref;
   B.frames = d_frame;
   B.state = d_state;
   B.x = d_x;
   B.vec = d_vec;
   B.fac = d_fac;
   B.Dtrial = d_D;
   B.partials = d_partials;
   B.active_frames = d_active;
   B.run_first = d_run_first;
   B.run_count = d_run_count;
   B.run_base = d_run_base;
   B.counters = d_counters;
   B.prm.radius = P.radius;
   B.prm.min_delta = P.min_delta;
   B.prm.rel_delta = P.rel_delta;
   B.prm.sub_eps = P.sub_eps;
   B.prm.max_iter = P.max_iter;
   B.prm.max_evals = P.max_evals;
   B.prm.sub_max_iter = P.sub_max_iter;
   B.prm.skip_repeat = P.skip_repeat;
   B.n_frames = 1;
   B.n_refs = 1;
   B.trace = nullptr;
   B.split_eig = r->dev.eig == 2 ? 1 : 0;
   B.n_eig_tiles = B.split_eig ? (n + 15) / 16 * 4 : 0;  // the kernels' tiles of one frame (no table)
   B.n_eig_frames = B.split_eig;
   // the launch shape and update variant a one-frame batch of this reference takes
   const int lds_floats = (int)std::min<int64_t>(lds_floats_for(r), kLdsCapFloats);
   PassArgs pa{};
   pa.refs = d_ref;
   pa.frames = d_frame;
   pa.Dtrial = d_D;
   pa.partials = d_partials;
   pa.active_frames = d_active;
   pa.run_first = d_run_first;
   pa.run_count = d_run_count;
   pa.run_base = d_run_base;
   pa.n_refs = 1;
   pa.counters = d_counters;
   pa.lds_cap = lds_floats;
   pa.use_lds = 1;
   pa.ref_lo = 0;
   pa.ref_hi = 1;
   const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(n_chunks, (int64_t)ctx->n_cu * kBlocksPerCu));
   const int need = tr_lds_need(r), need_full = tr_lds_full(r);
   const bool full = need_full <= kTrFullDoubles;
   const int tr_lds = tr_lds_for(ctx, full ? need_full : need);
   const int tr_full = full && tr_lds > 0 ? 1 : 0;
   const int tr_nt = tr_threads(r);

   hipStream_t s = ctx->stream;
   const RefDev rd = r->dev;
   const Counters zero{};
   HIPCHK(hipMemcpyAsync(d_ref, &rd, sizeof rd, hipMemcpyHostToDevice, s));
   HIPCHK(hipMemcpyAsync(d_frame, &bf, sizeof bf, hipMemcpyHostToDevice, s));
   HIPCHK(hipMemcpyAsync(d_counters, &zero, sizeof zero, hipMemcpyHostToDevice, s));
   HIPCHK(hipMemcpyAsync(d_x, x_inout, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
   HIPCHK(launch_tr_init(B, 0, s));
   constexpr int K = 8;  // rounds between polls of the active count
   const int64_t limit = P.max_evals > 0 ? (int64_t)P.max_evals + 2 : 100000;
   Counters hc{};
   for (int64_t round = 0; round < limit;) {
      for (int k = 0; k < K && round < limit; k++, round++) {
         HIPCHK(launch_schedule(B, 1, s));
         HIPCHK(launch_pass(pa, r->n_dim, grid, s));
         HIPCHK(launch_tr_update(B, 0, tr_lds, tr_full, tr_nt, s));
         HIPCHK(launch_eig_subproblems(B, s));
      }
      HIPCHK(launch_schedule(B, 0, s));
      HIPCHK(hipMemcpyAsync(&hc, d_counters, sizeof hc, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (hc.n_active_frames == 0) break;
   }
   TrState st{};
   HIPCHK(hipMemcpyAsync(x_inout, d_x, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
   HIPCHK(hipMemcpyAsync(&st, d_state, sizeof st, hipMemcpyDeviceToHost, s));
   HIPCHK(hipStreamSynchronize(s));
   if (stats) {
      stats->iterations = st.iterations;
      stats->evaluations = st.evaluations;
      stats->rejected = st.rejected;
      stats->status = st.done ? st.status : FA_STOP_MAX_EVALS;
      stats->f_final = st.f;
      stats->radius_final = st.radius;
   }
   return FA_OK;
}

}  // namespace

// ---- C ABI ------------------------------------------------------------------------------------

extern "C" {

const char* fa_last_error(void) { return g_last_error.c_str(); }

void fa_tr_default(fa_tr_params* p)
{
   p->radius = 40;
   p->min_delta = 1e-8;
   p->rel_delta = 0;
   p->max_iter = 0;
   p->max_evals = 0;
   p->sub_eps = 0.1;
   p->sub_max_iter = 20;
   p->skip_repeat = 0;
}

int fa_create(int device, fa_context** out)
{
   try {
      if (!out) return set_err(FA_E_ARG, "null output");
      *out = nullptr;
      int count = 0;
      if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
         (void)hipGetLastError();
         return set_err(FA_E_DEVICE, "No supported GPU");
      }
      hipDeviceProp_t prop;
      HIPCHK(hipGetDeviceProperties(&prop, device));
      if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
         return set_err(FA_E_DEVICE, std::string("No supported GPU (found ") + prop.gcnArchName + ")");
      HIPCHK(hipSetDevice(device));
      scratch_init(device);  // the allocation-free entry point's array on this GPU
      auto* c = new fa_context();
      c->device = device;
      c->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
      c->lds_per_block = (int)std::min<size_t>(prop.sharedMemPerBlock, 160 * 1024);
      // raises the dynamic-LDS limit of the kernels that use more than 64 KB where the runtime
      // keeps one; launches report their own errors
      c->lds_config = (int)configure_kernels(c->lds_per_block);
      (void)hipGetLastError();
      bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
      for (int l = 0; ok && l < fa_context::kLaneStreams; l++)
         ok = hipStreamCreateWithFlags(&c->lane_stream[l], hipStreamNonBlocking) == hipSuccess;
      if (!ok) {
         (void)hipGetLastError();
         for (hipStream_t s : c->lane_stream)
            if (s) (void)hipStreamDestroy(s);
         if (c->stream) (void)hipStreamDestroy(c->stream);
         delete c;
         return set_err(FA_E_DEVICE, "stream creation failed");
      }
      *out = c;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_destroy(fa_context* ctx)
{
   try {
      if (!ctx) return FA_OK;
      if (ctx->live_refs > 0)
         return set_err(FA_E_STATE, "context still has live references; release them first");
      DeviceGuard g(ctx->device);
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipStreamDestroy(ctx->stream);
      for (hipStream_t s : ctx->lane_stream) {
         (void)hipStreamSynchronize(s);
         (void)hipStreamDestroy(s);
      }
      delete ctx;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

void* fa_context_stream(fa_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int fa_set_reference(fa_context* ctx, const float* ref, int32_t nz, int32_t ny, int32_t nx,
                     int32_t nD, const fa_scan_params* scan, int32_t chunk_px, fa_reference** out)
{
   try {
      if (!ctx || !ref || !scan || !out) return set_err(FA_E_ARG, "null argument");
      *out = nullptr;
      if (nz < 1 || ny < 2 || nx < 2) return set_err(FA_E_ARG, "Reference must be three dimensional");
      const int64_t n_px = (int64_t)nz * ny * nx;
      // pixel indices are 32-bit, with room for the pass to step a few blocks past an item's end
      if (n_px > (int64_t)INT32_MAX - 4 * kBlock) return set_err(FA_E_ARG, "volume too large");
      if (nD < 2 || nD - 1 > 65535 || nD <= nz) return set_err(FA_E_ARG, "invalid number of nodes");
      if (scan->bidirectional) return set_err(FA_E_ARG, "bidirectional scans are not supported");
      std::lock_guard<std::mutex> lk(ctx->mu);
      DeviceGuard g(ctx->device);
      hipStream_t s = ctx->stream;

      auto* r = new fa_reference();
      r->ctx = ctx;
      r->nz = nz;
      r->ny = ny;
      r->nx = nx;
      r->nD = nD;
      r->n_dim = nz > 1 ? 3 : 2;
      r->n_px = n_px;
      // Frames at least 768 wide whose width is a multiple of kStripWidth are cut into strips of
      // kStripWidth columns: a work item then stages a narrow tile of many rows instead of a few
      // whole rows (less staging per pixel, more pixels per item for the same LDS).
      const bool strips = nx >= 3 * kStripWidth && nx % kStripWidth == 0;
      r->item_w = strips ? kStripWidth : nx;
      // default work item: strips 4096 pixels (16 rows); else 8 rows within 2048..8192 pixels
      // (2-D), 2048 pixels or one row of a wider frame (3-D, which stages two planes per item)
      r->chunk_px = chunk_px > 0 ? chunk_px
                    : strips     ? 16 * kStripWidth
                    : nz > 1     ? std::max(2048, nx)
                                 : (int)std::min<int64_t>(8192, std::max<int64_t>(2048, 8 * (int64_t)nx));
      derive_scan(*scan, nz, ny, nx, nD, r->scan_host, &r->scan);
      const int m = r->n_dim;

      float* d_T = nullptr;
      uint16_t* d_Di = nullptr;
      int32_t* d_begin = nullptr;
      int32_t* d_flags = nullptr;
      double* d_hpart = nullptr;
      double* d_info = nullptr;
      auto cleanup = [&]() {
         dfree(d_T);
         dfree(d_Di);
         dfree(d_begin);
         dfree(d_flags);
         dfree(d_hpart);
         dfree(d_info);
      };
      auto fail = [&](int code) {
         cleanup();
         fa_reference_release(r);
         return code;
      };
      int rc;
      if ((rc = dalloc(&d_T, n_px)) || (rc = dalloc(&d_Di, n_px)) || (rc = dalloc(&d_begin, nD)) ||
          (rc = dalloc(&d_flags, 1)) || (rc = dalloc(&d_info, 2)) ||
          (rc = dalloc(&r->d_rg4, (size_t)n_px)) ||
          (m == 3 && (rc = dalloc(&r->d_df, (size_t)n_px))))
         return fail(rc);

      // node map on the device, interval starts back to the host
      std::vector<int32_t> begin(nD, INT32_MAX);
      uint16_t last_i = 0;
      int32_t flags = 0;
      if (hipMemcpyAsync(d_T, ref, n_px * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess ||
          hipMemcpyAsync(d_begin, begin.data(), nD * sizeof(int32_t), hipMemcpyHostToDevice, s) != hipSuccess ||
          hipMemsetAsync(d_flags, 0, sizeof(int32_t), s) != hipSuccess ||
          launch_node_map(r->scan, nx, ny, n_px, d_Di, nullptr, s) != hipSuccess ||
          launch_interval_begin(d_Di, n_px, nD, d_begin, d_flags, s) != hipSuccess ||
          hipMemcpyAsync(begin.data(), d_begin, nD * sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipMemcpyAsync(&last_i, d_Di + (n_px - 1), sizeof(uint16_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipMemcpyAsync(&flags, d_flags, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
         (void)hipGetLastError();
         return fail(set_err(FA_E_DEVICE, "node map failed"));
      }
      if (flags & 2)
         return fail(set_err(FA_E_ARG, "scan timing is not monotonic in scan order "
                                       "(interframe_duration < frame duration)"));
      if ((flags & 1) || last_i != nD - 2) return fail(set_err(FA_E_NODE_MAP, "Something went wrong"));

      // D_range exactly as precomputeInterp leaves it (FrameWarper.cpp:269-283)
      r->D_range.assign(2 * nD, 0);
      int last = -1;
      for (int i = 0; i <= nD - 2; i++) {
         if (begin[i] == INT32_MAX) continue;
         r->D_range[2 * i] = begin[i];
         if (last >= 0) r->D_range[2 * last + 1] = begin[i] - 1;
         last = i;
      }
      r->D_range[2 * last + 1] = (int32_t)(n_px - 1);

      // work chunks: the part of an interval in one plane is cut into the fewest items of at
      // most chunk_px pixels, of equal length (to one pixel), so no item is a short remainder
      const int64_t plane = (int64_t)nx * ny;
      r->ichunk.assign(nD, 0);
      r->item_rows = 1;
      for (int i = 0; i <= nD - 2; i++) {
         r->ichunk[i] = (int32_t)r->chunks.size();
         if (begin[i] == INT32_MAX) continue;
         int64_t b0 = r->D_range[2 * i], e0 = r->D_range[2 * i + 1];
         for (int64_t s0 = b0; s0 <= e0;) {
            const int64_t s1 = std::min<int64_t>(e0, (s0 / plane + 1) * plane - 1);  // segment end
            if (!strips) {
               const int64_t len = s1 - s0 + 1, m = (len + r->chunk_px - 1) / r->chunk_px;
               const int64_t base = len / m, extra = len % m;
               int64_t p = s0;
               for (int64_t t = 0; t < m; t++) {
                  const int64_t q = p + base + (t < extra ? 1 : 0) - 1;
                  r->chunks.push_back(Chunk{(int32_t)p, (int32_t)(q + 1), i, (int32_t)e0, 0, nx});
                  r->item_rows = std::max<int>(r->item_rows, (int)(q / nx - p / nx + 1));
                  p = q + 1;
               }
            } else {
               // the segment's rows in the fewest equal groups of at most chunk_px / kStripWidth
               // rows, each group cut into strips; an item runs from the strip's first pixel in
               // the segment to its last
               const int64_t ra = s0 / nx, rb = s1 / nx, nrow = rb - ra + 1;
               const int64_t rmax = std::max<int64_t>(1, r->chunk_px / kStripWidth);
               const int64_t m = (nrow + rmax - 1) / rmax, base = nrow / m, extra = nrow % m;
               int64_t ga = ra;
               for (int64_t t = 0; t < m; t++) {
                  const int64_t gb = ga + base + (t < extra ? 1 : 0) - 1;
                  for (int cx = 0; cx < nx; cx += kStripWidth) {
                     // first pixel: in row ga unless the segment starts past the strip there
                     int64_t p0 = ga * nx + cx;
                     if (p0 + kStripWidth - 1 < s0) p0 = (ga + 1) * nx + cx;
                     p0 = std::max(p0, s0);
                     // last pixel: in row gb unless the segment ends before the strip there
                     int64_t pl = gb * nx + cx + kStripWidth - 1;
                     if (gb * nx + cx > s1) pl = (gb - 1) * nx + cx + kStripWidth - 1;
                     pl = std::min(pl, s1);
                     if (pl < p0) continue;
                     r->chunks.push_back(Chunk{(int32_t)p0, (int32_t)(pl + 1), i, (int32_t)e0, cx, kStripWidth});
                     r->item_rows = std::max<int>(r->item_rows, (int)(pl / nx - p0 / nx + 1));
                  }
                  ga = gb + 1;
               }
            }
            s0 = s1 + 1;
         }
      }
      r->ichunk[nD - 1] = (int32_t)r->chunks.size();
      const int nc = (int)r->chunks.size();
      const int mm = m * m;
      if ((rc = dalloc(&r->d_chunks, nc)) || (rc = dalloc(&r->d_ichunk, nD)) ||
          (rc = dalloc(&r->d_Hdiag, (size_t)nD * mm)) || (rc = dalloc(&r->d_Hsub, (size_t)nD * mm)) ||
          (rc = dalloc(&r->d_L0, (size_t)3 * nD * mm)) ||
          (rc = dalloc(&d_hpart, (size_t)nc * kHessStride)))
         return fail(rc);

      RefDev& R = r->dev;
      R.rg4 = r->d_rg4;
      R.df = r->d_df;
      R.chunks = r->d_chunks;
      R.ichunk = r->d_ichunk;
      R.Hdiag = r->d_Hdiag;
      R.Hsub = r->d_Hsub;
      R.L0 = r->d_L0;
      R.scan = r->scan;
      R.nx = nx;
      R.ny = ny;
      R.nz = nz;
      R.nD = nD;
      R.n_dim = m;
      R.n_chunks = nc;
      R.h_pd = 0;
      R.eig_lo = 0;
      R.n_px = n_px;

      double info[2] = {0, 0};
      r->Hdiag.assign((size_t)nD * mm, 0);
      r->Hsub.assign((size_t)nD * mm, 0);
      if (hipMemcpyAsync(r->d_chunks, r->chunks.data(), nc * sizeof(Chunk), hipMemcpyHostToDevice, s) != hipSuccess ||
          hipMemcpyAsync(r->d_ichunk, r->ichunk.data(), nD * sizeof(int32_t), hipMemcpyHostToDevice, s) != hipSuccess ||
          hipMemsetAsync(r->d_Hsub, 0, (size_t)nD * mm * sizeof(double), s) != hipSuccess ||
          hipMemsetAsync(r->d_L0, 0, (size_t)3 * nD * mm * sizeof(double), s) != hipSuccess ||
          launch_reference_record(d_T, nz, ny, nx, m, r->scan, r->d_rg4, r->d_df, s) != hipSuccess ||
          launch_hessian(R, d_hpart, r->d_Hdiag, r->d_Hsub, r->d_L0, d_info, s) != hipSuccess ||
          hipMemcpyAsync(info, d_info, 2 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipMemcpyAsync(r->Hdiag.data(), r->d_Hdiag, (size_t)nD * mm * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipMemcpyAsync(r->Hsub.data(), r->d_Hsub, (size_t)nD * mm * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
         (void)hipGetLastError();
         return fail(set_err(FA_E_DEVICE, "reference set-up failed"));
      }
      r->h_pd = info[0] != 0.0;
      r->eig_lo = info[1];
      R.h_pd = r->h_pd;
      R.eig_lo = r->eig_lo;
      // small problems: eigendecomposition of H for the trust-region subproblem
      R.eig = 0;
      R.q_lds = 0;
      R.ev_min = 0;
      const int nn = nD * m, mp = nn + (nn & 1);
      if (nn <= kEigMax && (int64_t)(2 * mp * mp + mp) * (int64_t)sizeof(double) <= ctx->lds_per_block) {
         if ((rc = dalloc(&r->d_Q, 2 * (size_t)nn * nn)) || (rc = dalloc(&r->d_ev, nn))) return fail(rc);  // Q, Q^T
         R.Q = r->d_Q;
         R.ev = r->d_ev;
         if (launch_eigen(R, r->d_Q, r->d_ev, d_info, s) != hipSuccess ||
             hipMemcpyAsync(info, d_info, 2 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
             hipStreamSynchronize(s) != hipSuccess) {
            (void)hipGetLastError();
            return fail(set_err(FA_E_DEVICE, "reference set-up failed"));
         }
         R.eig = info[0] != 0.0 ? 1 : 0;
         R.ev_min = info[1];
         R.q_lds = R.eig && tr_lds_full_base(r) + tr_lds_q(r) <= kTrFullDoubles ? 1 : 0;
      } else if (nn >= kEigBatchMin && nn <= kEigBatchMax) {
         large_eigen(r, s);  // else the cyclic-reduction subproblem
      }
      cleanup();
      ctx->live_refs++;
      r->counted = true;
      *out = r;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_reference_release(fa_reference* r)
{
   try {
      if (!r) return FA_OK;
      if (r->live_frames > 0)
         return set_err(FA_E_STATE, "reference still has registered frames or servers; release them first");
      DeviceGuard g(r->ctx->device);
      if (r->counted) r->ctx->live_refs--;
      dfree(r->d_rg4);
      dfree(r->d_df);
      dfree(r->d_chunks);
      dfree(r->d_ichunk);
      dfree(r->d_Hdiag);
      dfree(r->d_Hsub);
      dfree(r->d_L0);
      dfree(r->d_Q);
      dfree(r->d_ev);
      delete r;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_reference_info(const fa_reference* r, int32_t* o)
{
   try {
      if (!r || !o) return set_err(FA_E_ARG, "null argument");
      o[0] = r->nz;
      o[1] = r->ny;
      o[2] = r->nx;
      o[3] = r->nD;
      o[4] = r->n_dim;
      o[5] = (int32_t)r->chunks.size();
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_get_hessian(const fa_reference* r, double* H)
{
   try {
      if (!r || !H) return set_err(FA_E_ARG, "null argument");
      const int m = r->n_dim, n = r->nD * m, mm = m * m;
      std::fill(H, H + (size_t)n * n, 0.0);
      for (int k = 0; k < r->nD; k++)
         for (int a = 0; a < m; a++)
            for (int c = 0; c < m; c++) {
               H[(size_t)(k * m + a) * n + k * m + c] = r->Hdiag[(size_t)k * mm + a * m + c];
               if (k + 1 < r->nD) {
                  double v = r->Hsub[(size_t)k * mm + a * m + c];
                  H[(size_t)((k + 1) * m + a) * n + k * m + c] = v;
                  H[(size_t)(k * m + c) * n + (k + 1) * m + a] = v;
               }
            }
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_get_node_map(const fa_reference* r, uint16_t* Di, float* Df, int32_t* D_range)
{
   try {
      if (!r) return set_err(FA_E_ARG, "null argument");
      if (D_range) std::memcpy(D_range, r->D_range.data(), r->D_range.size() * sizeof(int32_t));
      if (!Di && !Df) return FA_OK;
      std::lock_guard<std::mutex> lk(r->ctx->mu);
      DeviceGuard g(r->ctx->device);
      hipStream_t s = r->ctx->stream;
      uint16_t* d_Di = nullptr;
      float* d_Df = nullptr;
      CHK(dalloc(&d_Di, r->n_px));
      if (dalloc(&d_Df, r->n_px) != FA_OK) {
         dfree(d_Di);
         return FA_E_OOM;
      }
      int rc = FA_OK;
      if (launch_node_map(r->scan, r->nx, r->ny, r->n_px, d_Di, d_Df, s) != hipSuccess ||
          (Di && hipMemcpyAsync(Di, d_Di, r->n_px * sizeof(uint16_t), hipMemcpyDeviceToHost, s) != hipSuccess) ||
          (Df && hipMemcpyAsync(Df, d_Df, r->n_px * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess) ||
          hipStreamSynchronize(s) != hipSuccess) {
         (void)hipGetLastError();
         rc = set_err(FA_E_DEVICE, "node map failed");
      }
      dfree(d_Di);
      dfree(d_Df);
      return rc;
   } catch (...) {
      return fa_caught();
   }
}

int fa_get_gradient(const fa_reference* r, float* gx, float* gy, float* gz)
{
   try {
      if (!r) return set_err(FA_E_ARG, "null argument");
      std::lock_guard<std::mutex> lk(r->ctx->mu);
      DeviceGuard g(r->ctx->device);
      const int m = r->n_dim;
      std::vector<float> rg((size_t)r->n_px * 4);
      HIPCHK(hipMemcpyAsync(rg.data(), r->d_rg4, rg.size() * sizeof(float), hipMemcpyDeviceToHost, r->ctx->stream));
      HIPCHK(hipStreamSynchronize(r->ctx->stream));
      for (int64_t p = 0; p < r->n_px; p++) {
         if (gx) gx[p] = rg[p * 4 + 1];
         if (gy) gy[p] = rg[p * 4 + 2];
         if (gz) gz[p] = m == 3 ? rg[p * 4 + 3] : 0.f;
      }
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_register_frame(fa_reference* r, const void* data, int32_t dtype, float add_offset, fa_frame** out)
{
   try {
      if (!r || !data || !out) return set_err(FA_E_ARG, "null argument");
      *out = nullptr;
      if (dtype != FA_F32 && dtype != FA_U16) return set_err(FA_E_ARG, "unsupported frame dtype");
      std::lock_guard<std::mutex> lk(r->ctx->mu);
      DeviceGuard g(r->ctx->device);
      auto* f = new fa_frame();
      f->ref = r;
      f->dtype = dtype;
      f->offset = add_offset;
      size_t bytes = (size_t)r->n_px * (dtype == FA_U16 ? 2 : 4);
      char* d = nullptr;
      int rc = dalloc(&d, bytes);
      if (rc) {
         delete f;
         return rc;
      }
      f->d_data = d;
      if (hipMemcpyAsync(d, data, bytes, hipMemcpyHostToDevice, r->ctx->stream) != hipSuccess ||
          hipStreamSynchronize(r->ctx->stream) != hipSuccess) {
         (void)hipGetLastError();
         dfree(d);
         delete f;
         return set_err(FA_E_DEVICE, "frame upload failed");
      }
      r->live_frames++;
      r->frames.push_back(f);
      *out = f;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_deregister_frame(fa_frame* f)
{
   try {
      if (!f) return FA_OK;
      std::lock_guard<std::mutex> lk(f->ref->ctx->mu);
      std::unique_lock<std::mutex> lf(f->mu);  // waits for a call on the frame still running
      DeviceGuard g(f->ref->ctx->device);
      delete f->single;
      f->single = nullptr;
      void* d = f->d_data;
      if (f->own_data) dfree(d);
      fa_reference* r = f->ref;
      r->live_frames--;
      r->frames.erase(std::remove(r->frames.begin(), r->frames.end(), f), r->frames.end());
      lf.unlock();
      delete f;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_clear_temp(fa_reference* r)
{
   try {
      if (!r) return set_err(FA_E_ARG, "null reference");
      std::lock_guard<std::mutex> lk(r->ctx->mu);
      DeviceGuard g(r->ctx->device);
      for (fa_frame* f : r->frames) {
         std::lock_guard<std::mutex> lf(f->mu);
         delete f->single;  // rebuilt at the frame's next fa_align or single evaluation
         f->single = nullptr;
      }
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_get_error(fa_frame* f, const double* D, double* err)
{
   try {
      if (!f || !D || !err) return set_err(FA_E_ARG, "null argument");
      std::lock_guard<std::mutex> lk(f->mu);
      DeviceGuard g(f->ref->ctx->device);
      fa_batch* b;
      CHK(single_batch(f, &b));
      return batch_eval(b, D, err, nullptr);
   } catch (...) {
      return fa_caught();
   }
}

int fa_get_jacobian(fa_frame* f, const double* D, double* jac, double* err)
{
   try {
      if (!f || !D || !jac) return set_err(FA_E_ARG, "null argument");
      std::lock_guard<std::mutex> lk(f->mu);
      DeviceGuard g(f->ref->ctx->device);
      fa_batch* b;
      CHK(single_batch(f, &b));
      double e = 0;
      CHK(batch_eval(b, D, &e, jac));
      if (err) *err = e;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_warp_interpolated(fa_frame* f, const double* D, float* out, float invalid_value)
{
   try {
      if (!f || !D || !out) return set_err(FA_E_ARG, "null argument");
      std::lock_guard<std::mutex> lk(f->mu);
      DeviceGuard g(f->ref->ctx->device);
      fa_batch* b;
      CHK(single_batch(f, &b));
      fa_reference* r = f->ref;
      hipStream_t s = b->stream;
      if (!b->d_warp) {
         CHK(dalloc(&b->d_warp, r->n_px));
         b->warp_px = r->n_px;
      }
      FrameDev F;
      F.data = f->d_data;
      F.dtype = f->dtype;
      F.offset = f->offset;
      int grid = (int)std::min<int64_t>((int64_t)r->chunks.size(), (int64_t)r->ctx->n_cu * kBlocksPerCu);
      HIPCHK(hipMemcpyAsync(b->d_D, D, 3 * (size_t)r->nD * sizeof(double), hipMemcpyHostToDevice, s));
      HIPCHK(launch_warp(r->dev, F, b->d_D, b->d_warp, invalid_value, b->tune.lds_floats, b->tune.use_lds, grid, s));
      HIPCHK(hipMemcpyAsync(out, b->d_warp, r->n_px * sizeof(float), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_warp_image(fa_frame* f, const double* D, float* wimg, float* coverage, int32_t apply_offset)
{
   try {
      if (!f || !D || !wimg || !coverage) return set_err(FA_E_ARG, "null argument");
      std::lock_guard<std::mutex> lk(f->mu);
      DeviceGuard g(f->ref->ctx->device);
      fa_batch* b;
      CHK(single_batch(f, &b));
      fa_reference* r = f->ref;
      hipStream_t s = b->stream;
      if (!b->d_warp) {
         CHK(dalloc(&b->d_warp, r->n_px));
         b->warp_px = r->n_px;
      }
      if (!b->d_cov) CHK(dalloc(&b->d_cov, r->n_px));
      FrameDev F;
      F.data = f->d_data;
      F.dtype = f->dtype;
      F.offset = f->offset;
      int grid = (int)std::min<int64_t>((int64_t)r->chunks.size(), (int64_t)r->ctx->n_cu * kBlocksPerCu);
      HIPCHK(hipMemcpyAsync(b->d_D, D, 3 * (size_t)r->nD * sizeof(double), hipMemcpyHostToDevice, s));
      HIPCHK(hipMemsetAsync(b->d_warp, 0, r->n_px * sizeof(float), s));
      HIPCHK(hipMemsetAsync(b->d_cov, 0, r->n_px * sizeof(float), s));
      HIPCHK(launch_splat(r->dev, F, b->d_D, b->d_warp, b->d_cov, apply_offset ? f->offset : 0.f, grid, s));
      HIPCHK(hipMemcpyAsync(wimg, b->d_warp, r->n_px * sizeof(float), hipMemcpyDeviceToHost, s));
      HIPCHK(hipMemcpyAsync(coverage, b->d_cov, r->n_px * sizeof(float), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_warp_point(const fa_reference* r, const double* D, int32_t x, int32_t y, int32_t z,
                  int32_t binning, double* out3)
{
   try {
      if (!r || !D || !out3) return set_err(FA_E_ARG, "null argument");
      if (binning < 1) return set_err(FA_E_ARG, "spatial binning must be >= 1");
      double factor = 1.0 / binning;
      int xs = (int)(x / factor);
      int ys = (int)(y / factor);
      out3[0] = out3[1] = out3[2] = 0;
      if (xs < 0 || xs >= r->nx || ys < 0 || ys >= r->ny || z < 0 || z >= r->nz) return FA_OK;
      double t = (double)z * r->scan.interframe + (double)ys * r->scan.interline + (double)xs * r->scan.pixel;
      double u = t / r->scan.stack * r->scan.nDm1;
      double ip = std::trunc(u);
      double f = (double)(float)(u - ip);
      int i = (int)ip;
      for (int c = 0; c < 3; c++) {
         double a = (c < r->n_dim) ? D[3 * (i + 1) + c] : 0.0, b = (c < r->n_dim) ? D[3 * i + c] : 0.0;
         out3[c] = (f * a + (1 - f) * b) * factor;
      }
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_align(fa_frame* f, const fa_tr_params* prm, double* x_inout, fa_align_stats* stats)
{
   try {
      if (!f || !x_inout) return set_err(FA_E_ARG, "null argument");
      std::lock_guard<std::mutex> lk(f->mu);
      DeviceGuard g(f->ref->ctx->device);
      fa_batch* b;
      CHK(single_batch(f, &b));
      if (prm) b->prm = *prm;
      else fa_tr_default(&b->prm);
      return batch_run(b, x_inout, stats, nullptr);
   } catch (...) {
      return fa_caught();
   }
}

int fa_align_noalloc(fa_frame* f, const fa_tr_params* prm, double* x_inout, fa_align_stats* stats)
{
   try {
      if (!f || !x_inout) return set_err(FA_E_ARG, "null argument");
      std::lock_guard<std::mutex> lk(f->ref->ctx->mu);
      DeviceGuard g(f->ref->ctx->device);
      return align_noalloc(f, prm, x_inout, stats);
   } catch (...) {
      return fa_caught();
   }
}

int fa_batch_create(fa_frame* const* frames, int32_t n, const fa_tr_params* prm, fa_batch** out)
{
   try {
      if (!out || !frames || n <= 0 || !frames[0] || !frames[0]->ref) return set_err(FA_E_ARG, "invalid batch");
      fa_context* ctx = frames[0]->ref->ctx;
      std::lock_guard<std::mutex> lk(ctx->mu);
      DeviceGuard g(ctx->device);
      return batch_create_split(frames, n, prm, out);
   } catch (...) {
      return fa_caught();
   }
}

int fa_batch_destroy(fa_batch* b)
{
   try {
      if (!b) return FA_OK;
      std::lock_guard<std::mutex> lk(b->ctx->mu);
      DeviceGuard g(b->ctx->device);
      delete b;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_batch_set_timing(fa_batch* b, int32_t enable)
{
   try {
      if (!b) return set_err(FA_E_ARG, "null batch");
      b->timing = enable ? 1 : 0;
      for (fa_batch* q : b->parts) q->timing = b->timing;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_batch_set_trace(fa_batch* b, int32_t max_records)
{
   try {
      if (!b) return set_err(FA_E_ARG, "null batch");
      if (max_records < 0 || max_records > 100000) return set_err(FA_E_ARG, "max_records out of range");
      DeviceGuard g(b->ctx->device);
      const int rec = kTraceFields + 3 * max_params(b);
      if (b->parts.empty()) return leaf_set_trace(b, max_records, rec);
      for (fa_batch* q : b->parts) CHK(leaf_set_trace(q, max_records, rec));
      b->trace_cap = max_records;
      b->trace_rec = max_records > 0 ? rec : 0;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_batch_get_trace(fa_batch* b, double* out)
{
   try {
      if (!b || !out) return set_err(FA_E_ARG, "null argument");
      if (b->trace_cap <= 0) return set_err(FA_E_STATE, "tracing is off");
      DeviceGuard g(b->ctx->device);
      if (b->parts.empty()) return leaf_get_trace(b, out, 0, b->trace_cap, b->trace_rec);
      for (size_t i = 0; i < b->parts.size(); i++)
         CHK(leaf_get_trace(b->parts[i], out, b->part_first[i], b->trace_cap, b->trace_rec));
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_batch_get_tuning(const fa_batch* b, fa_tuning* t)
{
   try {
      if (!b || !t) return set_err(FA_E_ARG, "null argument");
      *t = b->tune;
      const fa_batch* leaf = b->parts.empty() ? b : b->parts[0];
      bool any = false;
      for (int l : leaf->trc_lds) any = any || l > 0;
      t->tr_lds = b->tune.tr_lds && any;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_batch_set_tuning(fa_batch* b, const fa_tuning* t)
{
   try {
      if (!b || !t) return set_err(FA_E_ARG, "null argument");
      if (t->lds_floats < 0 || t->lds_floats > 15360 || t->blocks_per_cu < 1 || t->blocks_per_cu > 64 ||
          t->check_every < 1 || t->check_every > 64 || t->tr_lds < 0 || t->tr_lds > 1)
         return set_err(FA_E_ARG, "tuning out of range (lds_floats <= 15360, 1 <= blocks_per_cu <= 64)");
      std::lock_guard<std::mutex> lk(b->mu);
      b->tune = *t;
      if (b->parts.empty()) {
         set_grids(b);
         b->drop_graph();
      }
      for (fa_batch* q : b->parts) {
         q->tune = *t;
         set_grids(q);
         q->drop_graph();
      }
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_batch_run(fa_batch* b, double* x_inout, fa_align_stats* stats, fa_batch_timing* timing)
{
   try {
      if (!b || !x_inout) return set_err(FA_E_ARG, "null argument");
      std::lock_guard<std::mutex> lk(b->mu);
      DeviceGuard g(b->ctx->device);
      return batch_run(b, x_inout, stats, timing);
   } catch (...) {
      return fa_caught();
   }
}

int fa_batch_eval(fa_batch* b, const double* D, double* err, double* jac)
{
   try {
      if (!b || !D) return set_err(FA_E_ARG, "null argument");
      std::lock_guard<std::mutex> lk(b->mu);
      DeviceGuard g(b->ctx->device);
      return batch_eval(b, D, err, jac);
   } catch (...) {
      return fa_caught();
   }
}

int fa_align_batch(fa_frame* const* frames, int32_t n, const fa_tr_params* prm, double* x_inout,
                   fa_align_stats* stats)
{
   try {
      fa_batch* b = nullptr;
      CHK(fa_batch_create(frames, n, prm, &b));
      int rc = fa_batch_run(b, x_inout, stats, nullptr);
      fa_batch_destroy(b);
      return rc;
   } catch (...) {
      return fa_caught();
   }
}

int fa_context_info(const fa_context* ctx, int32_t* o)
{
   try {
      if (!ctx || !o) return set_err(FA_E_ARG, "null argument");
      o[0] = ctx->device;
      o[1] = ctx->n_cu;
      o[2] = ctx->lds_per_block;
      o[3] = 64;
      o[4] = ctx->lds_config;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_context_mem_info(fa_context* ctx, uint64_t* free_bytes, uint64_t* total_bytes)
{
   try {
      if (!ctx || !free_bytes || !total_bytes) return set_err(FA_E_ARG, "null argument");
      DeviceGuard g(ctx->device);
      size_t f = 0, t = 0;
      HIPCHK(hipMemGetInfo(&f, &t));
      *free_bytes = f;
      *total_bytes = t;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_server_create(fa_context* ctx, fa_reference* const* refs, int32_t n_refs, int32_t slots,
                     uint64_t arena_bytes, const fa_tr_params* prm, fa_server** out)
{
   try {
      if (!ctx || !refs || n_refs <= 0 || slots <= 0 || !out) return set_err(FA_E_ARG, "invalid server arguments");
      *out = nullptr;
      for (int i = 0; i < n_refs; i++)
         if (!refs[i] || refs[i]->ctx != ctx) return set_err(FA_E_ARG, "reference of another context");
      std::lock_guard<std::mutex> lk(ctx->mu);
      DeviceGuard g(ctx->device);
      auto* sv = new fa_server();
      sv->ctx = ctx;
      size_t largest = 0;
      for (int i = 0; i < n_refs; i++)
         if (std::find(sv->refs.begin(), sv->refs.end(), refs[i]) == sv->refs.end()) sv->refs.push_back(refs[i]);
      for (fa_reference* r : sv->refs) {
         sv->nmax = std::max(sv->nmax, r->nD * r->n_dim);
         largest = std::max(largest, fa_server::round_up(frame_bytes(r, FA_F32)));
      }
      // lanes: by kind, then pixels; a new lane where a reference has more than kLaneSpan times
      // the pixels of its lane's first; past kMaxLanes, the largest references share the last
      {
         std::vector<fa_reference*> rs = sv->refs;
         std::stable_sort(rs.begin(), rs.end(), [](const fa_reference* a, const fa_reference* c) {
            if (a->n_dim != c->n_dim) return a->n_dim < c->n_dim;
            return a->n_px < c->n_px;
         });
         for (fa_reference* r : rs) {
            if (sv->lanes.empty() || ((int)sv->lanes.size() < kMaxLanes &&
                                      (sv->lanes.back().refs.front()->n_dim != r->n_dim ||
                                       r->n_px > kLaneSpan * sv->lanes.back().refs.front()->n_px)))
               sv->lanes.emplace_back();
            sv->lanes.back().refs.push_back(r);
         }
         sv->ref_lane.resize(sv->refs.size());
         for (size_t i = 0; i < sv->refs.size(); i++)
            for (int l = 0; l < (int)sv->lanes.size(); l++)
               if (std::find(sv->lanes[l].refs.begin(), sv->lanes[l].refs.end(), sv->refs[i]) != sv->lanes[l].refs.end())
                  sv->ref_lane[i] = l;
         for (auto& L : sv->lanes) order_refs(L.refs);  // the slot batch's launch order
      }
      sv->cap = slots;
      // every lane can hold `slots` frames; a lone lane with enough slots runs as two parts
      const int P1 = sv->lanes.size() == 1 && slots >= kServerSplitSlots ? std::min(kServerParts, slots / 8) : 1;
      int rc = FA_OK;
      for (int l = 0, j = 0; l < (int)sv->lanes.size() && rc == FA_OK; l++) {
         fa_server::Lane& L = sv->lanes[l];
         L.part_lo = (int)sv->parts.size();
         for (int pi = 0; pi < P1 && rc == FA_OK; pi++) {
            const int sp = slots / P1 + (pi < slots % P1 ? 1 : 0);
            fa_batch* q = nullptr;
            // part k of the server runs on the context's lane stream k
            rc = slot_batch_build(ctx, L.refs, sp, prm, ctx->lane_stream[sv->parts.size() % fa_context::kLaneStreams], &q);
            if (rc != FA_OK) break;
            sv->parts.push_back(q);
            sv->first.push_back(j);
            sv->part_lane.push_back(l);
            j += sp;
         }
         L.part_hi = (int)sv->parts.size();
      }
      sv->S = 0;
      for (fa_batch* q : sv->parts) sv->S += q->n_slots;
      // the frame arena: by default room for every slot's frame, within the device's free memory
      // less a margin (GpuFrameWarper.cpp:113-130 sizes its frame count the same way)
      if (rc == FA_OK) {
         size_t free_b = 0, total_b = 0;
         if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            (void)hipGetLastError();
            rc = set_err(FA_E_DEVICE, "hipMemGetInfo failed");
         } else {
            const size_t avail = free_b > kArenaMargin ? free_b - kArenaMargin : 0;
            size_t want = arena_bytes ? fa_server::round_up(arena_bytes) : largest * (size_t)slots;
            if (!arena_bytes) want = std::min(want, avail / 256 * 256);
            if (want > avail || (!arena_bytes && want < largest))
               rc = set_err(FA_E_OOM, "Not enough video memory to use GPU");
            sv->arena_bytes = want;
         }
      }
      const int S = sv->S;
      if (rc == FA_OK)
         if ((rc = halloc(&sv->h_bf, S)) || (rc = halloc(&sv->h_x, (size_t)S * sv->nmax)) ||
             (rc = halloc(&sv->h_slots, S)) || (rc = halloc(&sv->h_state, S)) ||
             (rc = halloc(&sv->h_xall, (size_t)S * sv->nmax)) || (rc = dalloc(&sv->d_arena, sv->arena_bytes))) {
         }
      if (rc == FA_OK &&
          (hipHostGetDevicePointer((void**)&sv->dh_bf, sv->h_bf, 0) != hipSuccess ||
           hipHostGetDevicePointer((void**)&sv->dh_x, sv->h_x, 0) != hipSuccess ||
           hipHostGetDevicePointer((void**)&sv->dh_slots, sv->h_slots, 0) != hipSuccess ||
           hipHostGetDevicePointer((void**)&sv->dh_state, sv->h_state, 0) != hipSuccess ||
           hipHostGetDevicePointer((void**)&sv->dh_xall, sv->h_xall, 0) != hipSuccess)) {
         (void)hipGetLastError();
         rc = set_err(FA_E_DEVICE, "pinned staging is not mapped for the device");
      }
      sv->cstream = ctx->stream;  // uploads: the context's stream, whose queue no lane shares
      for (int j = 0; rc == FA_OK && j < S + (int)sv->parts.size(); j++) {
         hipEvent_t e;
         if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            rc = set_err(FA_E_DEVICE, "event creation failed");
            break;
         }
         if (j < S) sv->up_ev.push_back(e);
         else sv->done_ev.push_back(e);
      }
      if (rc != FA_OK) {
         delete sv;
         return rc;
      }
      sv->free_blocks[0] = sv->arena_bytes;
      sv->slot_req.resize(S);
      sv->sstate.assign(S, fa_server::kFree);
      sv->blk_off.assign(S, 0);
      sv->blk_size.assign(S, 0);
      sv->launched.assign(sv->parts.size(), 0);
      sv->free_slots.assign(sv->parts.size(), {});
      for (int j = S - 1; j >= 0; j--) sv->free_slots[sv->part_of(j)].push_back(j);
      for (fa_reference* r : sv->refs) r->live_frames++;  // the references outlive the server
      *out = sv;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_server_destroy(fa_server* sv)
{
   try {
      if (!sv) return FA_OK;
      std::lock_guard<std::mutex> lk(sv->ctx->mu);
      DeviceGuard g(sv->ctx->device);
      for (fa_reference* r : sv->refs) r->live_frames--;
      delete sv;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_server_submit(fa_server* sv, fa_frame* frame, double* x_inout, int64_t tag)
{
   try {
      if (!sv || !frame || !x_inout) return set_err(FA_E_ARG, "null argument");
      if (std::find(sv->refs.begin(), sv->refs.end(), frame->ref) == sv->refs.end())
         return set_err(FA_E_ARG, "frame of a reference the server does not serve");
      std::lock_guard<std::mutex> lk(sv->ctx->mu);
      fa_server::Req r;
      r.f = frame;
      r.x = x_inout;
      r.tag = tag;
      r.seq = sv->next_seq++;
      sv->lanes[sv->lane_of(frame->ref)].pending.push_back(r);
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_server_submit_host(fa_server* sv, fa_reference* ref, const void* data, int32_t dtype,
                          float add_offset, double* x_inout, int64_t tag)
{
   try {
      if (!sv || !ref || !data || !x_inout) return set_err(FA_E_ARG, "null argument");
      if (dtype != FA_F32 && dtype != FA_U16) return set_err(FA_E_ARG, "unsupported frame type");
      if (std::find(sv->refs.begin(), sv->refs.end(), ref) == sv->refs.end())
         return set_err(FA_E_ARG, "frame of a reference the server does not serve");
      if (fa_server::round_up(frame_bytes(ref, dtype)) > sv->arena_bytes)
         return set_err(FA_E_OOM, "frame larger than the server's frame memory");
      std::lock_guard<std::mutex> lk(sv->ctx->mu);
      fa_server::Req r;
      r.ref = ref;
      r.host = data;
      r.dtype = dtype;
      r.offset = add_offset;
      r.x = x_inout;
      r.tag = tag;
      r.seq = sv->next_seq++;
      sv->lanes[sv->lane_of(ref)].pending.push_back(r);
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_server_step(fa_server* sv, int32_t rounds, int32_t* in_flight, int32_t* queued)
{
   try {
      if (!sv || rounds < 1) return set_err(FA_E_ARG, "invalid argument");
      std::lock_guard<std::mutex> lk(sv->ctx->mu);
      DeviceGuard g(sv->ctx->device);
      CHK(server_step(sv, rounds));
      if (in_flight) {
         int n = 0;
         for (char c : sv->sstate) n += c != fa_server::kFree;
         *in_flight = n;
      }
      if (queued) *queued = (int32_t)sv->queued();
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_server_memory(const fa_server* sv, uint64_t* used, uint64_t* total)
{
   try {
      if (!sv || !used || !total) return set_err(FA_E_ARG, "null argument");
      *used = sv->arena_used;
      *total = sv->arena_bytes;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_host_alloc(uint64_t bytes, void** out)
{
   try {
      if (!out) return set_err(FA_E_ARG, "null argument");
      *out = nullptr;
      if (hipHostMalloc(out, std::max<uint64_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) {
         (void)hipGetLastError();
         *out = nullptr;
         return set_err(FA_E_OOM, "Not enough memory to allocate");
      }
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_host_free(void* p)
{
   try {
      if (p) HIPCHK(hipHostFree(p));
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_server_results(fa_server* sv, fa_server_result* out, int32_t max, int32_t* n_out)
{
   try {
      if (!sv || !n_out || (max > 0 && !out)) return set_err(FA_E_ARG, "null argument");
      int n = 0;
      while (n < max && !sv->results.empty()) {
         out[n++] = sv->results.front();
         sv->results.pop_front();
      }
      *n_out = n;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_server_run_stream(fa_server* sv, int64_t n, const double* arrival_s, fa_reference* const* refs,
                         const void* const* data, const int32_t* dtype, const float* add_offset,
                         double* x_inout, const int64_t* x_offset, int32_t rounds, double* latency_s,
                         fa_align_stats* stats, fa_stream_counts* counts)
{
   try {
      if (!sv || n < 0 || rounds < 1 || (n > 0 && (!arrival_s || !refs || !data || !dtype || !add_offset ||
                                                   !x_inout || !x_offset || !latency_s || !stats)))
         return set_err(FA_E_ARG, "invalid argument");
      for (int64_t i = 1; i < n; i++)
         if (arrival_s[i] < arrival_s[i - 1]) return set_err(FA_E_ARG, "arrivals out of order");
      // results are matched to arrivals by tag, so frames submitted before the call (queued,
      // admitted or with results not yet taken) would be taken for arrivals of this stream
      if (sv->admitted > 0 || sv->queued() > 0 || !sv->results.empty())
         return set_err(FA_E_ARG, "the server holds frames submitted before the stream");
      using clock = std::chrono::steady_clock;
      const auto t0 = clock::now();
      auto since = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };
      fa_stream_counts c{};
      std::vector<fa_server_result> buf(256);
      int64_t next = 0, done = 0, busy = 0;
      // a frame that cannot be submitted ends the stream: the frames already submitted are run
      // to their results (the server then holds nothing that points into the caller's arrays),
      // and its error is returned
      int stop_rc = FA_OK;
      std::string stop_msg;
      while (done < next || (stop_rc == FA_OK && done < n)) {
         const double now = since();
         while (stop_rc == FA_OK && next < n && arrival_s[next] <= now) {
            const int rc = fa_server_submit_host(sv, refs[next], data[next], dtype[next], add_offset[next],
                                                 x_inout + x_offset[next], next);
            if (rc != FA_OK) {
               stop_rc = rc;
               stop_msg = fa_last_error();
               break;
            }
            next++;
            busy++;
         }
         if (busy == 0) {
            if (stop_rc != FA_OK) break;
            // nothing submitted and not done: wait for the next arrival
            std::this_thread::sleep_for(std::chrono::duration<double>(std::max(0.0, arrival_s[next] - since())));
            continue;
         }
         int32_t in_flight = 0, queued = 0;
         CHK(fa_server_step(sv, rounds, &in_flight, &queued));
         c.steps++;
         c.max_in_flight = std::max<int64_t>(c.max_in_flight, in_flight);
         c.max_queued = std::max<int64_t>(c.max_queued, queued);
         for (;;) {
            int32_t got = 0;
            CHK(fa_server_results(sv, buf.data(), (int32_t)buf.size(), &got));
            const double t = since();
            for (int32_t k = 0; k < got; k++) {
               const int64_t i = buf[k].tag;
               if (i < 0 || i >= next) return set_err(FA_E_STATE, "result of a frame not in the stream");
               latency_s[i] = t - arrival_s[i];
               stats[i] = buf[k].stats;
               done++;
               busy--;
            }
            if (got < (int32_t)buf.size()) break;
         }
      }
      c.wall_s = since();
      if (counts) *counts = c;
      if (stop_rc != FA_OK) return set_err(stop_rc, stop_msg);
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

}  // extern "C"
