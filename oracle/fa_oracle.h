/*
 * CPU oracle for the frame-aligner warp-registration hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP product
 * path is compared against; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it. Nothing in the product links it.
 *
 * It restates, in plain C, the semantics of the reference's CPU warper
 * (flimfit/frame-aligner, FrameAligner/FrameWarper.cpp, Sobel3d.cpp,
 * ImageScanParameters.h, OptimisationModel.cpp) and of the dlib trust-region
 * loop the reference calls (FrameWarpAligner.cpp:205-209).
 *
 * Two arithmetic modes share one code path:
 *   fp64 == 0  "reference-faithful": every value the reference stores or
 *              computes in float (Sobel passes, Df, steepest-descent images,
 *              warped pixels, residuals, float*float products) is rounded to
 *              float at the same point; accumulations are double, as in the
 *              reference. (A float op emulated as a double op followed by a
 *              rounding to float is exact for + - * /, since 53 >= 2*24+2.)
 *   fp64 == 1  the same algorithm with no float rounding anywhere.
 *
 * Parity pins (tests/test_oracle_pins.py, tests/test_tr_loop_pins.py): closed
 * forms for the scan timing, node map and Sobel ramps; brute-force per-pixel
 * JtJ / Jtr sums; exact warp identities (zero, integer and half-pixel shifts);
 * SSD zero at the truth; trust-region subproblem solutions against the
 * secular equation; the trust-region loop's control flow (thresholds, radius
 * factors and cap, expansion rule, stale steps, collapse, stops) on
 * closed-form models; dense vs block-tridiagonal solver agreement.
 * Parity unpinned: the lambda safeguarding inside dlib's subproblem solver in
 * several dimensions and the radius at which dlib gives up (dlib is not
 * available here and its version is not pinned by the reference), and the
 * -ffast-math reorderings of the reference's Release build.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_warper orc_warper;

/* Trust-region controls (dlib find_min_trust_region + objective_delta_stop_strategy). */
typedef struct {
   double radius;       /* initial radius; the reference passes 40 (FrameWarpAligner.cpp:208) */
   double min_delta;    /* stop when |f - f_prev| < min_delta (reference: 1e-8) */
   double rel_delta;    /* stop when |f - f_prev| < rel_delta*|f_prev|; 0 disables */
   int max_iter;        /* stop-strategy iteration cap; 0 = unlimited (reference) */
   int max_evals;       /* hard cap on objective evaluations; 0 = unlimited */
   double sub_eps;      /* subproblem boundary tolerance (dlib: 0.1) */
   int sub_max_iter;    /* subproblem iteration cap (dlib: 20) */
   int skip_repeat;     /* 1: do not re-evaluate a rejected step that would be
                           proposed again unchanged (same trajectory, fewer evaluations) */
   int dense;           /* 1: factorise B as a dense n x n matrix at every subproblem iteration,
                           as dlib does with the general_matrix the reference passes
                           (OptimisationModel.cpp:55-63); 0: block-tridiagonal (same result,
                           O(n) instead of O(n^3)) */
} orc_tr_params;

typedef struct {
   int iterations;      /* stop-strategy calls */
   int evaluations;     /* objective evaluations, including the first */
   int rejected;        /* rejected steps */
   int status;          /* stop reason, ORC_STOP_* */
   double f_final;
   double radius_final;
} orc_tr_stats;

enum {
   ORC_STOP_ABS_DELTA = 1,
   ORC_STOP_REL_DELTA = 2,
   ORC_STOP_MAX_ITER = 3,
   ORC_STOP_NO_PROGRESS = 4,
   ORC_STOP_NONFINITE = 5,
   ORC_STOP_RADIUS = 6,
   ORC_STOP_MAX_EVALS = 7
};

const char* orc_last_error(void);

orc_warper* orc_create(int fp64);
void orc_destroy(orc_warper* w);

/* AbstractFrameWarper::setReference (FrameWarper.cpp:186-205). ref is (nz,ny,nx), x fastest,
   already smoothed and offset by the caller as the reference's aligner does. */
int orc_set_reference(orc_warper* w, const float* ref, int nz, int ny, int nx, int nD,
                      double line_duration, double interline_duration, double interframe_duration);

int orc_n_dim(const orc_warper* w);
/* Derived scan timing after ImageScanParameters' ctor and processScanParameters:
   out[0..5] = line, interline, pixel, frame, interframe, stack durations. */
void orc_get_scan(const orc_warper* w, double* out6);
void orc_get_node_map(const orc_warper* w, uint16_t* Di, double* Df, int32_t* D_range);
void orc_get_gradient(const orc_warper* w, double* gx, double* gy, double* gz);
void orc_get_vi_ranges(const orc_warper* w, int64_t* first_last);
void orc_get_hessian(const orc_warper* w, double* H);

/* CpuFrameWarper::getError (FrameWarper.cpp:169-176). D is nD x 3 (x,y,z). Optional outputs. */
double orc_get_error(const orc_warper* w, const float* frame, const double* D,
                     double* err_img, double* wimg);
/* Jtr at D: getError followed by getJacobian (FrameWarper.cpp:147-167,178-182). */
double orc_get_jacobian(const orc_warper* w, const float* frame, const double* D, double* jac);
/* CpuFrameWarper::warpImageInterpolated (FrameWarper.cpp:31-69). */
void orc_warp_interpolated(const orc_warper* w, const float* frame, const double* D,
                           double invalid_value, double* out);
/* AbstractFrameWarper::warpImage forward splat with coverage (FrameWarper.cpp:71-136). */
void orc_warp_image(const orc_warper* w, const float* frame, const double* D,
                    double* wimg, double* coverage);
/* AbstractFrameWarper::warpPoint (FrameWarper.cpp:397-419). */
void orc_warp_point(const orc_warper* w, const double* D, int x, int y, int z,
                    int spatial_binning, double* out3);

void orc_tr_default(orc_tr_params* p);
/* Trust-region alignment of one frame from x_inout (nD*n_dim, D2col layout). */
int orc_align(const orc_warper* w, const float* frame, const orc_tr_params* prm,
              double* x_inout, orc_tr_stats* stats);

/* Per-evaluation trace of the trust region. Record r describes objective evaluation r + 1
   (r = 0: the start point) and the state after it was processed: the decision on it, the stop
   check that followed and the next subproblem. Each record is ORC_TRACE_FIELDS doubles, then
   x, g and p (n = nD*n_dim each) after the record's update. */
enum {
   ORC_T_EVALS = 0,   /* evaluations so far (r + 1) */
   ORC_T_FNEW = 1,    /* f of this evaluation */
   ORC_T_F = 2,       /* f at the current x after the decision */
   ORC_T_PRED = 3,    /* predicted decrease of the step evaluated (0 for the start) */
   ORC_T_MEAS = 4,    /* measured decrease */
   ORC_T_RHO = 5,     /* meas / |pred| */
   ORC_T_ACCEPT = 6,  /* 1 if the evaluated point became x */
   ORC_T_RADIUS = 7,  /* radius after the decision */
   ORC_T_ITER = 8,    /* stop-strategy iterations after the stop check */
   ORC_T_STATUS = 9,  /* ORC_STOP_* if the loop ended here, else 0 */
   ORC_T_LAMBDA = 10, /* lambda of the next subproblem (0 if none ran) */
   ORC_T_SUBIT = 11,  /* iterations of the next subproblem (0 if none ran) */
   ORC_T_PNORM = 12,  /* |p| of the next step (0 if none) */
   ORC_T_GNORM = 13,  /* |g| at x */
   ORC_T_GP = 14,     /* g'p of the next step */
   ORC_T_REJ = 15,    /* rejected steps so far */
   ORC_TRACE_FIELDS = 16
};
/* orc_align with a trace of at most cap records (trace: cap x (ORC_TRACE_FIELDS + 3n)). */
int orc_align_trace(const orc_warper* w, const float* frame, const orc_tr_params* prm,
                    double* x_inout, orc_tr_stats* stats, double* trace, int cap);

/* The trust-region loop over any model of the OptimisationModel concept
   (OptimisationModel.h:6-27) with a constant block-tridiagonal B, for tests that drive the
   loop's control flow with closed-form objectives. value(x) returns f(x); gradient(g) gives
   the derivative at the point of the last value() call whose point was accepted (the loop
   calls it right after such a value()). */
typedef struct orc_tr_model {
   int nb, m;          /* n = nb*m parameters; B: nb diagonal blocks of m x m */
   const double* diag; /* nb x m*m, row-major */
   const double* sub;  /* nb-1 blocks (k+1, k), m*m each */
   double (*value)(void* user, const double* x);
   void (*gradient)(void* user, double* g);
   void* user;
} orc_tr_model;
int orc_tr_run(const orc_tr_model* model, const orc_tr_params* prm, double* x_inout,
               orc_tr_stats* stats, double* trace, int cap);
/* Aligns n frames of n_px each with n_threads workers (one frame per task). */
int orc_align_many(const orc_warper* w, const float* frames, int n_frames,
                   const orc_tr_params* prm, double* x_inout, orc_tr_stats* stats,
                   double* seconds_per_frame, int n_threads);

/* Linear algebra used by the trust region, exposed for tests.
   Block-tridiagonal storage: diag[nb][m*m], sub[nb-1][m*m] (block (k+1,k)), row-major. */
int orc_bt_cholesky(int nb, int m, const double* diag, const double* sub, double lambda,
                    double* Ldiag, double* Lsub);
void orc_bt_solve(int nb, int m, const double* Ldiag, const double* Lsub, const double* b, double* x);
int orc_tr_subproblem_bt(int nb, int m, const double* diag, const double* sub, const double* g,
                         double radius, double eps, int max_iter, double* p, double* lambda_out);
int orc_tr_subproblem_dense(int n, const double* B, const double* g, double radius, double eps,
                            int max_iter, double* p, double* lambda_out);

#ifdef __cplusplus
}
#endif
