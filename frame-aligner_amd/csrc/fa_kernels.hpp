// Launchers of the gfx950 kernels (fa_kernels.hip).
#pragma once

#include "fa_internal.hpp"

namespace fa {

hipError_t launch_node_map(const ScanDev& sc, int nx, int ny, int64_t n, uint16_t* Di, float* Df,
                           hipStream_t s);
hipError_t launch_interval_begin(const uint16_t* Di, int64_t n, int nD, int32_t* begin,
                                 int32_t* flags, hipStream_t s);
hipError_t launch_reference_record(const float* T, int nz, int ny, int nx, int n_dim,
                                   const ScanDev& sc, float4* rg4, float* df, hipStream_t s);
hipError_t launch_hessian(const RefDev& R, double* partials, double* Hdiag, double* Hsub,
                          double* L0, double* info, hipStream_t s);
hipError_t launch_eigen(const RefDev& R, double* Q, double* ev, double* info, hipStream_t s);
hipError_t launch_pass(const PassArgs& a, int n_dim, int grid, hipStream_t s);
// count frames in one launch: frame i's samples, nD x 3 displacements and outputs follow the
// first's at i times their sizes
hipError_t launch_warp(const RefDev& R, const FrameDev& F, const double* D, float* out,
                       float invalid, int lds_cap, int use_lds, int grid, hipStream_t s, int count = 1);
hipError_t launch_splat(const RefDev& R, const FrameDev& F, const double* D, float* wimg, float* cov,
                        float offset, int grid, hipStream_t s, int count = 1);
hipError_t launch_schedule(const BatchDev& B, int account, hipStream_t s);
hipError_t launch_tr_init(const BatchDev& B, int eval_only, hipStream_t s);
hipError_t launch_tr_start_slots(const BatchDev& B, const BatchFrame* hbf, const double* hx, const int32_t* hslots,
                                 int n, int nmax, hipStream_t s);
hipError_t launch_publish(const BatchDev& B, TrState* hstate, double* hx, int64_t total_x, hipStream_t s);
hipError_t launch_tr_update(const BatchDev& B, int cls, int lds_doubles, int full, int threads, hipStream_t s);
hipError_t launch_tr_update_all(const BatchDev& B, int lds_doubles, uint32_t wide, uint32_t full, hipStream_t s);
int tr_update_all_per_cu(int lds_doubles);
hipError_t configure_kernels(int max_lds_bytes);
hipError_t launch_reduce_eval(const BatchDev& B, hipStream_t s);
// The subproblems the update left pending (references with RefDev::eig == 2): qv = Q^T g for
// all pending frames of a reference at once, the secular iteration per frame, p = -Q w, x + p,
// over B.eig_tiles and B.eig_frames (nothing when they are empty).
hipError_t launch_eig_subproblems(const BatchDev& B, hipStream_t s);
// Appends the product tiles of reference index ref (n = nD n_dim parameters) for up to
// `frames` active frames to tiles (4 ints each: ref, first row, first list entry, k group).
void eig_tiles_of(int ref, int n, int frames, std::vector<int32_t>& tiles);
// dst = src^T (n x n, row-major, distinct buffers)
hipError_t launch_transpose(const double* src, double* dst, int n, hipStream_t s);

}  // namespace fa
