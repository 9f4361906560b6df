// HipFrameWarper: the reference's warper interface (FrameAligner/FrameWarper.h:95-143,
// FrameAligner/GpuFrameWarper.h:31-72) in C++ over the C ABI of include/frame_aligner.h.
//
// Header-only and dependency-free: volumes are plain float/uint16 buffers with (nz, ny, nx)
// dims instead of cv::Mat, node displacements are std::vector<Point3d>, parameter vectors
// std::vector<double> in the D2col layout (OptimisationModel.cpp:14-27). Frames are keyed
// by their data pointer, as GpuFrameWarper keys them by cv::Mat::data. Status codes become
// std::runtime_error carrying the reference's messages.
#pragma once

#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "frame_aligner.h"

namespace fa_host {

struct Point3d {
   double x = 0, y = 0, z = 0;
};

// ImageScanParameters (ImageScanParameters.h:7-20): the constructor arguments.
struct ImageScanParameters {
   double line_duration = 100, interline_duration = 101, interframe_duration = 102;
   int n_x = 1, n_y = 1, n_z = 1;
};

inline void check(int rc)
{
   if (rc != FA_OK) throw std::runtime_error(fa_last_error());
}

// D2col / col2D (OptimisationModel.cpp:14-42).
inline std::vector<double> D2col(const std::vector<Point3d>& D, int n_dim)
{
   std::vector<double> c;
   c.reserve(D.size() * n_dim);
   for (const auto& p : D) {
      c.push_back(p.x);
      c.push_back(p.y);
      if (n_dim == 3) c.push_back(p.z);
   }
   return c;
}

inline std::vector<Point3d> col2D(const std::vector<double>& c, int n_dim)
{
   std::vector<Point3d> D(c.size() / n_dim);
   for (size_t i = 0; i < D.size(); i++) {
      D[i].x = c[i * n_dim];
      D[i].y = c[i * n_dim + 1];
      if (n_dim == 3) D[i].z = c[i * n_dim + 2];
   }
   return D;
}

class HipFrameWarper {
 public:
   explicit HipFrameWarper(int device = 0) { check(fa_create(device, &ctx_)); }
   HipFrameWarper(const HipFrameWarper&) = delete;
   HipFrameWarper& operator=(const HipFrameWarper&) = delete;
   ~HipFrameWarper()
   {
      release_frames();
      if (ref_) fa_reference_release(ref_);
      fa_destroy(ctx_);
   }

   // AbstractFrameWarper::setReference (FrameWarper.cpp:186-205). reference: nz*ny*nx floats.
   void setReference(const float* reference, const std::vector<int>& dims_zyx, int nD,
                     const ImageScanParameters& p)
   {
      if (dims_zyx.size() != 3) throw std::runtime_error("Reference must be three dimensional");
      release_frames();
      if (ref_) check(fa_reference_release(ref_));
      ref_ = nullptr;
      fa_scan_params s{p.line_duration, p.interline_duration, p.interframe_duration, p.n_x, p.n_y, p.n_z, 0};
      check(fa_set_reference(ctx_, reference, dims_zyx[0], dims_zyx[1], dims_zyx[2], nD, &s, 0, &ref_));
      dims = dims_zyx;
      this->nD = nD;
      n_dim = dims[0] > 1 ? 3 : 2;
   }

   // registerFrame / deregisterFrame (GpuFrameWarper.cpp:46-73). Frames are already offset
   // (as the reference's aligner passes them) unless add_offset is given.
   void registerFrame(const float* frame, float add_offset = 0.f) { add(frame, FA_F32, add_offset); }
   void registerFrame(const uint16_t* frame, float add_offset = 1.f) { add(frame, FA_U16, add_offset); }
   void deregisterFrame(const void* frame)
   {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = frames_.find(frame);
      if (it == frames_.end()) return;
      fa_deregister_frame(it->second);
      frames_.erase(it);
   }

   double getError(const void* frame, const std::vector<Point3d>& D)
   {
      double e = 0;
      auto Dv = flat(D);
      check(fa_get_error(get(frame), Dv.data(), &e));
      return e;
   }

   void getJacobian(const void* frame, const std::vector<Point3d>& D, std::vector<double>& jac)
   {
      jac.assign((size_t)nD * n_dim, 0.0);
      auto Dv = flat(D);
      check(fa_get_jacobian(get(frame), Dv.data(), jac.data(), nullptr));
   }

   void warpImageInterpolated(const void* frame, std::vector<float>& wimg,
                              const std::vector<Point3d>& D, int invalid_value = 0)
   {
      wimg.assign((size_t)dims[0] * dims[1] * dims[2], 0.f);
      auto Dv = flat(D);
      check(fa_warp_interpolated(get(frame), Dv.data(), wimg.data(), (float)invalid_value));
   }

   Point3d warpPoint(const std::vector<Point3d>& D, int x, int y, int z, int spatial_binning = 1)
   {
      auto Dv = flat(D);
      double o[3];
      check(fa_warp_point(ref_, Dv.data(), x, y, z, spatial_binning, o));
      return Point3d{o[0], o[1], o[2]};
   }

   // The trust-region call of FrameWarpAligner::addFrame (FrameWarpAligner.cpp:200-215).
   fa_align_stats align(const void* frame, std::vector<double>& x, const fa_tr_params* prm = nullptr)
   {
      fa_align_stats st{};
      x.resize((size_t)nD * n_dim, 0.0);
      check(fa_align(get(frame), prm, x.data(), &st));
      return st;
   }

   // AbstractFrameWarper::H, dense.
   std::vector<double> H() const
   {
      const size_t n = (size_t)nD * n_dim;
      std::vector<double> h(n * n);
      check(fa_get_hessian(ref_, h.data()));
      return h;
   }

   // GpuFrameWarper::clearTemp (GpuFrameWarper.cpp:154-158): frees the cached working spaces;
   // registered frames stay registered.
   void clearTemp()
   {
      std::lock_guard<std::mutex> lk(mu_);
      if (ref_) check(fa_clear_temp(ref_));
   }

   int n_dim = 2;
   int nD = 0;
   std::vector<int> dims;

 private:
   void release_frames()
   {
      std::lock_guard<std::mutex> lk(mu_);
      for (auto& kv : frames_) fa_deregister_frame(kv.second);
      frames_.clear();
   }

   void add(const void* frame, int dtype, float off)
   {
      if (!ref_) throw std::runtime_error("Reference imaging not yet set");
      fa_frame* f = nullptr;
      check(fa_register_frame(ref_, frame, dtype, off, &f));
      std::lock_guard<std::mutex> lk(mu_);
      auto it = frames_.find(frame);
      if (it != frames_.end()) fa_deregister_frame(it->second);
      frames_[frame] = f;
   }

   fa_frame* get(const void* frame)
   {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = frames_.find(frame);
      if (it == frames_.end()) throw std::runtime_error("Unrecognised frame");
      return it->second;
   }

   std::vector<double> flat(const std::vector<Point3d>& D) const
   {
      if ((int)D.size() != nD) throw std::runtime_error("D has the wrong number of nodes");
      std::vector<double> v(3 * D.size());
      for (size_t i = 0; i < D.size(); i++) {
         v[3 * i] = D[i].x;
         v[3 * i + 1] = D[i].y;
         v[3 * i + 2] = D[i].z;
      }
      return v;
   }

   fa_context* ctx_ = nullptr;
   fa_reference* ref_ = nullptr;
   std::mutex mu_;
   std::map<const void*, fa_frame*> frames_;
};

}  // namespace fa_host
