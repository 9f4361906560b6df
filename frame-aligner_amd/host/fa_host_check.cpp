// Drives HipFrameWarper the way FrameWarpAligner::addFrame drives a warper
// (FrameWarpAligner.cpp:117-223): setReference, registerFrame, getError, getJacobian,
// the trust region, warpImageInterpolated, deregisterFrame. Inputs and outputs are raw
// little-endian files in one directory, so a test can compare them with the oracle:
//   in:  meta.txt ("nz ny nx nD"), ref.f32, frame.f32, D.f64 (nD x 3)
//   out: err.f64 (SSD at D), jac.f64, x.f64 (aligned parameters), warp.f32, stats.txt
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "HipFrameWarper.hpp"

template <typename T>
static std::vector<T> read(const std::string& path, size_t n)
{
   std::vector<T> v(n);
   std::ifstream f(path, std::ios::binary);
   if (!f.read(reinterpret_cast<char*>(v.data()), n * sizeof(T))) throw std::runtime_error("cannot read " + path);
   return v;
}

template <typename T>
static void write(const std::string& path, const std::vector<T>& v)
{
   std::ofstream f(path, std::ios::binary);
   f.write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T));
}

int main(int argc, char** argv)
{
   if (argc != 2) {
      std::cerr << "usage: fa_host_check <dir>\n";
      return 2;
   }
   const std::string dir = argv[1];
   try {
      int nz, ny, nx, nD;
      std::ifstream meta(dir + "/meta.txt");
      meta >> nz >> ny >> nx >> nD;
      const size_t n = (size_t)nz * ny * nx;
      auto ref = read<float>(dir + "/ref.f32", n);
      auto frame = read<float>(dir + "/frame.f32", n);
      auto Dv = read<double>(dir + "/D.f64", 3 * (size_t)nD);
      std::vector<fa_host::Point3d> D(nD);
      for (int i = 0; i < nD; i++) D[i] = {Dv[3 * i], Dv[3 * i + 1], Dv[3 * i + 2]};

      fa_host::HipFrameWarper w(0);
      fa_host::ImageScanParameters p;
      p.line_duration = 1;
      p.interline_duration = 1;
      p.interframe_duration = ny;
      p.n_x = nx;
      p.n_y = ny;
      p.n_z = nz;
      w.setReference(ref.data(), {nz, ny, nx}, nD, p);
      w.registerFrame(frame.data());
      std::vector<double> err{w.getError(frame.data(), D)};
      std::vector<double> jac;
      w.getJacobian(frame.data(), D, jac);
      fa_tr_params prm;
      fa_tr_default(&prm);
      prm.rel_delta = 1e-4;
      prm.max_iter = 50;
      prm.skip_repeat = 1;
      std::vector<double> x;
      fa_align_stats st = w.align(frame.data(), x, &prm);
      std::vector<float> warped;
      w.warpImageInterpolated(frame.data(), warped, fa_host::col2D(x, w.n_dim));
      w.deregisterFrame(frame.data());
      bool unknown_rejected = false;
      try {
         w.getError(frame.data(), D);
      } catch (const std::runtime_error& e) {
         unknown_rejected = std::string(e.what()) == "Unrecognised frame";
      }
      write(dir + "/err.f64", err);
      write(dir + "/jac.f64", jac);
      write(dir + "/x.f64", x);
      write(dir + "/warp.f32", warped);
      std::ofstream s(dir + "/stats.txt");
      s << st.evaluations << " " << st.status << " " << (unknown_rejected ? 1 : 0) << "\n";
   } catch (const std::exception& e) {
      std::cerr << "fa_host_check: " << e.what() << "\n";
      return 1;
   }
   return 0;
}
