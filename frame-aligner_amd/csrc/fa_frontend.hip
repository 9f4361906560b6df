// Front end of the warp aligner on the device: FrameWarpAligner::setReference and addFrame
// (FrameAligner/FrameWarpAligner.cpp:46-223) around the warper of fa_api.hip.
//
//  * smoothStack: per z-slice Gaussian blur, sigma x = smoothing and sigma y = 1 (the
//    reference passes 1 as cv::GaussianBlur's sigmaY), border reflect-101, row pass then column
//    pass in float; then + 1.0f.
//  * rigid start: VolumePhaseCorrelator (FrameAligner/VolumePhaseCorrelator.cpp): the raw
//    frame binned in x and y (AbstractFrameAligner.cpp:37-56), times the square-root Hann
//    window in y and x, real-to-complex 3-D FFT in FP64, cross-power spectrum with the
//    reference's conjugate spectrum normalised by |.| + DBL_MIN, inverse FFT, first maximum,
//    indices above half the size wrapped negative, x and y scaled by the binning.
//  * starts {0, 0, rigid}: the first with the smallest SSD (FrameWarpAligner.cpp:150-198),
//    then the trust region of all frames as one batch (type Warp).
// The FFTs are rocFFT plans (double precision) on the aligner's own stream.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "fa_internal.hpp"
#include "fa_kernels.hpp"

#include <cstdio>

using namespace fa;

namespace {

// Host copies between pageable caller memory and the aligner's pinned staging: pieces of
// `parts` (dst, src, bytes) triples spread over up to kCopyThreads threads (a single thread
// moves a few GB/s into freshly allocated pages; the DMA side runs at PCIe rate).
constexpr int kCopyThreads = 8;
struct CopyPiece {
   void* dst;
   const void* src;
   size_t bytes;
};

void par_copy(const std::vector<CopyPiece>& parts)
{
   size_t total = 0;
   for (const auto& c : parts) total += c.bytes;
   const int T = (int)std::max<size_t>(1, std::min<size_t>({(size_t)kCopyThreads, parts.size() * 4, total >> 21}));
   if (T <= 1) {
      for (const auto& c : parts) std::memcpy(c.dst, c.src, c.bytes);
      return;
   }
   // thread t copies bytes [t total / T, (t + 1) total / T) of the concatenated pieces
   auto run = [&](int t) {
      const size_t lo = total * t / T, hi = total * (t + 1) / T;
      size_t at = 0;
      for (const auto& c : parts) {
         const size_t a = std::max(lo, at), b = std::min(hi, at + c.bytes);
         if (a < b) std::memcpy((char*)c.dst + (a - at), (const char*)c.src + (a - at), b - a);
         at += c.bytes;
         if (at >= hi) break;
      }
   };
   // a thread that cannot be started (no exception may leave the C interface) leaves its share
   // and the ones after it to this thread
   std::vector<std::thread> th;
   int started = 1;
   try {
      th.reserve(T - 1);
      for (int t = 1; t < T; t++) {
         th.emplace_back(run, t);
         started++;
      }
   } catch (...) {
   }
   for (int t = started; t < T; t++) run(t);
   run(0);
   for (auto& x : th) x.join();
}


#define FCHK(expr)                                                                                \
   do {                                                                                           \
      hipError_t e_ = (expr);                                                                     \
      if (e_ != hipSuccess) return set_error(FA_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
   } while (0)

#define RCHK(expr)                                                                                \
   do {                                                                                           \
      if ((expr) != rocfft_status_success) return set_error(FA_E_DEVICE, std::string("rocFFT: ") + #expr); \
   } while (0)

#define OK(expr)                                                                                  \
   do {                                                                                           \
      int r_ = (expr);                                                                            \
      if (r_ != FA_OK) return r_;                                                                 \
   } while (0)

constexpr int kFB = 256;

template <typename T>
int falloc(T** p, size_t n)
{
   *p = nullptr;
   if (hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      *p = nullptr;
      return set_error(FA_E_OOM, "Not enough memory to allocate");
   }
   return FA_OK;
}

template <typename T>
void ffree(T*& p)
{
   if (p) (void)hipFree((void*)p);
   p = nullptr;
}

// ---- kernels ---------------------------------------------------------------------------------

__global__ void k_to_float(const void* raw, int dtype, int64_t n, float* out)
{
   const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   out[i] = dtype == FA_U16 ? (float)((const uint16_t*)raw)[i] : ((const float*)raw)[i];
}

__device__ __forceinline__ int reflect101(int i, int n)
{
   if (n == 1) return 0;
   const int period = 2 * n - 2;
   i = abs(i) % period;
   return i >= n ? period - i : i;
}

// Symmetric filter along x (axis = 0) or y (axis = 1) of every slice:
// out = k[r] v[i] + sum_{j=1..r} k[r+j] (v[i-j] + v[i+j]), j ascending, in float; add is
// added to the result (the aligner's +1 after the last pass).
__global__ void k_blur(const float* in, float* out, int nz, int ny, int nx, const float* k, int r, int axis,
                       float add)
{
   const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   const int64_t n = (int64_t)nz * ny * nx;
   if (i >= n) return;
   const int x = (int)(i % nx);
   const int64_t row = i / nx;
   const int y = (int)(row % ny);
   const int64_t base = row - y;  // first row of the slice
   float s;
   if (axis == 0) {
      const float* v = in + row * nx;
      s = k[r] * v[x];
      for (int j = 1; j <= r; j++) s = s + k[r + j] * (v[reflect101(x - j, nx)] + v[reflect101(x + j, nx)]);
   } else {
      const float* v = in + base * nx + x;
      s = k[r] * v[(int64_t)y * nx];
      for (int j = 1; j <= r; j++)
         s = s + k[r + j] * (v[(int64_t)reflect101(y - j, ny) * nx] + v[(int64_t)reflect101(y + j, ny) * nx]);
   }
   out[i] = add != 0.f ? s + add : s;
}

__global__ void k_add(float* v, int64_t n, float add)
{
   const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) v[i] = v[i] + add;
}

// Binned (x, y summed in bins of f, y outer and x inner, in float) volumes times the window,
// for `count` frames (inputs nz*ny*nx floats apart, outputs nz*ph*pw doubles apart).
__global__ void k_bin_window(const float* v, int nz, int ny, int nx, int f, int ph, int pw, const double* win,
                             double* out, int count)
{
   const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   const int64_t n = (int64_t)nz * ph * pw;
   if (t >= n * count) return;
   const int64_t fr = t / n, i = t - fr * n;
   const float* vf = v + fr * ((int64_t)nz * ny * nx);
   const int X = (int)(i % pw);
   const int Y = (int)((i / pw) % ph);
   const int z = (int)(i / ((int64_t)pw * ph));
   float s = 0.f;
   for (int y = Y * f; y < min(ny, Y * f + f); y++)
      for (int x = X * f; x < min(nx, X * f + f); x++) s += vf[((int64_t)z * ny + y) * nx + x];
   out[t] = (double)s * win[i];
}

__global__ void k_conj(const double2* in, double2* out, int64_t n)
{
   const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) out[i] = make_double2(in[i].x, -in[i].y);
}

// Cross-power spectra of `count` frames' spectra (n apart) with the reference's.
__global__ void k_cross_power(double2* s, const double2* ref, int64_t n, int count)
{
   const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n * count) return;
   const double2 a = s[i], b = ref[i % n];
   const double re = a.x * b.x - a.y * b.y, im = a.x * b.y + a.y * b.x;
   const double nrm = hypot(re, im) + DBL_MIN;
   s[i] = make_double2(re / nrm, im / nrm);
}

// First maximum: per workgroup (value, index), ties to the lower index; the last stage is one
// workgroup over the partials. Frame blockIdx.y of the first stage reads v + y n and writes
// partials y * gridDim.x + blockIdx.x; workgroup b of the last stage reduces frame b's m.
__device__ __forceinline__ void better(double& v, int64_t& i, double v2, int64_t i2)
{
   if (v2 > v || (v2 == v && i2 < i)) {
      v = v2;
      i = i2;
   }
}

__global__ __launch_bounds__(kFB) void k_argmax(const double* v, int64_t n, double* pv, int64_t* pi)
{
   __shared__ double sv[kFB];
   __shared__ int64_t si[kFB];
   v += (int64_t)blockIdx.y * n;
   pv += (int64_t)blockIdx.y * gridDim.x;
   pi += (int64_t)blockIdx.y * gridDim.x;
   double bv = -DBL_MAX;
   int64_t bi = INT64_MAX;
   for (int64_t i = (int64_t)blockIdx.x * kFB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kFB)
      better(bv, bi, v[i], i);
   sv[threadIdx.x] = bv;
   si[threadIdx.x] = bi;
   __syncthreads();
   for (int off = kFB / 2; off >= 1; off >>= 1) {
      if (threadIdx.x < off) better(sv[threadIdx.x], si[threadIdx.x], sv[threadIdx.x + off], si[threadIdx.x + off]);
      __syncthreads();
   }
   if (threadIdx.x == 0) {
      pv[blockIdx.x] = sv[0];
      pi[blockIdx.x] = si[0];
   }
}

__global__ __launch_bounds__(kFB) void k_argmax_final(const double* pv, const int64_t* pi, int m, int64_t* out)
{
   __shared__ double sv[kFB];
   __shared__ int64_t si[kFB];
   pv += (int64_t)blockIdx.x * m;
   pi += (int64_t)blockIdx.x * m;
   out += blockIdx.x;
   double bv = -DBL_MAX;
   int64_t bi = INT64_MAX;
   for (int i = threadIdx.x; i < m; i += kFB) better(bv, bi, pv[i], pi[i]);
   sv[threadIdx.x] = bv;
   si[threadIdx.x] = bi;
   __syncthreads();
   for (int off = kFB / 2; off >= 1; off >>= 1) {
      if (threadIdx.x < off) better(sv[threadIdx.x], si[threadIdx.x], sv[threadIdx.x + off], si[threadIdx.x + off]);
      __syncthreads();
   }
   if (threadIdx.x == 0) *out = si[0];
}

// float -> uint16 as cv::Mat::convertTo(CV_16U) does: round to nearest even, saturate to
// [0, 65535]; NaN, infinities and values beyond the int range convert to 0 (the x86 integer
// conversion's out-of-range value, then saturated).
__device__ __forceinline__ uint16_t to_u16(float v)
{
   if (!(fabsf(v) < 2147483648.f)) return 0;
   const float r = rintf(v);
   return (uint16_t)(r < 0.f ? 0.f : (r > 65535.f ? 65535.f : r));
}

__global__ void k_to_u16(const float* in, int64_t n, uint16_t* out)
{
   const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) out[i] = to_u16(in[i]);
}

// float -> uint8 as cv::Mat::convertTo(CV_8U) does (the same rounding, saturated to [0, 255]).
__device__ __forceinline__ uint8_t to_u8(float v)
{
   if (!(fabsf(v) < 2147483648.f)) return 0;
   const float r = rintf(v);
   return (uint8_t)(r < 0.f ? 0.f : (r > 255.f ? 255.f : r));
}

__global__ void k_to_u8(const float* in, int64_t n, uint8_t* out)
{
   const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) out[i] = to_u8(in[i]);
}

// getWarpedFrame: the splat divided by its coverage (float division), to uint16.
__global__ void k_preserving(const float* wimg, const float* cov, int64_t n, uint16_t* out)
{
   const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) out[i] = to_u16(wimg[i] / cov[i]);
}

// Masked moments for correlation() (Cv3dUtils.cpp:64-88): over pixels with cov != 0,
// stage 1: count, sum x, sum x^2, sum y, sum y^2 in double (y = the record's reference value);
// stage 2: sum of float((x - mx) * (y - my)) with mx, my the float means of stage 1's sums
// (s1: kMoments per frame; mx = float(sum x / count)). Frame blockIdx.y reads x and cov n
// apart; per-workgroup partials, summed in a fixed order by k_sum_partials.
constexpr int kMoments = 5;

__global__ __launch_bounds__(kFB) void k_moments(const float* x, const float4* rec, const float* cov, int64_t n,
                                                 int stage, const double* s1, double* part)
{
   __shared__ double sm[kFB][kMoments];
   const int fr = blockIdx.y;
   x += (int64_t)fr * n;
   cov += (int64_t)fr * n;
   float mx = 0.f, my = 0.f;
   if (stage == 2) {
      const double* h = s1 + (int64_t)fr * kMoments;
      if (h[0] != 0) {
         mx = (float)(h[1] / h[0]);
         my = (float)(h[3] / h[0]);
      }
   }
   double a[kMoments] = {0, 0, 0, 0, 0};
   for (int64_t i = (int64_t)blockIdx.x * kFB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kFB) {
      if (cov[i] == 0.f) continue;
      const float xv = x[i], yv = rec[i].x;
      if (stage == 1) {
         a[0] += 1.0;
         a[1] += xv;
         a[2] += (double)xv * xv;
         a[3] += yv;
         a[4] += (double)yv * yv;
      } else {
         const float dx = xv - mx, dy = yv - my;
         a[0] += (double)(dx * dy);
      }
   }
   for (int k = 0; k < kMoments; k++) sm[threadIdx.x][k] = a[k];
   __syncthreads();
   for (int off = kFB / 2; off >= 1; off >>= 1) {
      if (threadIdx.x < off)
         for (int k = 0; k < kMoments; k++) sm[threadIdx.x][k] += sm[threadIdx.x + off][k];
      __syncthreads();
   }
   if (threadIdx.x < kMoments)
      part[((int64_t)fr * gridDim.x + blockIdx.x) * kMoments + threadIdx.x] = sm[0][threadIdx.x];
}

// Frame blockIdx.x's m partials: lane l sums partials l, l + 64, ..., then a fixed tree over
// the 64 lanes.
__global__ __launch_bounds__(64) void k_sum_partials(const double* part, int m, double* out)
{
   __shared__ double sm[64][kMoments];
   part += (int64_t)blockIdx.x * m * kMoments;
   double a[kMoments] = {0, 0, 0, 0, 0};
   for (int b = threadIdx.x; b < m; b += 64)
      for (int k = 0; k < kMoments; k++) a[k] += part[(int64_t)b * kMoments + k];
   for (int k = 0; k < kMoments; k++) sm[threadIdx.x][k] = a[k];
   __syncthreads();
   for (int off = 32; off >= 1; off >>= 1) {
      if (threadIdx.x < off)
         for (int k = 0; k < kMoments; k++) sm[threadIdx.x][k] += sm[threadIdx.x + off][k];
      __syncthreads();
   }
   if (threadIdx.x < kMoments) out[(int64_t)blockIdx.x * kMoments + threadIdx.x] = sm[0][threadIdx.x];
}

// correlation() from the two stages' sums of one frame (NaN where no pixel is covered).
double correlation_of(const double* h, const double* h2)
{
   const double cnt = h[0];
   if (cnt == 0) return NAN;
   const double m1 = h[1] / cnt, m2 = h[3] / cnt;
   const double s1 = std::sqrt(std::max(h[2] / cnt - m1 * m1, 0.0)), s2 = std::sqrt(std::max(h[4] / cnt - m2 * m2, 0.0));
   return (h2[0] / cnt) / (s1 * s2);
}

unsigned grid_of(int64_t n) { return (unsigned)((n + kFB - 1) / kFB); }

// Float Gaussian weights of cv::GaussianBlur(.., Size(0, 0), sigma) on a float image:
// ksize = round(8 sigma + 1) | 1; exp(-(i - c)^2 / (2 sigma^2)) rounded to float, then scaled
// by 1 / (their double sum) and rounded again.
std::vector<float> gaussian_weights(double sigma)
{
   const int n = (int)std::nearbyint(sigma * 8 + 1) | 1;
   std::vector<float> k(n);
   const double c = (n - 1) * 0.5, s2 = -0.5 / (sigma * sigma);
   double sum = 0;
   for (int i = 0; i < n; i++) {
      k[i] = (float)std::exp(s2 * (i - c) * (i - c));
      sum += k[i];
   }
   const double inv = 1.0 / sum;
   for (int i = 0; i < n; i++) k[i] = (float)(k[i] * inv);
   return k;
}

std::vector<double> hann(int n)
{
   std::vector<double> w(n, 1.0);
   if (n > 1) {
      const double f = 2.0 * 3.14159265358979323846 / (n - 1);
      for (int i = 0; i < n; i++) w[i] = std::sqrt(0.5 * (1 - std::cos(f * i)));
   }
   return w;
}

std::once_flag g_fft_setup;

}  // namespace

struct fa_aligner {
   fa_context* ctx = nullptr;
   fa_realign_params rp{};
   fa_scan_params scan{};
   int nz = 0, ny = 0, nx = 0, nD = 0, n_dim = 0;
   int64_t n_px = 0;
   fa_reference* ref = nullptr;
   hipStream_t stream = nullptr;
   std::mutex mu;
   // phase correlation over the binned volume (pz, ph, pw)
   int pz = 0, ph = 0, pw = 0;
   int64_t n_el = 0, n_cplx = 0;
   rocfft_plan fwd = nullptr, inv = nullptr;
   rocfft_execution_info info = nullptr;
   void* work = nullptr;
   double* d_win = nullptr;
   double* d_real = nullptr;
   double2* d_spec = nullptr;
   double2* d_refspec = nullptr;
   double* d_pv = nullptr;
   int64_t* d_pi = nullptr;
   int64_t* d_idx = nullptr;
   int n_part = 0;
   // blur
   float* d_kx = nullptr;
   float* d_ky = nullptr;
   int rx = 0, ry = 0;
   float* d_raw = nullptr;  // raw frame as float
   float* d_tmp = nullptr;
   void* d_up = nullptr;    // raw upload
   // add_frames state kept between calls: its device arrays (grown when a call needs more) and
   // the frames and batch of the last call, reused while the frame count stays the same (the
   // frames are views of the preprocessed-frame array, whose address then stays the same)
   enum { kUp, kRawF, kPre, kBlurTmp, kReal, kSpec, kPv, kPi, kIdx, kD, kPart, kSums, kW, kW2, kWimg, kCov, kU16, kBufs };
   struct Buf {
      void* p = nullptr;
      size_t bytes = 0;
   };
   Buf bufs[kBufs];
   // pinned host staging of uploads and outputs, two buffers used in turn, and the events of
   // the copies that last used them
   Buf hbuf[2];
   hipEvent_t hev[2] = {nullptr, nullptr};
   std::vector<fa_frame*> cfr;
   fa_batch* cbatch = nullptr;

   void drop_cached()
   {
      if (cbatch) fa_batch_destroy(cbatch);
      cbatch = nullptr;
      for (fa_frame* f : cfr) fa_deregister_frame(f);
      cfr.clear();
   }

   // buffer `slot` with room for count values of T (its contents are not kept when it grows)
   template <typename T>
   int grow(int slot, size_t count, T** out)
   {
      const size_t need = std::max<size_t>(count, 1) * sizeof(T);
      Buf& b = bufs[slot];
      if (b.bytes < need) {
         FCHK(hipStreamSynchronize(stream));
         if (slot == kPre) drop_cached();  // the cached frames view the old array
         if (b.p) (void)hipFree(b.p);
         b.p = nullptr;
         b.bytes = 0;
         if (hipMalloc(&b.p, need) != hipSuccess) {
            (void)hipGetLastError();
            b.p = nullptr;
            return set_error(FA_E_OOM, "Not enough memory to allocate");
         }
         b.bytes = need;
      }
      *out = (T*)b.p;
      return FA_OK;
   }

   // both pinned staging buffers with room for `bytes` (the stream drained first when they
   // grow) and their events
   int grow_host(size_t bytes)
   {
      for (int i = 0; i < 2; i++) {
         if (!hev[i] && hipEventCreateWithFlags(&hev[i], hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            hev[i] = nullptr;
            return set_error(FA_E_DEVICE, "event creation failed");
         }
         Buf& b = hbuf[i];
         if (b.bytes >= bytes) continue;
         FCHK(hipStreamSynchronize(stream));
         if (b.p) (void)hipHostFree(b.p);
         b.p = nullptr;
         b.bytes = 0;
         if (hipHostMalloc(&b.p, std::max<size_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            b.p = nullptr;
            return set_error(FA_E_OOM, "Not enough memory to allocate");
         }
         b.bytes = bytes;
      }
      return FA_OK;
   }

   void free_host()
   {
      for (Buf& b : hbuf) {
         if (b.p) (void)hipHostFree(b.p);
         b.p = nullptr;
         b.bytes = 0;
      }
   }

   ~fa_aligner()
   {
      if (stream) (void)hipStreamSynchronize(stream);
      drop_cached();
      free_host();
      for (hipEvent_t e : hev)
         if (e) (void)hipEventDestroy(e);
      for (Buf& b : bufs)
         if (b.p) (void)hipFree(b.p);
      for (auto& kv : bplans) {
         if (kv.second.fwd) rocfft_plan_destroy(kv.second.fwd);
         if (kv.second.inv) rocfft_plan_destroy(kv.second.inv);
         if (kv.second.info) rocfft_execution_info_destroy(kv.second.info);
         if (kv.second.work) (void)hipFree(kv.second.work);
      }
      if (fwd) rocfft_plan_destroy(fwd);
      if (inv) rocfft_plan_destroy(inv);
      if (info) rocfft_execution_info_destroy(info);
      if (work) (void)hipFree(work);
      ffree(d_win);
      ffree(d_real);
      ffree(d_spec);
      ffree(d_refspec);
      ffree(d_pv);
      ffree(d_pi);
      ffree(d_idx);
      ffree(d_kx);
      ffree(d_ky);
      ffree(d_raw);
      ffree(d_tmp);
      if (d_up) (void)hipFree(d_up);
      if (ref) fa_reference_release(ref);
      if (stream) (void)hipStreamDestroy(stream);
   }

   size_t raw_bytes(int dtype) const { return (size_t)n_px * (dtype == FA_U16 ? 2 : 4); }

   // d_raw = float(raw)
   int upload(const void* raw, int dtype)
   {
      FCHK(hipMemcpyAsync(d_up, raw, raw_bytes(dtype), hipMemcpyHostToDevice, stream));
      hipLaunchKernelGGL(k_to_float, dim3(grid_of(n_px)), dim3(kFB), 0, stream, (const void*)d_up, dtype, n_px, d_raw);
      FCHK(hipGetLastError());
      return FA_OK;
   }

   // spectrum of the binned, windowed d_raw into d_spec
   int spectrum()
   {
      const int f = std::max(1, rp.spatial_binning);
      hipLaunchKernelGGL(k_bin_window, dim3(grid_of(n_el)), dim3(kFB), 0, stream, (const float*)d_raw, nz, ny, nx, f,
                         ph, pw, (const double*)d_win, d_real, 1);
      FCHK(hipGetLastError());
      void* in[1] = {d_real};
      void* out[1] = {d_spec};
      RCHK(rocfft_execute(fwd, in, out, info));
      return FA_OK;
   }

   // FP64 real-to-complex and complex-to-real plans for `count` frames at once (cached by count)
   struct Plans {
      rocfft_plan fwd = nullptr, inv = nullptr;
      rocfft_execution_info info = nullptr;
      void* work = nullptr;
   };
   std::map<int, Plans> bplans;
   int batch_plans(int count, Plans** out)
   {
      auto it = bplans.find(count);
      if (it == bplans.end()) {
         Plans& P = bplans[count];
         const size_t dims = pz > 1 ? 3 : 2;
         const size_t len[3] = {(size_t)pw, (size_t)ph, (size_t)pz};
         RCHK(rocfft_plan_create(&P.fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                                 rocfft_precision_double, dims, len, (size_t)count, nullptr));
         RCHK(rocfft_plan_create(&P.inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                                 rocfft_precision_double, dims, len, (size_t)count, nullptr));
         size_t w1 = 0, w2 = 0;
         RCHK(rocfft_plan_get_work_buffer_size(P.fwd, &w1));
         RCHK(rocfft_plan_get_work_buffer_size(P.inv, &w2));
         RCHK(rocfft_execution_info_create(&P.info));
         if (std::max(w1, w2) > 0) {
            if (hipMalloc(&P.work, std::max(w1, w2)) != hipSuccess) {
               (void)hipGetLastError();
               P.work = nullptr;
               return set_error(FA_E_OOM, "Not enough memory to allocate");
            }
            RCHK(rocfft_execution_info_set_work_buffer(P.info, P.work, std::max(w1, w2)));
         }
         RCHK(rocfft_execution_info_set_stream(P.info, stream));
         it = bplans.find(count);
      }
      *out = &it->second;
      return FA_OK;
   }

   // Rigid shifts of `count` frames (raw values as float, n_px apart) as flat indices of the
   // correlation maximum in idx[0..count) (device), with the work arrays the caller provides
   // (count times n_el doubles, n_cplx complex, n_part maxima).
   int shifts(const float* rawf, int count, double* w_real, double2* w_spec, double* w_pv, int64_t* w_pi,
              int64_t* idx)
   {
      Plans* P = nullptr;
      OK(batch_plans(count, &P));
      const int f = std::max(1, rp.spatial_binning);
      hipLaunchKernelGGL(k_bin_window, dim3(grid_of(n_el * count)), dim3(kFB), 0, stream, rawf, nz, ny, nx, f, ph,
                         pw, (const double*)d_win, w_real, count);
      FCHK(hipGetLastError());
      void* in[1] = {w_real};
      void* out[1] = {w_spec};
      RCHK(rocfft_execute(P->fwd, in, out, P->info));
      hipLaunchKernelGGL(k_cross_power, dim3(grid_of(n_cplx * count)), dim3(kFB), 0, stream, w_spec,
                         (const double2*)d_refspec, n_cplx, count);
      FCHK(hipGetLastError());
      void* in2[1] = {w_spec};
      void* out2[1] = {w_real};
      RCHK(rocfft_execute(P->inv, in2, out2, P->info));
      hipLaunchKernelGGL(k_argmax, dim3(n_part, count), dim3(kFB), 0, stream, (const double*)w_real, n_el, w_pv, w_pi);
      hipLaunchKernelGGL(k_argmax_final, dim3(count), dim3(kFB), 0, stream, (const double*)w_pv,
                         (const int64_t*)w_pi, n_part, idx);
      FCHK(hipGetLastError());
      return FA_OK;
   }

   // the shift of a flat maximum index: indices above half the size wrapped negative, x and y
   // scaled by the binning
   void shift_of(int64_t idx, double* s3) const
   {
      int x = (int)(idx % pw), y = (int)((idx / pw) % ph), z = (int)(idx / ((int64_t)pw * ph));
      if (x > pw / 2) x -= pw;
      if (y > ph / 2) y -= ph;
      if (z > pz / 2) z -= pz;
      const int f = std::max(1, rp.spatial_binning);
      s3[0] = (double)x * f;
      s3[1] = (double)y * f;
      s3[2] = (double)z;
   }

   // rigid shift of d_raw
   int shift(double* s3)
   {
      OK(shifts(d_raw, 1, d_real, d_spec, d_pv, d_pi, d_idx));
      int64_t idx = 0;
      FCHK(hipMemcpyAsync(&idx, d_idx, sizeof(int64_t), hipMemcpyDeviceToHost, stream));
      FCHK(hipStreamSynchronize(stream));
      shift_of(idx, s3);
      return FA_OK;
   }

   // out = smoothStack(in) + 1 for `count` frames (n_px apart); tmp: count frames (blur only)
   int smooth_plus_one(const float* in, float* out, int count, float* tmp)
   {
      const int64_t n = n_px * count;
      if (rp.smoothing > 0) {
         // the blur runs slice by slice: the frames are count * nz slices
         hipLaunchKernelGGL(k_blur, dim3(grid_of(n)), dim3(kFB), 0, stream, in, tmp, nz * count, ny, nx,
                            (const float*)d_kx, rx, 0, 0.f);
         hipLaunchKernelGGL(k_blur, dim3(grid_of(n)), dim3(kFB), 0, stream, (const float*)tmp, out, nz * count, ny, nx,
                            (const float*)d_ky, ry, 1, 1.f);
      } else {
         if (out != in) FCHK(hipMemcpyAsync(out, in, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
         hipLaunchKernelGGL(k_add, dim3(grid_of(n)), dim3(kFB), 0, stream, out, n, 1.f);
      }
      FCHK(hipGetLastError());
      return FA_OK;
   }
   int smooth_plus_one(float* out) { return smooth_plus_one(d_raw, out, 1, d_tmp); }

   // Frames handled together by the batched phases of add_frames: at most 32, and at most
   // about 1 GB of work arrays.
   int chunk_frames(int n) const
   {
      const int64_t per = n_px * 4 * 6 + n_el * 8 + n_cplx * 16 + n_part * 16;
      return (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)n, 32, ((int64_t)1 << 30) / per}));
   }

   // Output images and metrics of n frames at D_all (host, n x nD x 3): the raw frame's splat
   // (mask, preserving), its interpolated warp (realigned), the warp of the blurred frame + 1
   // (warped) and the correlations and coverage (FrameWarpAligner.cpp:218-240,317-346). rawf:
   // the raw frames as float, pre: the warper's frames (device, n_px apart). Frames run C at a
   // time; the metrics' sums come back in one copy at the end.
   int products(int n, const double* D_all, const float* rawf, const float* pre, const fa_add_outputs* o,
                fa_add_result* res)
   {
      const RefDev& R = ref->dev;
      const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(R.n_chunks, (int64_t)ctx->n_cu * 8));
      const int parts = (int)std::min<int64_t>(1024, (n_px + kFB - 1) / kFB);
      const int C = chunk_frames(n);
      const int64_t cpx = (int64_t)C * n_px;
      double *d_D = nullptr, *d_part = nullptr, *d_sums = nullptr;
      float *d_w = nullptr, *d_w2 = nullptr, *d_wimg = nullptr, *d_cov = nullptr;
      uint16_t* d_u16 = nullptr;
      auto done = [&](int rc) {
         (void)hipStreamSynchronize(stream);
         return rc;
      };
      int rc;
      if ((rc = grow(kD, 3 * (size_t)nD * n, &d_D)) || (rc = grow(kPart, (size_t)C * parts * kMoments, &d_part)) ||
          (rc = grow(kSums, (size_t)4 * n * kMoments, &d_sums)) || (rc = grow(kW, cpx, &d_w)) ||
          (rc = grow(kW2, cpx, &d_w2)) || (rc = grow(kWimg, cpx, &d_wimg)) || (rc = grow(kCov, cpx, &d_cov)) ||
          (rc = grow(kU16, cpx, &d_u16)))
         return done(rc);
      if (hipMemcpyAsync(d_D, D_all, 3 * (size_t)nD * n * sizeof(double), hipMemcpyHostToDevice, stream) != hipSuccess)
         return done(set_error(FA_E_DEVICE, "output warp failed"));
      auto fail = [&](const char* what) { (void)hipGetLastError(); return done(set_error(FA_E_DEVICE, what)); };
      // host outputs leave through the two pinned buffers in turn: a chunk's images are copied
      // there by the GPU, then into the caller's arrays by the host threads while the GPU works
      // on the next chunk. Per chunk: mask (float), preserving, realigned (uint16), warped (float).
      const size_t o_mask = 0, o_pres = o_mask + (o && o->mask ? (size_t)cpx * 4 : 0);
      const size_t o_real = o_pres + (o && o->preserving ? (size_t)cpx * 2 : 0);
      const size_t o_warp = o_real + (o && o->realigned ? (size_t)cpx * 2 : 0);
      const size_t o_end = o_warp + (o && o->warped ? (size_t)cpx * 4 : 0);
      if (o_end && (rc = grow_host(o_end))) return done(rc);
      struct Staged {
         int k = -1;
         int64_t off = 0, m = 0;
      } prev;
      auto deliver = [&](const Staged& q) -> int {
         if (q.k < 0 || !o_end) return FA_OK;
         if (hipEventSynchronize(hev[q.k & 1]) != hipSuccess) return set_error(FA_E_DEVICE, "output copy failed");
         const char* st = (const char*)hbuf[q.k & 1].p;
         std::vector<CopyPiece> pc;
         if (o->mask) pc.push_back({o->mask + q.off, st + o_mask, (size_t)q.m * 4});
         if (o->preserving) pc.push_back({o->preserving + q.off, st + o_pres, (size_t)q.m * 2});
         if (o->realigned) pc.push_back({o->realigned + q.off, st + o_real, (size_t)q.m * 2});
         if (o->warped) pc.push_back({o->warped + q.off, st + o_warp, (size_t)q.m * 4});
         par_copy(pc);
         return FA_OK;
      };
      for (int j0 = 0, k = 0; j0 < n; j0 += C, k++) {
         const int c = std::min(C, n - j0);
         const int64_t off = (int64_t)j0 * n_px, m = (int64_t)c * n_px;
         char* st = o_end ? (char*)hbuf[k & 1].p : nullptr;
         if (hipMemsetAsync(d_wimg, 0, m * sizeof(float), stream) != hipSuccess ||
             hipMemsetAsync(d_cov, 0, m * sizeof(float), stream) != hipSuccess)
            return fail("output warp failed");
         const FrameDev Fraw{rawf + off, FA_F32, 0.f}, Fsm{pre + off, FA_F32, 0.f};
         const double* Dc = d_D + 3 * (size_t)nD * j0;
         if (launch_splat(R, Fraw, Dc, d_wimg, d_cov, 0.f, grid, stream, c) != hipSuccess)
            return fail("output warp failed");
         if (o && o->mask &&
             hipMemcpyAsync(st + o_mask, d_cov, m * sizeof(float), hipMemcpyDeviceToHost, stream) != hipSuccess)
            return fail("output copy failed");
         if (o && o->preserving) {
            hipLaunchKernelGGL(k_preserving, dim3(grid_of(m)), dim3(kFB), 0, stream, (const float*)d_wimg,
                               (const float*)d_cov, m, d_u16);
            if (hipMemcpyAsync(st + o_pres, d_u16, m * sizeof(uint16_t), hipMemcpyDeviceToHost, stream) !=
                hipSuccess)
               return fail("output copy failed");
         }
         if (o && o->realigned) {
            if (launch_warp(R, Fraw, Dc, d_w, 0.f, 0, 0, grid, stream, c) != hipSuccess)
               return fail("output warp failed");
            hipLaunchKernelGGL(k_to_u16, dim3(grid_of(m)), dim3(kFB), 0, stream, (const float*)d_w, m, d_u16);
            if (hipMemcpyAsync(st + o_real, d_u16, m * sizeof(uint16_t), hipMemcpyDeviceToHost, stream) !=
                hipSuccess)
               return fail("output copy failed");
         }
         if (launch_warp(R, Fsm, Dc, d_w2, 0.f, 0, 0, grid, stream, c) != hipSuccess)
            return fail("output warp failed");
         if (o && o->warped &&
             hipMemcpyAsync(st + o_warp, d_w2, m * sizeof(float), hipMemcpyDeviceToHost, stream) != hipSuccess)
            return fail("output copy failed");
         if (res) {
            // correlation of the warped blurred frame and of the blurred frame itself with the
            // reference over the covered pixels: stage 1 (whose count is the coverage), then
            // stage 2 at its means
            const float* xs[2] = {d_w2, pre + off};
            for (int q = 0; q < 2; q++) {
               double* s1 = d_sums + ((int64_t)2 * q * n + j0) * kMoments;
               double* s2 = d_sums + ((int64_t)(2 * q + 1) * n + j0) * kMoments;
               hipLaunchKernelGGL(k_moments, dim3(parts, c), dim3(kFB), 0, stream, xs[q], R.rg4, (const float*)d_cov,
                                  n_px, 1, (const double*)nullptr, d_part);
               hipLaunchKernelGGL(k_sum_partials, dim3(c), dim3(64), 0, stream, (const double*)d_part, parts, s1);
               hipLaunchKernelGGL(k_moments, dim3(parts, c), dim3(kFB), 0, stream, xs[q], R.rg4, (const float*)d_cov,
                                  n_px, 2, (const double*)s1, d_part);
               hipLaunchKernelGGL(k_sum_partials, dim3(c), dim3(64), 0, stream, (const double*)d_part, parts, s2);
            }
         }
         if (hipGetLastError() != hipSuccess) return fail("output warp failed");
         if (o_end && hipEventRecord(hev[k & 1], stream) != hipSuccess) return fail("output copy failed");
         // the previous chunk's images into the caller's arrays while this chunk runs (its
         // buffer is the other one, free since that chunk's delivery two chunks ago)
         if ((rc = deliver(prev))) return done(rc);
         prev = Staged{k, off, m};
      }
      if ((rc = deliver(prev))) return done(rc);
      if (res) {
         std::vector<double> sums((size_t)4 * n * kMoments);
         if (hipMemcpyAsync(sums.data(), d_sums, sums.size() * sizeof(double), hipMemcpyDeviceToHost, stream) != hipSuccess ||
             hipStreamSynchronize(stream) != hipSuccess)
            return fail("metrics failed");
         for (int i = 0; i < n; i++) {
            const double* h = sums.data() + (size_t)i * kMoments;
            res[i].coverage = h[0] / (double)n_px;  // pixels with coverage != 0 (an exact count)
            res[i].correlation = correlation_of(h, h + (size_t)n * kMoments);
            res[i].unaligned_correlation = correlation_of(h + (size_t)2 * n * kMoments, h + (size_t)3 * n * kMoments);
         }
      }
      if (hipStreamSynchronize(stream) != hipSuccess) return fail("output copy failed");
      return done(FA_OK);
   }
};

namespace {

int aligner_init(fa_aligner* a, const void* reference, int dtype)
{
   const int f = std::max(1, a->rp.spatial_binning);
   a->pz = a->nz;
   a->ph = (a->ny + f - 1) / f;
   a->pw = (a->nx + f - 1) / f;
   a->n_el = (int64_t)a->pz * a->ph * a->pw;
   a->n_cplx = (int64_t)a->pz * a->ph * (a->pw / 2 + 1);
   a->n_part = (int)std::min<int64_t>(1024, (a->n_el + kFB - 1) / kFB);
   FCHK(hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking));
   OK(falloc(&a->d_win, a->n_el));
   OK(falloc(&a->d_real, a->n_el));
   OK(falloc(&a->d_spec, a->n_cplx));
   OK(falloc(&a->d_refspec, a->n_cplx));
   OK(falloc(&a->d_pv, a->n_part));
   OK(falloc(&a->d_pi, a->n_part));
   OK(falloc(&a->d_idx, 1));
   OK(falloc(&a->d_raw, a->n_px));
   OK(falloc(&a->d_tmp, a->n_px));
   if (hipMalloc(&a->d_up, a->n_px * 4) != hipSuccess) {
      (void)hipGetLastError();
      a->d_up = nullptr;
      return set_error(FA_E_OOM, "Not enough memory to allocate");
   }
   // window: 1 in z, square-root Hann in y and x
   {
      const std::vector<double> wy = hann(a->ph), wx = hann(a->pw);
      std::vector<double> w((size_t)a->n_el);
      size_t i = 0;
      for (int z = 0; z < a->pz; z++)
         for (int y = 0; y < a->ph; y++)
            for (int x = 0; x < a->pw; x++) w[i++] = wy[y] * wx[x];
      FCHK(hipMemcpy(a->d_win, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
   }
   if (a->rp.smoothing > 0) {
      const std::vector<float> kx = gaussian_weights(a->rp.smoothing), ky = gaussian_weights(1.0);
      a->rx = (int)kx.size() / 2;
      a->ry = (int)ky.size() / 2;
      OK(falloc(&a->d_kx, kx.size()));
      OK(falloc(&a->d_ky, ky.size()));
      FCHK(hipMemcpy(a->d_kx, kx.data(), kx.size() * sizeof(float), hipMemcpyHostToDevice));
      FCHK(hipMemcpy(a->d_ky, ky.data(), ky.size() * sizeof(float), hipMemcpyHostToDevice));
   }
   // FP64 real-to-complex and complex-to-real plans (lengths fastest first)
   std::call_once(g_fft_setup, [] { rocfft_setup(); });
   const size_t dims = a->pz > 1 ? 3 : 2;
   const size_t len[3] = {(size_t)a->pw, (size_t)a->ph, (size_t)a->pz};
   RCHK(rocfft_plan_create(&a->fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                           rocfft_precision_double, dims, len, 1, nullptr));
   RCHK(rocfft_plan_create(&a->inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                           rocfft_precision_double, dims, len, 1, nullptr));
   size_t w1 = 0, w2 = 0;
   RCHK(rocfft_plan_get_work_buffer_size(a->fwd, &w1));
   RCHK(rocfft_plan_get_work_buffer_size(a->inv, &w2));
   RCHK(rocfft_execution_info_create(&a->info));
   if (std::max(w1, w2) > 0) {
      if (hipMalloc(&a->work, std::max(w1, w2)) != hipSuccess) {
         (void)hipGetLastError();
         a->work = nullptr;
         return set_error(FA_E_OOM, "Not enough memory to allocate");
      }
      RCHK(rocfft_execution_info_set_work_buffer(a->info, a->work, std::max(w1, w2)));
   }
   RCHK(rocfft_execution_info_set_stream(a->info, a->stream));
   // reference: conjugate spectrum of the raw reference, and the warper's reference
   OK(a->upload(reference, dtype));
   OK(a->spectrum());
   hipLaunchKernelGGL(k_conj, dim3(grid_of(a->n_cplx)), dim3(kFB), 0, a->stream, (const double2*)a->d_spec,
                      a->d_refspec, a->n_cplx);
   FCHK(hipGetLastError());
   float* d_sm = nullptr;
   OK(falloc(&d_sm, a->n_px));
   std::vector<float> sm((size_t)a->n_px);
   int rc = a->smooth_plus_one(d_sm);
   if (rc == FA_OK && (hipMemcpyAsync(sm.data(), d_sm, a->n_px * sizeof(float), hipMemcpyDeviceToHost, a->stream) != hipSuccess ||
                       hipStreamSynchronize(a->stream) != hipSuccess)) {
      (void)hipGetLastError();
      rc = set_error(FA_E_DEVICE, "reference preprocessing failed");
   }
   ffree(d_sm);
   OK(rc);
   return fa_set_reference(a->ctx, sm.data(), a->nz, a->ny, a->nx, a->nD, &a->scan, 0, &a->ref);
}

}  // namespace

extern "C" {

int fa_aligner_create(fa_context* ctx, const fa_realign_params* rp, const fa_scan_params* scan,
                      const void* reference, int32_t dtype, fa_aligner** out)
{
   try {
      if (!ctx || !rp || !scan || !reference || !out) return set_error(FA_E_ARG, "null argument");
      *out = nullptr;
      if (dtype != FA_F32 && dtype != FA_U16) return set_error(FA_E_ARG, "unsupported frame dtype");
      if (rp->type < FA_REALIGN_NONE || rp->type > FA_REALIGN_WARP || rp->spatial_binning < 1 ||
          rp->n_resampling_points < 1 || !(rp->smoothing >= 0))
         return set_error(FA_E_ARG, "invalid realignment parameters");
      if (scan->n_x < 1 || scan->n_y < 1 || scan->n_z < 1) return set_error(FA_E_ARG, "Reference must be three dimensional");
      if (hipSetDevice(ctx->device) != hipSuccess) {
         (void)hipGetLastError();
         return set_error(FA_E_DEVICE, "No supported GPU");
      }
      auto* a = new fa_aligner();
      a->ctx = ctx;
      a->rp = *rp;
      a->scan = *scan;
      a->nz = scan->n_z;
      a->ny = scan->n_y;
      a->nx = scan->n_x;
      a->n_px = (int64_t)a->nz * a->ny * a->nx;
      a->n_dim = a->nz > 1 ? 3 : 2;
      a->nD = rp->n_resampling_points * a->nz + 1;
      const int rc = aligner_init(a, reference, dtype);
      if (rc != FA_OK) {
         delete a;
         return rc;
      }
      *out = a;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_aligner_destroy(fa_aligner* a)
{
   try {
      if (!a) return FA_OK;
      (void)hipSetDevice(a->ctx->device);
      delete a;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_aligner_clear_temp(fa_aligner* a)
{
   try {
      if (!a) return set_error(FA_E_ARG, "null argument");
      std::lock_guard<std::mutex> lk(a->mu);
      FCHK(hipSetDevice(a->ctx->device));
      FCHK(hipStreamSynchronize(a->stream));
      a->drop_cached();
      for (auto& b : a->bufs) {
         if (b.p) (void)hipFree(b.p);
         b.p = nullptr;
         b.bytes = 0;
      }
      a->free_host();
      return fa_clear_temp(a->ref);
   } catch (...) {
      return fa_caught();
   }
}

int fa_aligner_reference(fa_aligner* a, fa_reference** ref, int32_t* nD)
{
   try {
      if (!a) return set_error(FA_E_ARG, "null argument");
      if (ref) *ref = a->ref;
      if (nD) *nD = a->nD;
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_aligner_preprocess(fa_aligner* a, const void* raw, int32_t dtype, float* out)
{
   try {
      if (!a || !raw || !out) return set_error(FA_E_ARG, "null argument");
      if (dtype != FA_F32 && dtype != FA_U16) return set_error(FA_E_ARG, "unsupported frame dtype");
      std::lock_guard<std::mutex> lk(a->mu);
      FCHK(hipSetDevice(a->ctx->device));
      OK(a->upload(raw, dtype));
      OK(a->smooth_plus_one(a->d_raw));  // in place (the blur passes through d_tmp)
      FCHK(hipMemcpyAsync(out, a->d_raw, a->n_px * sizeof(float), hipMemcpyDeviceToHost, a->stream));
      FCHK(hipStreamSynchronize(a->stream));
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_aligner_rigid_shift(fa_aligner* a, const void* raw, int32_t dtype, double* shift3)
{
   try {
      if (!a || !raw || !shift3) return set_error(FA_E_ARG, "null argument");
      if (dtype != FA_F32 && dtype != FA_U16) return set_error(FA_E_ARG, "unsupported frame dtype");
      std::lock_guard<std::mutex> lk(a->mu);
      FCHK(hipSetDevice(a->ctx->device));
      OK(a->upload(raw, dtype));
      return a->shift(shift3);
   } catch (...) {
      return fa_caught();
   }
}

int fa_aligner_add_frames(fa_aligner* a, const void* const* frames, int32_t n, int32_t dtype,
                          const fa_tr_params* prm, double* D_out, const fa_add_outputs* outputs,
                          fa_add_result* results)
{
   try {
      if (!a || !frames || n <= 0 || !D_out) return set_error(FA_E_ARG, "null argument");
      if (dtype != FA_F32 && dtype != FA_U16) return set_error(FA_E_ARG, "unsupported frame dtype");
      for (int i = 0; i < n; i++)
         if (!frames[i]) return set_error(FA_E_ARG, "null frame");
      std::lock_guard<std::mutex> lk(a->mu);
      FCHK(hipSetDevice(a->ctx->device));
      const int nD = a->nD, m = a->n_dim, np = nD * m;
      const int64_t npx = a->n_px;
      const size_t rb = a->raw_bytes(dtype);
      const int C = a->chunk_frames(n);
      std::vector<double> shifts(3 * (size_t)n, 0.0);
      // the call's device arrays (kept in the aligner between calls): raw uploads, raw values as
      // float, the warper's frames (blur + 1, registered without a copy), and the rigid shift's
      // work arrays for C frames
      uint8_t* d_up = nullptr;
      float *d_rawf = nullptr, *d_pre = nullptr, *d_btmp = nullptr;
      double *w_real = nullptr, *w_pv = nullptr;
      double2* w_spec = nullptr;
      int64_t *w_pi = nullptr, *d_idx = nullptr;
      // on an error the cached frames and batch are dropped (rebuilt by the next call)
      auto release = [&]() {
         (void)hipStreamSynchronize(a->stream);
         a->drop_cached();
      };
      int rc = FA_OK;
      if ((rc = a->grow(fa_aligner::kUp, (size_t)n * rb, &d_up)) || (rc = a->grow(fa_aligner::kRawF, (size_t)n * npx, &d_rawf)) ||
          (rc = a->grow(fa_aligner::kPre, (size_t)n * npx, &d_pre)) ||
          (a->rp.smoothing > 0 && (rc = a->grow(fa_aligner::kBlurTmp, (size_t)C * npx, &d_btmp))) ||
          (rc = a->grow(fa_aligner::kReal, (size_t)C * a->n_el, &w_real)) ||
          (rc = a->grow(fa_aligner::kSpec, (size_t)C * a->n_cplx, &w_spec)) ||
          (rc = a->grow(fa_aligner::kPv, (size_t)C * a->n_part, &w_pv)) ||
          (rc = a->grow(fa_aligner::kPi, (size_t)C * a->n_part, &w_pi)) || (rc = a->grow(fa_aligner::kIdx, (size_t)n, &d_idx))) {
         release();
         return rc;
      }
      // 1. the frames uploaded U at a time through the two pinned buffers in turn (the host
      //    threads copy a chunk while the previous chunk's DMA and kernels run), each chunk
      //    converted, then C frames at a time: rigid shifts of the raw frames and the warper's
      //    frames (blur + 1); the shifts come back in one copy. U is a multiple of C, so the
      //    rigid batches are the same as over all frames at once.
      int U = (int)std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)(32 << 20) / (int64_t)rb));
      if (U >= C) U = U / C * C;
      if (rc == FA_OK) rc = a->grow_host((size_t)U * rb);
      for (int j0 = 0, k = 0; j0 < n && rc == FA_OK; j0 += U, k++) {
         const int c = std::min(U, n - j0);
         uint8_t* st = (uint8_t*)a->hbuf[k & 1].p;
         if (k >= 2 && hipEventSynchronize(a->hev[k & 1]) != hipSuccess) {
            rc = set_error(FA_E_DEVICE, "frame upload failed");
            break;
         }
         std::vector<CopyPiece> pc;
         for (int i = 0; i < c; i++) pc.push_back({st + (size_t)i * rb, frames[j0 + i], rb});
         par_copy(pc);
         if (hipMemcpyAsync(d_up + (size_t)j0 * rb, st, (size_t)c * rb, hipMemcpyHostToDevice, a->stream) != hipSuccess ||
             hipEventRecord(a->hev[k & 1], a->stream) != hipSuccess) {
            rc = set_error(FA_E_DEVICE, "frame upload failed");
            break;
         }
         hipLaunchKernelGGL(k_to_float, dim3(grid_of((int64_t)c * npx)), dim3(kFB), 0, a->stream,
                            (const void*)(d_up + (size_t)j0 * rb), dtype, (int64_t)c * npx, d_rawf + (int64_t)j0 * npx);
         if (hipGetLastError() != hipSuccess) {
            rc = set_error(FA_E_DEVICE, "frame upload failed");
            break;
         }
         for (int i0 = j0; i0 < j0 + c && rc == FA_OK; i0 += C) {
            const int cc = std::min(C, j0 + c - i0);
            if ((rc = a->shifts(d_rawf + (int64_t)i0 * npx, cc, w_real, w_spec, w_pv, w_pi, d_idx + i0)) == FA_OK)
               rc = a->smooth_plus_one(d_rawf + (int64_t)i0 * npx, d_pre + (int64_t)i0 * npx, cc, d_btmp);
         }
      }
      std::vector<int64_t> idx(n);
      if (rc == FA_OK && (hipMemcpyAsync(idx.data(), d_idx, n * sizeof(int64_t), hipMemcpyDeviceToHost, a->stream) != hipSuccess ||
                          hipStreamSynchronize(a->stream) != hipSuccess)) {
         (void)hipGetLastError();
         rc = set_error(FA_E_DEVICE, "frame preprocessing failed");
      }
      for (int i = 0; i < n && rc == FA_OK; i++) a->shift_of(idx[i], &shifts[3 * (size_t)i]);
      // the frames (views of d_pre) and their batch: those of the last call when the count is
      // the same, else new ones
      if (rc == FA_OK && (int)a->cfr.size() != n) {
         a->drop_cached();
         a->cfr.assign(n, nullptr);
         for (int i = 0; i < n && rc == FA_OK; i++)
            rc = adopt_device_frame(a->ref, d_pre + (int64_t)i * npx, FA_F32, 0.f, &a->cfr[i], false);
      }
      if (rc != FA_OK) {
         release();
         return rc;
      }
      fa_tr_params p;
      if (prm) p = *prm;
      else fa_tr_default(&p);
      if (!a->cbatch) {
         if ((rc = fa_batch_create(a->cfr.data(), n, &p, &a->cbatch)) != FA_OK) {
            release();
            return rc;
         }
      } else {
         batch_set_params(a->cbatch, p);
      }
      fa_batch* b = a->cbatch;
      // 2. SSD at the zero start (twice in the reference) and at the rigid start
      std::vector<double> D0(3 * (size_t)nD * n, 0.0), Dr(3 * (size_t)nD * n), e0(n), er(n);
      for (int i = 0; i < n; i++)
         for (int k = 0; k < nD; k++)
            for (int c = 0; c < 3; c++) Dr[((size_t)i * nD + k) * 3 + c] = shifts[3 * (size_t)i + c];
      if ((rc = fa_batch_eval(b, D0.data(), e0.data(), nullptr)) != FA_OK ||
          (rc = fa_batch_eval(b, Dr.data(), er.data(), nullptr)) != FA_OK) {
         release();
         return rc;
      }
      // 3. start: the first smallest of {e0, e0, er}; x in D2col order
      std::vector<double> x((size_t)np * n, 0.0);
      std::vector<int> start(n, 0);
      for (int i = 0; i < n; i++) {
         start[i] = er[i] < e0[i] ? 2 : 0;
         if (start[i] == 2)
            for (int k = 0; k < nD; k++)
               for (int c = 0; c < m; c++) x[(size_t)i * np + k * m + c] = shifts[3 * (size_t)i + c];
      }
      // 4. trust region (Warp); None keeps D = 0 (NullFrameAligner), Translation the start
      std::vector<fa_align_stats> st(n);
      std::memset(st.data(), 0, st.size() * sizeof(fa_align_stats));
      if (a->rp.type == FA_REALIGN_WARP) {
         if ((rc = fa_batch_run(b, x.data(), st.data(), nullptr)) != FA_OK) {
            release();
            return rc;
         }
      } else if (a->rp.type == FA_REALIGN_NONE) {
         std::fill(x.begin(), x.end(), 0.0);
      }
      for (int i = 0; i < n; i++) {
         double* D = D_out + (size_t)i * nD * 3;
         for (int k = 0; k < nD; k++)
            for (int c = 0; c < 3; c++) D[k * 3 + c] = c < m ? x[(size_t)i * np + k * m + c] : 0.0;
         if (results) {
            fa_add_result& r = results[i];
            std::memset(&r, 0, sizeof(r));
            r.start = start[i];
            for (int c = 0; c < 3; c++) r.rigid_shift[c] = shifts[3 * (size_t)i + c];
            r.start_error[0] = r.start_error[1] = e0[i];
            r.start_error[2] = er[i];
            r.stats = st[i];
         }
      }
      // 5. output images and metrics from the raw frame at the final D
      if (outputs || results) rc = a->products(n, D_out, d_rawf, d_pre, outputs, results);
      if (rc != FA_OK) release();
      return rc;
   } catch (...) {
      return fa_caught();
   }
}

int fa_aligner_realign_as_frame(fa_aligner* a, const double* D, const void* frame, int32_t dtype,
                                int32_t interpolate_missing, int32_t cv_type, void* out)
{
   try {
      if (!a || !D || !frame || !out) return set_error(FA_E_ARG, "null argument");
      if (dtype != FA_F32 && dtype != FA_U16) return set_error(FA_E_ARG, "unsupported frame dtype");
      if (!interpolate_missing && cv_type != 0 && cv_type != 2 && cv_type != 5)
         return set_error(FA_E_ARG, "cv_type must be 0 (CV_8U), 2 (CV_16U) or 5 (CV_32F)");
      std::lock_guard<std::mutex> lk(a->mu);
      FCHK(hipSetDevice(a->ctx->device));
      const RefDev& R = a->ref->dev;
      const int64_t n_px = a->n_px;
      const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(R.n_chunks, (int64_t)a->ctx->n_cu * 8));
      double* d_D = nullptr;
      float *d_w = nullptr, *d_cov = nullptr;
      void* d_out = nullptr;
      auto done = [&](int rc) {
         ffree(d_D);
         ffree(d_w);
         ffree(d_cov);
         if (d_out) (void)hipFree(d_out);
         return rc;
      };
      int rc;
      if ((rc = falloc(&d_D, 3 * (size_t)a->nD)) || (rc = falloc(&d_w, n_px)) || (rc = falloc(&d_cov, n_px)))
         return done(rc);
      OK(a->upload(frame, dtype));  // d_raw = the frame as float, as given
      const FrameDev Fraw{a->d_raw, FA_F32, 0.f};
      if (hipMemcpyAsync(d_D, D, 3 * (size_t)a->nD * sizeof(double), hipMemcpyHostToDevice, a->stream) != hipSuccess)
         return done(set_error(FA_E_DEVICE, "upload failed"));
      size_t out_bytes = n_px * sizeof(float);
      const void* src = d_w;
      if (interpolate_missing) {
         // warpImageInterpolated, invalid value 0
         if (launch_warp(R, Fraw, d_D, d_w, 0.f, 0, 0, grid, a->stream) != hipSuccess)
            return done(set_error(FA_E_DEVICE, "warp failed"));
      } else {
         // warpImage splat; the weighted values, not divided by the coverage, converted
         if (hipMemsetAsync(d_w, 0, n_px * sizeof(float), a->stream) != hipSuccess ||
             hipMemsetAsync(d_cov, 0, n_px * sizeof(float), a->stream) != hipSuccess ||
             launch_splat(R, Fraw, d_D, d_w, d_cov, 0.f, grid, a->stream) != hipSuccess)
            return done(set_error(FA_E_DEVICE, "warp failed"));
         if (cv_type != 5) {
            out_bytes = n_px * (cv_type == 0 ? 1 : 2);
            if (hipMalloc(&d_out, out_bytes) != hipSuccess) {
               (void)hipGetLastError();
               return done(set_error(FA_E_DEVICE, "device allocation failed"));
            }
            if (cv_type == 0)
               hipLaunchKernelGGL(k_to_u8, dim3(grid_of(n_px)), dim3(kFB), 0, a->stream, (const float*)d_w, n_px,
                                  (uint8_t*)d_out);
            else
               hipLaunchKernelGGL(k_to_u16, dim3(grid_of(n_px)), dim3(kFB), 0, a->stream, (const float*)d_w, n_px,
                                  (uint16_t*)d_out);
            src = d_out;
         }
      }
      if (hipGetLastError() != hipSuccess || hipMemcpyAsync(out, src, out_bytes, hipMemcpyDeviceToHost, a->stream) != hipSuccess ||
          hipStreamSynchronize(a->stream) != hipSuccess)
         return done(set_error(FA_E_DEVICE, "realignment failed"));
      return done(FA_OK);
   } catch (...) {
      return fa_caught();
   }
}

int fa_shift_pixel(const fa_reference* ref, const double* D, double* xyz)
{
   try {
      if (!ref || !D || !xyz) return set_error(FA_E_ARG, "null argument");
      double loc[3];
      OK(fa_warp_point(ref, D, (int32_t)xyz[0], (int32_t)xyz[1], (int32_t)xyz[2], 1, loc));
      for (int c = 0; c < 3; c++) xyz[c] -= loc[c];
      return FA_OK;
   } catch (...) {
      return fa_caught();
   }
}

int fa_write_realignment_info(const char* filename, const fa_scan_params* scan, int32_t n, int32_t nD,
                              const double* D, const double* unaligned_correlation, const double* correlation,
                              const double* coverage)
{
   try {
      if (!filename || !scan || n < 0 || nD < 1 || (n > 0 && (!D || !unaligned_correlation || !correlation || !coverage)))
         return set_error(FA_E_ARG, "invalid argument");
      if (n == 0) return FA_OK;  // nothing stored: the reference writes no file
      FILE* fh = std::fopen(filename, "w");
      if (!fh) return set_error(FA_E_ARG, std::string("cannot write ") + filename);
      // ImageScanParameters constructor (ImageScanParameters.h:7-20)
      const int ny = std::max(scan->n_y, 1);
      double interline = scan->interline_duration;
      if (interline < scan->line_duration) interline = 1.01 * scan->line_duration;
      const double frame = ny * interline;
      // std::ostream's default formatting of doubles is printf's %g
      std::fprintf(fh, "Realignment Information Version, 2\n");
      std::fprintf(fh, "Line Duration,%g\n", scan->line_duration);
      std::fprintf(fh, "Interline Duration,%g\n", interline);
      std::fprintf(fh, "Frame Duration,%g\n", frame);
      std::fprintf(fh, "Interframe Duration,%g\n", scan->interframe_duration);
      std::fprintf(fh, "Lines,%d\n", ny);
      std::fprintf(fh, "Frames Per Stack,%d\n", scan->n_z);
      std::fprintf(fh, "Frame, UnalignedCorrelation, Correlation, Coverage");
      for (int j = 0; j < nD; j++) std::fprintf(fh, ", dx_%d, dy_%d, dz_%d", j, j, j);
      std::fprintf(fh, "\n");
      for (int i = 0; i < n; i++) {
         std::fprintf(fh, "%d,%g,%g, %g", i, unaligned_correlation[i], correlation[i], coverage[i]);
         for (int j = 0; j < nD; j++) {
            const double* d = D + ((size_t)i * nD + j) * 3;
            std::fprintf(fh, ", %g, %g, %g", d[0], d[1], d[2]);
         }
         std::fprintf(fh, "\n");
      }
      const bool ok = std::fclose(fh) == 0;
      return ok ? FA_OK : set_error(FA_E_ARG, std::string("cannot write ") + filename);
   } catch (...) {
      return fa_caught();
   }
}

}  // extern "C"
