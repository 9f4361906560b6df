// gfx950 kernels of the warp-registration path.
//
// Precision: the node map (which interval a pixel belongs to, its fraction) and the sample
// position (which taps, whether a sample exists) are computed in FP64 with the reference's
// operation order, because they decide integer indices that must match the reference
// exactly. Interpolation, residuals and per-workgroup sums run in FP32 with compensated
// (Kahan) accumulation; partial sums cross workgroups as FP64, and JtJ / Jtr reductions and
// the trust-region solve are FP64.
#include "fa_kernels.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <type_traits>

namespace fa {

// ---- compensated FP32 accumulation -------------------------------------------------------
// The pair (s, c) holds the value s - c (Kahan's running compensation).
struct KF {
   float s, c;
};

__device__ __forceinline__ void kadd(KF& a, float x)
{
   float y = x - a.c;
   float t = a.s + y;
   a.c = (t - a.s) - y;
   a.s = t;
}

// Sum of two compensated values: TwoSum of the leading parts, errors folded into c.
__device__ __forceinline__ KF kcomb(KF a, KF b)
{
   float s = a.s + b.s;
   float bb = s - a.s;
   float err = (a.s - (s - bb)) + (b.s - bb);
   KF r;
   r.s = s;
   r.c = (a.c + b.c) - err;
   return r;
}

template <int NV>
__device__ __forceinline__ void wave_reduce(KF (&acc)[NV])
{
#pragma unroll
   for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
      for (int v = 0; v < NV; v++) {
         KF o;
         o.s = __shfl_xor(acc[v].s, off, 64);
         o.c = __shfl_xor(acc[v].c, off, 64);
         acc[v] = kcomb(acc[v], o);
      }
   }
}

// Block (kBlock threads) reduction; thread 0 writes NV doubles to out. Ends with a barrier.
template <int NV>
__device__ __forceinline__ void block_reduce_store(KF (&acc)[NV], KF (*sred)[NV], double* out)
{
   wave_reduce<NV>(acc);
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   if (lane == 0)
#pragma unroll
      for (int v = 0; v < NV; v++) sred[wave][v] = acc[v];
   __syncthreads();
   if (threadIdx.x == 0) {
#pragma unroll
      for (int v = 0; v < NV; v++) {
         KF t = sred[0][v];
         for (int w = 1; w < kBlock / 64; w++) t = kcomb(t, sred[w][v]);
         out[v] = (double)t.s - (double)t.c;
      }
   }
   __syncthreads();
}

// ---- node map (FrameWarper.cpp:239-293) -------------------------------------------------
// u = (z*interframe + y*interline + x*pixel) / stack * (nD-1), i = integer part, f = rest,
// each operation separately rounded as the reference's C++ states them.
__device__ __forceinline__ double node_coord(const ScanDev& sc, int z, int y, int x)
{
#pragma clang fp contract(off)
   double t = (double)z * sc.interframe + (double)y * sc.interline + (double)x * sc.pixel;
   return t / sc.stack * sc.nDm1;
}

__global__ void k_node_map(ScanDev sc, int nx, int ny, int64_t n, uint16_t* Di, float* Df)
{
   int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (p >= n) return;
   int x = (int)(p % nx);
   int64_t r = p / nx;
   int y = (int)(r % ny);
   int z = (int)(r / ny);
   double u = node_coord(sc, z, y, x);
   double ip = trunc(u);
   double f = u - ip;
   Di[p] = (uint16_t)(int)ip;
   if (Df) Df[p] = (float)f;
}

// First pixel of every interval; flags: 1 = interval index out of range, 2 = the map is not
// monotonic in scan order.
__global__ void k_interval_begin(const uint16_t* Di, int64_t n, int nD, int32_t* begin,
                                 int32_t* flags)
{
   int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (p >= n) return;
   int i = Di[p];
   if (i > nD - 2) atomicOr(flags, 1);
   if (p == 0 || Di[p - 1] != i) {
      if (p > 0 && Di[p - 1] > i) atomicOr(flags, 2);
      if (i < nD) atomicMin(&begin[i], (int32_t)p);
   }
}

// ---- Sobel (Sobel3d.cpp:8-84), bit-exact -------------------------------------------------
// One 3-tap pass at position pos of an axis of length n: interior c0*m(-1) + c1*m(0) +
// c2*m(+1); first sample e1[0]*m(0) + e1[1]*m(+1); last sample e2[0]*m(0) + e2[1]*m(-1); all
// in float, left to right. A single-plane Z pass is the identity (Sobel3d.cpp:40-44).
struct Taps {
   float c0, c1, c2, a0, a1, b0, b1;
};

template <class G>
__device__ __forceinline__ float filt(const Taps& t, int pos, int n, bool identity, G get)
{
#pragma clang fp contract(off)
   if (identity) return get(0);
   if (pos == 0) return t.a0 * get(0) + t.a1 * get(1);
   if (pos == n - 1) return t.b0 * get(0) + t.b1 * get(-1);
   return t.c0 * get(-1) + t.c1 * get(0) + t.c2 * get(1);
}

// Builds the per-pixel reference record: Sobel gradient (bit-exact to Sobel3d.cpp) and the
// node fraction Df (FrameWarper.cpp:258-267).
__global__ void k_reference_record(const float* __restrict__ T, int nz, int ny, int nx, int n_dim,
                                   ScanDev sc, float4* __restrict__ rg4, float* __restrict__ df)
{
   int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   int64_t n = (int64_t)nz * ny * nx;
   if (p >= n) return;
   const int x = (int)(p % nx);
   const int64_t r = p / nx;
   const int y = (int)(r % ny);
   const int z = (int)(r / ny);
   const Taps S{1.0f / 4.0f, 2.0f / 4.0f, 1.0f / 4.0f, 2.0f / 3.0f, 2.0f / 3.0f, 2.0f / 3.0f, 2.0f / 3.0f};
   const Taps Dv{-0.5f, 0.0f, 0.5f, -0.5f, 0.5f, 0.5f, -0.5f};
   const bool zid = nz == 1;
   auto Tat = [&](int zz, int yy, int xx) { return T[((int64_t)zz * ny + yy) * nx + xx]; };
   auto Sx = [&](int zz, int yy, int xx) {
      return filt(S, xx, nx, false, [&](int o) { return Tat(zz, yy, xx + o); });
   };
   auto Sy = [&](int zz, int yy, int xx) {
      return filt(S, yy, ny, false, [&](int o) { return Tat(zz, yy + o, xx); });
   };
   // gx = Dx(Sz(Sy(T)))
   float gx = filt(Dv, x, nx, false, [&](int ox) {
      return filt(S, z, nz, zid, [&](int oz) { return Sy(z + oz, y, x + ox); });
   });
   // gy = Dy(Sz(Sx(T)))
   float gy = filt(Dv, y, ny, false, [&](int oy) {
      return filt(S, z, nz, zid, [&](int oz) { return Sx(z + oz, y + oy, x); });
   });
   double u = node_coord(sc, z, y, x);
   float f = (float)(u - trunc(u));
   if (n_dim == 3) {
      // gz = Dz(Sy(Sx(T)))
      float gz = filt(Dv, z, nz, zid, [&](int oz) {
         return filt(S, y, ny, false, [&](int oy) { return Sx(z + oz, y + oy, x); });
      });
      rg4[p] = make_float4(T[p], gx, gy, gz);
      df[p] = f;
   } else {
      rg4[p] = make_float4(T[p], gx, gy, f);
   }
}

// ---- shared per-pixel helpers -------------------------------------------------------------

template <typename FT>
__device__ __forceinline__ float fload(const FrameDev& F, int64_t i)
{
   return (float)((const FT*)F.data)[i] + F.offset;
}

// p[i] through a global-address-space pointer with a 32-bit index: the compiler then issues a
// global load with a scalar base and a 32-bit offset instead of a flat load, whose completion
// is also counted against LDS operations.
// A pointer the whole wave holds the same value of, in scalar registers (the compiler cannot
// prove that for pointers read from a table indexed by the work item), so loads through it take
// a scalar base and a 32-bit vector offset.
template <typename T>
__device__ __forceinline__ const T* uniform_ptr(const T* p)
{
   const unsigned long long v = (unsigned long long)p;
   const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
   const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
   return (const T*)(((unsigned long long)hi << 32) | lo);
}

// *(p + byte offset off), off 32-bit: a global load with a scalar base and a 32-bit vector
// offset when p is uniform (uniform_ptr).
template <typename T>
__device__ __forceinline__ T gld_bytes(const T* p, unsigned off)
{
#if defined(__HIP_DEVICE_COMPILE__)
   return *(const __attribute__((address_space(1))) T*)((const __attribute__((address_space(1))) char*)p + off);
#else
   return *(const T*)((const char*)p + off);  // host pass: never executed
#endif
}

template <typename T>
__device__ __forceinline__ T gld(const T* p, unsigned i)
{
#if defined(__HIP_DEVICE_COMPILE__)
   return ((const __attribute__((address_space(1))) T*)p)[i];
#else
   return p[i];  // host pass: never executed
#endif
}

// Pixel coordinates of the pixels a thread visits in [p0, p1) of one plane: start at
// p0 + tid and advance by kBlock.
struct PixelWalk {
   int x, y;
   __device__ __forceinline__ void init(int q, int nx)
   {
      y = (int)((float)q * (1.0f / (float)nx));
      x = q - y * nx;
      while (x < 0) { x += nx; y--; }
      while (x >= nx) { x -= nx; y++; }
   }
   __device__ __forceinline__ void step(int nx)
   {
      x += kBlock;
      while (x >= nx) { x -= nx; y++; }
   }
};

// ---- JtJ partial sums (FrameWarper.cpp:296-395) -----------------------------------------
// Per pixel p of interval k (except its last pixel): a = (1-f) grad T (node k),
// b = f grad T (node k+1), each product formed in double and stored as float as the
// reference stores VI_dW_dp; float products a_i a_j, b_i b_j, b_i a_j accumulated.
template <int M>
__device__ void hess_chunk(const RefDev& R, const Chunk ch, double* out, KF (*sred)[21])
{
   constexpr int NA = M * (M + 1) / 2;
   constexpr int NV = 2 * NA + M * M;
   KF acc[NV];
#pragma unroll
   for (int v = 0; v < NV; v++) acc[v] = KF{0.f, 0.f};
   const int64_t plane_sz = (int64_t)R.nx * R.ny;
   for (int p = ch.p0 + threadIdx.x; p < ch.p1; p += kBlock) {
      if (p == ch.end_p) continue;
      if (ch.w != R.nx) {  // strip item: only its columns
         const int x = (int)((int64_t)p % plane_sz % R.nx);
         if (x < ch.x0 || x >= ch.x0 + ch.w) continue;
      }
      const float4 rec = R.rg4[p];
      const double fd = (double)(M == 3 ? R.df[p] : rec.w);
      const float g[3] = {rec.y, rec.z, rec.w};
      float a[M], b[M];
#pragma unroll
      for (int c = 0; c < M; c++) {
         a[c] = (float)((double)g[c] * (1.0 - fd));
         b[c] = (float)((double)g[c] * fd);
      }
      int v = 0;
#pragma unroll
      for (int i = 0; i < M; i++)
#pragma unroll
         for (int j = i; j < M; j++) kadd(acc[v++], a[i] * a[j]);
#pragma unroll
      for (int i = 0; i < M; i++)
#pragma unroll
         for (int j = i; j < M; j++) kadd(acc[v++], b[i] * b[j]);
#pragma unroll
      for (int i = 0; i < M; i++)
#pragma unroll
         for (int j = 0; j < M; j++) kadd(acc[v++], b[i] * a[j]);
   }
   block_reduce_store<NV>(acc, (KF(*)[NV])sred, out);
}

__global__ __launch_bounds__(kBlock) void k_hess_partials(RefDev R, double* partials)
{
   __shared__ KF sred[kBlock / 64][21];
   for (int c = blockIdx.x; c < R.n_chunks; c += gridDim.x) {
      Chunk ch = R.chunks[c];
      double* out = partials + (int64_t)c * kHessStride;
      if (R.n_dim == 2) hess_chunk<2>(R, ch, out, sred);
      else hess_chunk<3>(R, ch, out, sred);
   }
}

// ---- small FP64 linear algebra (block-tridiagonal Cholesky) -----------------------------
// Templated on the block size so the per-block work stays in registers; the operation order
// is the same as the oracle's (oracle/fa_oracle.c).

template <int M>
__device__ __forceinline__ bool chol_block(double (&A)[M][M], double (&il)[M])
{
#pragma unroll
   for (int j = 0; j < M; j++) {
      double s = A[j][j];
#pragma unroll
      for (int k = 0; k < j; k++) s -= A[j][k] * A[j][k];
      if (!(s > 0) || !isfinite(s)) return false;
      const double l = sqrt(s);
      il[j] = 1.0 / l;
      A[j][j] = l;
#pragma unroll
      for (int i = j + 1; i < M; i++) {
         double t = A[i][j];
#pragma unroll
         for (int k = 0; k < j; k++) t -= A[i][k] * A[j][k];
         A[i][j] = t * il[j];
      }
#pragma unroll
      for (int i = 0; i < j; i++) A[i][j] = 0;
   }
   return true;
}

// ---- block cyclic reduction ---------------------------------------------------------------
// Cholesky factorisation of T = H + lambda I (block-tridiagonal: nb diagonal blocks, blocks
// (k+1, k) below) in odd-even elimination order. At level s = 1, 2, 4, ... the blocks
// i = s, 3s, 5s, ... < nb are eliminated; no two of them are neighbours in the remaining chain,
// so a whole level runs in parallel, and the Schur complement on the blocks 0, 2s, 4s, ... is
// block-tridiagonal again. Block 0 is eliminated last. This is the Cholesky factor of a
// symmetric permutation of T: it exists exactly when T is positive definite, and
// ||L^-1 v||^2 = v^T T^-1 v as with the natural order. Depth ceil(log2 nb) + 1 levels.
//
// Storage, nb M x M blocks each (row-major):
//   A[i] = L_i, the Cholesky factor of block i's Schur complement when it is eliminated,
//          with the reciprocals of its diagonal in place of the diagonal;
//   W[i] = L_i^-1 T(i, i-s), the link to its left neighbour at its level s;
//   C[i] = L_i^-1 T(i+s, i)^T, the link to its right neighbour (when i + s < nb).
// While the factorisation runs, C[j] of a block j still in the chain holds T(j+s, j) at the
// current level. NT threads share the work and all of them call these functions; each level
// ends with a workgroup barrier.

template <int M>
__device__ __forceinline__ void ldb(const double* p, double (&X)[M][M])
{
#pragma unroll
   for (int i = 0; i < M; i++)
#pragma unroll
      for (int j = 0; j < M; j++) X[i][j] = p[i * M + j];
}

template <int M>
__device__ __forceinline__ void stb(double* p, const double (&X)[M][M])
{
#pragma unroll
   for (int i = 0; i < M; i++)
#pragma unroll
      for (int j = 0; j < M; j++) p[i * M + j] = X[i][j];
}

// X := L^-1 X for a lower-triangular L with reciprocal diagonal il.
template <int M>
__device__ __forceinline__ void trsm_lower(const double (&L)[M][M], const double (&il)[M], double (&X)[M][M])
{
#pragma unroll
   for (int c = 0; c < M; c++)
#pragma unroll
      for (int i = 0; i < M; i++) {
         double t = X[i][c];
#pragma unroll
         for (int q = 0; q < i; q++) t -= L[i][q] * X[q][c];
         X[i][c] = t * il[i];
      }
}

// v := L^-1 v and v := L^-T v for one diagonal factor block (reciprocal diagonal stored).
template <int M>
__device__ __forceinline__ void fwd_block(const double* Lp, double (&v)[M])
{
#pragma unroll
   for (int i = 0; i < M; i++) {
      double t = v[i];
#pragma unroll
      for (int q = 0; q < i; q++) t -= Lp[i * M + q] * v[q];
      v[i] = t * Lp[i * M + i];
   }
}

template <int M>
__device__ __forceinline__ void bwd_block(const double* Lp, double (&v)[M])
{
#pragma unroll
   for (int i = M - 1; i >= 0; i--) {
      double t = v[i];
#pragma unroll
      for (int q = i + 1; q < M; q++) t -= Lp[q * M + i] * v[q];
      v[i] = t * Lp[i * M + i];
   }
}

// Stores a factor block with its diagonal replaced by the reciprocals il.
template <int M>
__device__ __forceinline__ void stb_factor(double* p, const double (&L)[M][M], const double (&il)[M])
{
#pragma unroll
   for (int i = 0; i < M; i++)
#pragma unroll
      for (int j = 0; j < M; j++) p[i * M + j] = i == j ? il[i] : L[i][j];
}

// Factor of Hd/Hs + lambda I into (A, C, W); false (on every thread) when not positive definite.
// With y (nb blocks of M), also y := L^-1 y within the same levels: the steps of cr_fwd, in its
// order, after each level's elimination and Schur update (y is left undefined on failure).
template <int M, int NT>
__device__ bool cr_factor(int nb, const double* __restrict__ Hd, const double* __restrict__ Hs, double lambda,
                          double* A, double* C, double* W, double* y = nullptr)
{
   constexpr int MM = M * M;
   const int t = threadIdx.x;
   // H + lambda I into (A, C): each thread loads kCopy entries of both before it stores any,
   // so the loads (global memory or LDS) overlap instead of completing one by one
   constexpr int kCopy = 8;
   const int ne = nb * MM, nc = (nb - 1) * MM;
   for (int i0 = t; i0 < ne; i0 += kCopy * NT) {
      double hd[kCopy], hs[kCopy];
#pragma unroll
      for (int u = 0; u < kCopy; u++) {
         const int i = min(i0 + u * NT, ne - 1);
         hd[u] = Hd[i];
         hs[u] = Hs[min(i, nc > 0 ? nc - 1 : 0)];
      }
#pragma unroll
      for (int u = 0; u < kCopy; u++) {
         const int i = i0 + u * NT;
         if (i < ne) {
            A[i] = hd[u] + ((i % MM) % (M + 1) == 0 ? lambda : 0.0);
            C[i] = i < nc ? hs[u] : 0.0;
         }
      }
   }
   __syncthreads();
   for (int s = 1; s < nb; s <<= 1) {
      const int ne = (nb + s - 1) / (2 * s);  // i = s, 3s, ... < nb
      int fail = 0;
      for (int q = t; q < ne; q += NT) {
         const int i = s * (2 * q + 1);
         double L[M][M], X[M][M], il[M];
         ldb<M>(A + i * MM, L);
         if (!chol_block<M>(L, il)) {  // il: reciprocals of L's diagonal
            fail = 1;
            continue;
         }
         stb_factor<M>(A + i * MM, L, il);
         if (y) {
            double v[M];
#pragma unroll
            for (int a = 0; a < M; a++) v[a] = y[i * M + a];
            fwd_block<M>(A + i * MM, v);
#pragma unroll
            for (int a = 0; a < M; a++) y[i * M + a] = v[a];
         }
         ldb<M>(C + (i - s) * MM, X);
         trsm_lower<M>(L, il, X);
         stb<M>(W + i * MM, X);
         if (i + s < nb) {
#pragma unroll
            for (int a = 0; a < M; a++)
#pragma unroll
               for (int b = 0; b < M; b++) X[a][b] = C[i * MM + b * M + a];
            trsm_lower<M>(L, il, X);
            stb<M>(C + i * MM, X);
         }
      }
      if (__syncthreads_or(fail)) return false;
      const int nr = (nb + 2 * s - 1) / (2 * s);  // j = 0, 2s, 4s, ... < nb
      for (int q = t; q < nr; q += NT) {
         const int j = 2 * s * q;
         double Aj[M][M], X[M][M], V[M][M];
         ldb<M>(A + j * MM, Aj);
         if (j + s < nb) {
            ldb<M>(W + (j + s) * MM, X);
#pragma unroll
            for (int a = 0; a < M; a++)
#pragma unroll
               for (int b = 0; b < M; b++) {
                  double u = 0;
#pragma unroll
                  for (int r = 0; r < M; r++) u += X[r][a] * X[r][b];
                  Aj[a][b] -= u;
               }
            if (j + 2 * s < nb) {
               // new coupling T'(j+2s, j) = -T(j+2s, j+s) T(j+s, j+s)^-1 T(j+s, j) = -V^T W
               ldb<M>(C + (j + s) * MM, V);
               double N[M][M];
#pragma unroll
               for (int a = 0; a < M; a++)
#pragma unroll
                  for (int b = 0; b < M; b++) {
                     double u = 0;
#pragma unroll
                     for (int r = 0; r < M; r++) u += V[r][a] * X[r][b];
                     N[a][b] = -u;
                  }
               stb<M>(C + j * MM, N);
            }
         }
         if (j >= s) {
            ldb<M>(C + (j - s) * MM, V);
#pragma unroll
            for (int a = 0; a < M; a++)
#pragma unroll
               for (int b = 0; b < M; b++) {
                  double u = 0;
#pragma unroll
                  for (int r = 0; r < M; r++) u += V[r][a] * V[r][b];
                  Aj[a][b] -= u;
               }
         }
         stb<M>(A + j * MM, Aj);
         if (y) {
            double v[M];
#pragma unroll
            for (int a = 0; a < M; a++) v[a] = y[j * M + a];
            if (j + s < nb) {
               const double* X = W + (j + s) * MM;
#pragma unroll
               for (int a = 0; a < M; a++)
#pragma unroll
                  for (int r = 0; r < M; r++) v[a] -= X[r * M + a] * y[(j + s) * M + r];
            }
            if (j >= s) {
               const double* V = C + (j - s) * MM;
#pragma unroll
               for (int a = 0; a < M; a++)
#pragma unroll
                  for (int r = 0; r < M; r++) v[a] -= V[r * M + a] * y[(j - s) * M + r];
            }
#pragma unroll
            for (int a = 0; a < M; a++) y[j * M + a] = v[a];
         }
      }
      __syncthreads();
   }
   int fail = 0;
   if (t == 0) {
      double L[M][M], il[M] = {};
      ldb<M>(A, L);
      fail = chol_block<M>(L, il) ? 0 : 1;
      stb_factor<M>(A, L, il);
      if (y && !fail) {
         double v[M];
#pragma unroll
         for (int a = 0; a < M; a++) v[a] = y[a];
         fwd_block<M>(A, v);
#pragma unroll
         for (int a = 0; a < M; a++) y[a] = v[a];
      }
   }
   return __syncthreads_or(fail) == 0;
}

// y := L^-1 y with the factor of cr_factor (y: nb blocks of M).
template <int M, int NT>
__device__ void cr_fwd(int nb, const double* A, const double* C, const double* W, double* y)
{
   constexpr int MM = M * M;
   const int t = threadIdx.x;
   for (int s = 1; s < nb; s <<= 1) {
      const int ne = (nb + s - 1) / (2 * s);
      for (int q = t; q < ne; q += NT) {
         const int i = s * (2 * q + 1);
         double v[M];
#pragma unroll
         for (int a = 0; a < M; a++) v[a] = y[i * M + a];
         fwd_block<M>(A + i * MM, v);
#pragma unroll
         for (int a = 0; a < M; a++) y[i * M + a] = v[a];
      }
      __syncthreads();
      const int nr = (nb + 2 * s - 1) / (2 * s);
      for (int q = t; q < nr; q += NT) {
         const int j = 2 * s * q;
         double v[M];
#pragma unroll
         for (int a = 0; a < M; a++) v[a] = y[j * M + a];
         if (j + s < nb) {
            const double* X = W + (j + s) * MM;
#pragma unroll
            for (int a = 0; a < M; a++)
#pragma unroll
               for (int r = 0; r < M; r++) v[a] -= X[r * M + a] * y[(j + s) * M + r];
         }
         if (j >= s) {
            const double* V = C + (j - s) * MM;
#pragma unroll
            for (int a = 0; a < M; a++)
#pragma unroll
               for (int r = 0; r < M; r++) v[a] -= V[r * M + a] * y[(j - s) * M + r];
         }
#pragma unroll
         for (int a = 0; a < M; a++) y[j * M + a] = v[a];
      }
      __syncthreads();
   }
   if (t == 0) {
      double v[M];
#pragma unroll
      for (int a = 0; a < M; a++) v[a] = y[a];
      fwd_block<M>(A, v);
#pragma unroll
      for (int a = 0; a < M; a++) y[a] = v[a];
   }
   __syncthreads();
}

// y := L^-T y with the factor of cr_factor.
template <int M, int NT>
__device__ void cr_bwd(int nb, const double* A, const double* C, const double* W, double* y)
{
   constexpr int MM = M * M;
   const int t = threadIdx.x;
   if (t == 0) {
      double v[M];
#pragma unroll
      for (int a = 0; a < M; a++) v[a] = y[a];
      bwd_block<M>(A, v);
#pragma unroll
      for (int a = 0; a < M; a++) y[a] = v[a];
   }
   __syncthreads();
   int top = 1;
   while (2 * top < nb) top <<= 1;
   for (int s = nb > 1 ? top : 0; s >= 1; s >>= 1) {
      const int ne = (nb + s - 1) / (2 * s);
      for (int q = t; q < ne; q += NT) {
         const int i = s * (2 * q + 1);
         double v[M];
         const double* X = W + i * MM;
#pragma unroll
         for (int a = 0; a < M; a++) {
            double u = y[i * M + a];
#pragma unroll
            for (int r = 0; r < M; r++) u -= X[a * M + r] * y[(i - s) * M + r];
            v[a] = u;
         }
         if (i + s < nb) {
            const double* V = C + i * MM;
#pragma unroll
            for (int a = 0; a < M; a++)
#pragma unroll
               for (int r = 0; r < M; r++) v[a] -= V[a * M + r] * y[(i + s) * M + r];
         }
         bwd_block<M>(A + i * MM, v);
#pragma unroll
         for (int a = 0; a < M; a++) y[i * M + a] = v[a];
      }
      __syncthreads();
   }
}
__device__ __forceinline__ double d_norm(int n, const double* v)
{
   double s = 0;
   for (int i = 0; i < n; i++) s += v[i] * v[i];
   return sqrt(s);
}

// H entry (i, j) of the block-tridiagonal storage.
__device__ double h_get(const RefDev& R, int i, int j)
{
   const int m = R.n_dim;
   int bi = i / m, bj = j / m, a = i % m, b = j % m;
   if (bi == bj) return R.Hdiag[bi * m * m + a * m + b];
   if (bi == bj + 1) return R.Hsub[bj * m * m + a * m + b];
   if (bj == bi + 1) return R.Hsub[bi * m * m + b * m + a];
   return 0;
}

// ---- JtJ assembly, factorisation and spectrum bound (one workgroup) --------------------
constexpr int kHessThreads = 256;

__global__ __launch_bounds__(kHessThreads) void k_hess_assemble(RefDev R, const double* partials, double* Hdiag, double* Hsub,
                                double* L0, double* info)
{
   const int m = R.n_dim, mm = m * m;
   const int NA = m * (m + 1) / 2;
   for (int k = threadIdx.x; k < R.nD; k += blockDim.x) {
      double hd[9] = {0}, hs[9] = {0};
      // interval k: aa into the diagonal block of node k, ba into block (k+1, k)
      if (k <= R.nD - 2)
         for (int c = R.ichunk[k]; c < R.ichunk[k + 1]; c++) {
            const double* P = partials + (int64_t)c * kHessStride;
            int v = 0;
            for (int i = 0; i < m; i++)
               for (int j = i; j < m; j++) hd[i * m + j] += P[v++];
            v = 2 * NA;
            for (int i = 0; i < m; i++)
               for (int j = 0; j < m; j++) hs[i * m + j] += P[v++];
         }
      // interval k-1: bb into the diagonal block of node k
      if (k >= 1)
         for (int c = R.ichunk[k - 1]; c < R.ichunk[k]; c++) {
            const double* P = partials + (int64_t)c * kHessStride;
            int v = NA;
            for (int i = 0; i < m; i++)
               for (int j = i; j < m; j++) hd[i * m + j] += P[v++];
         }
      for (int i = 0; i < m; i++)
         for (int j = 0; j < i; j++) hd[i * m + j] = hd[j * m + i];
      for (int i = 0; i < mm; i++) {
         Hdiag[k * mm + i] = hd[i];
         if (k <= R.nD - 2) Hsub[k * mm + i] = hs[i];
      }
   }
   __syncthreads();
   R.Hdiag = Hdiag;
   R.Hsub = Hsub;
   // Gershgorin lower bound of the spectrum: min over rows of diagonal - sum |off-diagonal|
   __shared__ double s_min[kHessThreads / 64];
   const int n = R.nD * m;
   double lo = INFINITY;
   for (int i = threadIdx.x; i < n; i += kHessThreads) {
      double c = h_get(R, i, i), r = 0;
      int j0 = (i / m - 1) * m, j1 = (i / m + 2) * m;
      if (j0 < 0) j0 = 0;
      if (j1 > n) j1 = n;
      for (int j = j0; j < j1; j++)
         if (j != i) r += fabs(h_get(R, i, j));
      lo = fmin(lo, c - r);
   }
#pragma unroll
   for (int off = 32; off >= 1; off >>= 1) lo = fmin(lo, __shfl_xor(lo, off, 64));
   if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = lo;
   __syncthreads();
   const int64_t nbmm = (int64_t)R.nD * mm;
   const bool pd = m == 2 ? cr_factor<2, kHessThreads>(R.nD, Hdiag, Hsub, 0.0, L0, L0 + nbmm, L0 + 2 * nbmm)
                          : cr_factor<3, kHessThreads>(R.nD, Hdiag, Hsub, 0.0, L0, L0 + nbmm, L0 + 2 * nbmm);
   if (threadIdx.x == 0) {
      double e = s_min[0];
      for (int w = 1; w < kHessThreads / 64; w++) e = fmin(e, s_min[w]);
      info[0] = pd ? 1.0 : 0.0;
      info[1] = e;
   }
}

// ---- eigendecomposition of H (small problems) --------------------------------------------
// Parallel cyclic Jacobi (Golub & Van Loan, 8.5): every round applies m/2 disjoint rotations
// chosen by the circle method, so that each sweep of m - 1 rounds visits every pair (p, q)
// once; A <- J^T A J, V <- V J. A and V live in LDS (m <= kEigMax + 1). Sweeps stop when the
// off-diagonal Frobenius norm is below 1e-15 of ||H||_F (or after 40 sweeps, reported as not
// converged). info[0] = 1 if converged, info[1] = smallest eigenvalue.
constexpr int kEigThreads = 256;

__device__ __forceinline__ void round_pair(int m, int r, int k, int& p, int& q)
{
   if (k == 0) {
      p = r % (m - 1);
      q = m - 1;
   } else {
      p = (r + k) % (m - 1);
      q = (r - k + (m - 1)) % (m - 1);
   }
   if (p > q) {
      const int t = p;
      p = q;
      q = t;
   }
}

__device__ __forceinline__ double block_sum(double v, double* red)
{
#pragma unroll
   for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
   __syncthreads();
   if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
   __syncthreads();
   double t = 0;
   for (int w = 0; w < kEigThreads / 64; w++) t += red[w];
   return t;
}

__global__ __launch_bounds__(kEigThreads) void k_eigen(RefDev R, double* Q, double* ev, double* info)
{
   extern __shared__ double sm[];
   __shared__ double red[kEigThreads / 64];
   const int n0 = R.nD * R.n_dim, m = n0 + (n0 & 1), h = m / 2;
   double* A = sm;
   double* V = sm + m * m;
   double* cs = V + m * m;
   const int t = threadIdx.x;
   for (int i = t; i < m * m; i += kEigThreads) {
      const int r = i / m, c = i - r * m;
      A[i] = (r < n0 && c < n0) ? h_get(R, r, c) : 0.0;
      V[i] = r == c ? 1.0 : 0.0;
   }
   __syncthreads();
   double fro = 0;
   for (int i = t; i < m * m; i += kEigThreads) fro += A[i] * A[i];
   fro = sqrt(block_sum(fro, red));
   int sweep = 0;
   for (; sweep < 40; sweep++) {
      double off = 0;
      for (int i = t; i < m * m; i += kEigThreads) {
         const int r = i / m, c = i - r * m;
         if (r != c) off += A[i] * A[i];
      }
      if (!(sqrt(block_sum(off, red)) > 1e-15 * fro)) break;
      for (int rnd = 0; rnd < m - 1; rnd++) {
         for (int k = t; k < h; k += kEigThreads) {
            int p, q;
            round_pair(m, rnd, k, p, q);
            double c = 1.0, s = 0.0;
            const double apq = A[p * m + q];
            if (q < n0 && apq != 0.0) {
               const double th = (A[q * m + q] - A[p * m + p]) / (2.0 * apq);
               const double tt = th == 0.0 ? 1.0 : copysign(1.0, th) / (fabs(th) + sqrt(th * th + 1.0));
               c = 1.0 / sqrt(tt * tt + 1.0);
               s = tt * c;
            }
            cs[2 * k] = c;
            cs[2 * k + 1] = s;
         }
         __syncthreads();
         for (int e = t; e < h * m; e += kEigThreads) {
            const int k = e / m, j = e - k * m;
            int p, q;
            round_pair(m, rnd, k, p, q);
            const double c = cs[2 * k], s = cs[2 * k + 1];
            const double ap = A[p * m + j], aq = A[q * m + j];
            A[p * m + j] = c * ap - s * aq;
            A[q * m + j] = s * ap + c * aq;
         }
         __syncthreads();
         for (int e = t; e < h * m; e += kEigThreads) {
            const int k = e / m, i = e - k * m;
            int p, q;
            round_pair(m, rnd, k, p, q);
            const double c = cs[2 * k], s = cs[2 * k + 1];
            const double ap = A[i * m + p], aq = A[i * m + q];
            A[i * m + p] = c * ap - s * aq;
            A[i * m + q] = s * ap + c * aq;
            const double vp = V[i * m + p], vq = V[i * m + q];
            V[i * m + p] = c * vp - s * vq;
            V[i * m + q] = s * vp + c * vq;
         }
         __syncthreads();
      }
   }
   // Q, then its transpose (the subproblem reads Q column-wise for Q^T g and row-wise for
   // Q w; with the transpose both reads are coalesced)
   for (int i = t; i < n0 * n0; i += kEigThreads) {
      const int r = i / n0, c = i - r * n0;
      Q[i] = V[r * m + c];
      Q[n0 * n0 + c * n0 + r] = V[r * m + c];
   }
   double lo = INFINITY;
   for (int i = t; i < n0; i += kEigThreads) {
      ev[i] = A[i * m + i];
      lo = fmin(lo, A[i * m + i]);
   }
#pragma unroll
   for (int off = 32; off >= 1; off >>= 1) lo = fmin(lo, __shfl_xor(lo, off, 64));
   __syncthreads();
   if ((t & 63) == 0) red[t >> 6] = lo;
   __syncthreads();
   if (t == 0) {
      double e = red[0];
      for (int w = 1; w < kEigThreads / 64; w++) e = fmin(e, red[w]);
      info[0] = sweep < 40 ? 1.0 : 0.0;
      info[1] = e;
   }
}

// ---- the fused pass ------------------------------------------------------------------------
// For every pixel of a work item: node fraction, displacement d = f D[k+1] + (1-f) D[k]
// (warpPoint, FrameWarper.cpp:397-419), zero-padded bilinear/trilinear pull-sample of the
// frame at p + d (warpImageInterpolated, FrameWarper.cpp:31-69), residual against the
// reference and SSD over samples that exist (computeErrorImage, FrameWarper.cpp:4-28), and
// Jtr rows (computeJacobian, FrameWarper.cpp:147-167). The frame rows the work item can
// reach are staged in LDS when they fit.
enum { MODE_PASS = 0, MODE_WARP = 1 };

// v plus the value of the lane that DPP control CTRL selects, in FP64 (the two 32-bit halves
// are moved separately).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_add(double v)
{
   const unsigned long long b = (unsigned long long)__double_as_longlong(v);
   const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROW_MASK, 0xf, false);
   const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROW_MASK, 0xf, false);
   return v + __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Frame samples a work item can reach, staged in LDS: planes [za, za+nzs) and rows
// [ya, ya+nrows) of those planes, each row as kPad zeros, the nx samples, kPad zeros
// (stride floats). Rows and planes outside the frame are staged as zeros, which gives the
// zero padding of getZeroPadded (FrameWarper.h:39-47) without a test per tap.
struct Stage {
   int za, nzs, ya, nrows, stride;
};

// Each thread issues NB loads of staged rows before it stores any, so a wave has that many row
// segments in flight instead of waiting out one global-load latency per row. NB is 6 for 2-D
// items (the pass then fits 7 waves per SIMD) and 8 for 3-D items (two planes: twice the rows).
// Frames whose width is not a multiple of 4 are staged sample by sample over one flat index of
// (plane, row, column), kBlock apart per thread, with (zi, yr, c) advanced without a division.

struct StageWalk {
   int zi, yr, c;
   __device__ __forceinline__ void init(int i, int ncol, int nrows)
   {
      const int r = i / ncol;
      c = i - r * ncol;
      zi = r / nrows;
      yr = r - zi * nrows;
   }
   __device__ __forceinline__ void step(int ncol, int nrows, int dr, int dc)
   {
      c += dc;
      yr += dr;
      if (c >= ncol) {
         c -= ncol;
         yr++;
      }
      while (yr >= nrows) {
         yr -= nrows;
         zi++;
      }
   }
};

// Four consecutive samples as stored (uint2 for uint16, float4 for float) and their float
// values plus the frame's offset.
template <typename FT>
struct Raw4 {
   using type = float4;
};
template <>
struct Raw4<uint16_t> {
   using type = uint2;
};

__device__ __forceinline__ float4 cvt4(uint2 w, float off)
{
   return make_float4((float)(w.x & 0xffffu) + off, (float)(w.x >> 16) + off, (float)(w.y & 0xffffu) + off,
                      (float)(w.y >> 16) + off);
}

__device__ __forceinline__ float4 cvt4(float4 w, float off)
{
   return make_float4(w.x + off, w.y + off, w.z + off, w.w + off);
}

// The tile's columns: frame columns [cx - kPad, cx + cw + kPad) at tile columns 0, 1, ...;
// cw = nx (cx = 0) for whole-width items, whose padding columns lie outside the frame. Staged
// row R (0 .. nzs * nrows - 1) is frame plane za + R / nrows, row ya + R % nrows, at tile
// floats R * stride.
template <int M, int NB, typename FT>
__device__ void stage_frame(const RefDev& R, const FrameDev& F, const Stage& S, float* lds, int cx, int cw)
{
   const int nx = R.nx, ny = R.ny, nz = R.nz;
   const int rows = S.nzs * S.nrows;
   const int tid = threadIdx.x;
   const FT* __restrict__ src = uniform_ptr((const FT*)F.data);
   const float off = F.offset;
   const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
   if ((nx & 3) == 0 && cw == nx) {
      constexpr int P4 = kPad / 4;
      for (int i = tid; i < rows * 2 * P4; i += kBlock) {
         const int r = i / (2 * P4), j = i - r * (2 * P4);
         const int col = j < P4 ? 4 * j : kPad + nx + 4 * (j - P4);
         *(float4*)&lds[r * S.stride + col] = zero4;
      }
   }
   if ((nx & 3) == 0) {
      // column groups of four samples: the frame's own (whole width; padding zeroed above)
      // or the strip's with its padding, which holds frame samples where the frame has them
      const int ncol4 = cw == nx ? nx >> 2 : (cw + 2 * kPad) >> 2;
      const int x_org = cw == nx ? 0 : cx - kPad;  // frame column of group 0
      const int t_org = cw == nx ? kPad : 0;       // tile column of group 0
      using RT = typename Raw4<FT>::type;
      // A thread keeps one column group cg and takes staged rows tr, tr + rpi, ... (rpi = rows
      // per sweep of the workgroup), NB of them loaded before any is stored, so a wave has NB
      // row segments in flight; rows wider than the workgroup take several column groups per
      // thread. In 3-D the plane of staged row R is R / nrows, formed as a float product with
      // a half-row margin (exact for the few thousand rows a tile has). Rows and planes
      // outside the frame load from a clamped row (in bounds) and store zeros, so the loads
      // are not under branches. Byte offsets are 32-bit (frames hold fewer than 2^28 samples).
      const int rpi = ncol4 <= kBlock ? kBlock / ncol4 : 1;
      const int tr = ncol4 <= kBlock ? tid / ncol4 : 0;
      const float inv_nrows = 1.0f / (float)S.nrows;
      if (tr < rpi) {
         for (int cg = ncol4 <= kBlock ? tid - tr * ncol4 : tid; cg < ncol4; cg += kBlock) {
            const int xg = x_org + 4 * cg;
            const bool col_ok = (unsigned)xg < (unsigned)nx;
            const unsigned xs = col_ok ? (unsigned)xg : 0u;
            const int tcol = t_org + 4 * cg;
            for (int R0 = tr; R0 < rows; R0 += NB * rpi) {
               RT raw[NB];
               bool ok[NB];
#pragma unroll
               for (int u = 0; u < NB; u++) {
                  const int Rr = R0 + u * rpi;
                  int zi = 0, yr = Rr;
                  if (M == 3) {
                     zi = (int)(((float)Rr + 0.5f) * inv_nrows);
                     yr = Rr - __mul24(zi, S.nrows);
                  }
                  const int yy = S.ya + yr, zz = S.za + zi;
                  ok[u] = Rr < rows && col_ok && (unsigned)yy < (unsigned)ny && (M == 2 || (unsigned)zz < (unsigned)nz);
                  // rows and planes outside the frame read some row of the frame (discarded)
                  const unsigned yc = min((unsigned)yy, (unsigned)(ny - 1));
                  const unsigned zc = M == 3 ? min((unsigned)zz, (unsigned)(nz - 1)) : 0u;
                  raw[u] = gld_bytes((const RT*)src, (unsigned)sizeof(FT) * (__umul24(zc * (unsigned)ny + yc, (unsigned)nx) + xs));
               }
#pragma unroll
               for (int u = 0; u < NB; u++) {
                  const int Rr = R0 + u * rpi;
                  // samples outside the frame are zero: masked with a bitwise and, not a branch
                  const float4 v = cvt4(raw[u], off);
                  const unsigned mk = ok[u] ? 0xffffffffu : 0u;
                  const float4 w = make_float4(__uint_as_float(__float_as_uint(v.x) & mk), __uint_as_float(__float_as_uint(v.y) & mk),
                                               __uint_as_float(__float_as_uint(v.z) & mk), __uint_as_float(__float_as_uint(v.w) & mk));
                  if (Rr < rows) *(float4*)&lds[__mul24(Rr, S.stride) + tcol] = w;
               }
            }
         }
      }
   } else {
      const int total = rows * S.stride;
      const int dr = kBlock / S.stride, dc = kBlock - dr * S.stride;
      StageWalk w;
      w.init(tid, S.stride, S.nrows);
      for (int i0 = tid; i0 < total; i0 += NB * kBlock) {
         FT raw[NB];
         int dst[NB];
         bool ok[NB];
#pragma unroll
         for (int u = 0; u < NB; u++) {
            const int yy = S.ya + w.yr, zz = S.za + w.zi, xx = w.c - kPad;
            const bool live = i0 + u * kBlock < total;
            dst[u] = live ? (w.zi * S.nrows + w.yr) * S.stride + w.c : -1;
            ok[u] = live && (unsigned)yy < (unsigned)ny && (unsigned)zz < (unsigned)nz && (unsigned)xx < (unsigned)nx;
            const int yc = min(max(yy, 0), ny - 1), zc = min(max(zz, 0), nz - 1), xc = min(max(xx, 0), nx - 1);
            raw[u] = src[((int64_t)zc * ny + yc) * nx + xc];
            w.step(S.stride, S.nrows, dr, dc);
         }
#pragma unroll
         for (int u = 0; u < NB; u++)
            if (dst[u] >= 0) lds[dst[u]] = ok[u] ? (float)raw[u] + off : 0.f;
      }
   }
   __syncthreads();
}

// The work item's sums (SSD, then per component sum(g e) and sum((g e) f)) over the workgroup,
// through LDS, written to out as SSD, Jtr of node k (per component), Jtr of node k + 1. Every
// thread stores its values (red: NV x kBlock doubles), then lanes 8v .. 8v + 7 of the first
// wave each add a fixed eighth of value v's column and the eight partial sums are combined
// with DPP moves within each group of 8 lanes (lane pairs, then quad pairs, then the two quads
// by row_half_mirror), so every value is summed in a fixed order and a frame's sums do not
// depend on which work items or frames share the launch. Only the first wave does
// floating-point work; the other waves store and go on. red may alias the staging tile: the
// first barrier waits for every wave to finish sampling, and the caller's next staging
// follows the second.
template <int M>
__device__ __forceinline__ void block_reduce_lds(const double (&acc)[1 + 2 * M], double* red, double* out)
{
   constexpr int NV = 1 + 2 * M;
   __syncthreads();
#pragma unroll
   for (int v = 0; v < NV; v++) red[v * kBlock + threadIdx.x] = acc[v];
   __syncthreads();
   if (threadIdx.x < 64) {
      const int v = threadIdx.x >> 3, s = threadIdx.x & 7;
      double t = 0;
      if (v < NV)
#pragma unroll 8
         for (int i = 0; i < kBlock / 8; i++) t += red[v * kBlock + s + 8 * i];
      t = dpp_add<0xb1, 0xf>(t);   // quad_perm [1,0,3,2]
      t = dpp_add<0x4e, 0xf>(t);   // quad_perm [2,3,0,1]
      t = dpp_add<0x141, 0xf>(t);  // row_half_mirror: the other quad of the 8
      // node k's Jtr rows are sum(g e) - sum((g e) f), the latter held 8 M lanes up
      const double tm = __shfl(t, (int)((threadIdx.x + 8 * M) & 63), 64);
      if (s == 0 && v < NV) out[v] = (v >= 1 && v <= M) ? t - tm : t;
   }
   __syncthreads();
}

// For every pixel of a work item: displacement d = f D[k+1] + (1-f) D[k] (warpPoint,
// FrameWarper.cpp:397-419), zero-padded bilinear/trilinear pull-sample of the frame at p + d
// (warpImageInterpolated, FrameWarper.cpp:31-69), residual against the reference and SSD over
// samples that exist (computeErrorImage, FrameWarper.cpp:4-28), and the Jacobian rows
// (computeJacobian, FrameWarper.cpp:147-167).
//
// Precision. The displacement is FP64, as the reference forms it: d = D[k] + f (D[k+1] - D[k])
// relative to the pixel (one FP64 fma per component from the float node fraction f), split
// into an integer tap offset and a fraction; D[k+1] = D[k] gives D[k] exactly (integer shifts
// sample exactly). The interpolation, the residual e and the products with the gradient are
// FP32. Each thread sums its pixels' terms in FP64 (e*e, g e and (g e) f with g e a float
// product, formed exactly in FP64 by the fma; the Jtr row of node k is sum(g e) - sum((g e) f)),
// the workgroup in FP64 in a fixed order. On gfx950 an FP64 fma issues at the rate of an
// unpacked FP32 one, so these sums cost about what compensated FP32 sums of 4-pixel FP32
// partials did, and they are as accurate as the reference's double sums. (Displacements
// rounded to float move the sample point by ~1e-7 px per node, and FP32 partial sums lose ~1e-8
// of Jtr; both are small against the 1e-3 tolerance, but the trust region amplifies them: on
// 256^2 x 32 volumes with 513 nodes it took a different path from the reference's arithmetic.)
//
// Thread layout. A work item is the pixels of rows r0..r1 of one plane and columns
// [x0, x0 + w) between p0 and p1 - 1 in scan order. Each thread keeps one column (w >= kBlock:
// columns kBlock apart; narrower items run kBlock / w rows side by side) and walks down its
// rows, so a pixel's record, frame row and tile row advance by constants; the column's first and
// last rows follow from p0 and p1 (an item may start and end mid-row). kUnroll rows per step:
// their records are loaded before any is used; rows past the column's last are computed on
// its last row and masked.
template <int M, typename FT, int MODE>
__device__ void pass_chunk(const RefDev& R, const FrameDev& F, const Chunk ch,
                           const double* __restrict__ D, float* lds, int lds_cap, bool want_lds,
                           double* out, float* wout, float invalid)
{
   constexpr int NV = 1 + 2 * M;
   const int nx = R.nx, ny = R.ny, nz = R.nz;
   const int64_t plane_sz = (int64_t)nx * ny;
   const int k = ch.k;
   double Da[3], Dd[3];
   bool finite = true;
#pragma unroll
   for (int c = 0; c < 3; c++) {
      const double d0 = (c < M) ? D[3 * k + c] : 0.0, d1 = (c < M) ? D[3 * (k + 1) + c] : 0.0;
      finite = finite && fabs(d0) < 1e6 && fabs(d1) < 1e6;
      Da[c] = d0;
      Dd[c] = d1 - d0;
   }
   const int plane = (int)(ch.p0 / plane_sz);
   const int pb = plane * (int)plane_sz;
   const int r0 = (ch.p0 - pb) / nx, r1 = (ch.p1 - 1 - pb) / nx;

   // f rises with the pixel's scan time, so over the work item it lies in [f(p0), f(p1-1)]
   // and d between its values there; the margin covers the rounding of the per-pixel fma
   // (half an FP64 ulp of |d|; none when D[k+1] = D[k]).
   Stage S{0, 1, 0, 0, 0};
   bool use = false;
   // Positions d + K with Dsh = D[k] + K and integers K = -floor(lo) that make every pixel's
   // displacement d + K >= 0 over the work item: in a staged tile its integer part is a
   // truncation and its fraction one v_fract (d + K rounds at ~1e-15 px). The global-load
   // sampler takes the same d + K (floor, then -K), so an item gives the same bits staged or
   // not, whichever LDS budget its launch has.
   double Dsh[3] = {Da[0], Da[1], Da[2]};
   int Ksh[3] = {0, 0, 0};
   if (finite) {
      const double lim = kPad - 3;
      const double fa = (double)((M == 3) ? R.df[ch.p0] : R.rg4[ch.p0].w);
      const double fb = (double)((M == 3) ? R.df[ch.p1 - 1] : R.rg4[ch.p1 - 1].w);
      double lo[3], hi[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
         const double a = Da[c], dd = Dd[c];
         const double va = a + fa * dd, vb = a + fb * dd;
         const double e = dd == 0.0 ? 0.0 : 0x1p-40 * (fabs(a) + fabs(dd));
         lo[c] = fmin(va, vb) - e;
         hi[c] = fmax(va, vb) + e;
      }
#pragma unroll
      for (int c = 0; c < M; c++) {
         if (fabs(lo[c]) < 0x1p20) Ksh[c] = -(int)floor(lo[c]);
         Dsh[c] = Da[c] + (double)Ksh[c];
      }
      if (want_lds && lo[0] > -lim && hi[0] < lim) {
         // taps y0, y0 + 1 with y0 = y + floor(d_y); likewise in z
         S.ya = r0 + (int)floor(lo[1]);
         S.nrows = r1 + (int)floor(hi[1]) + 1 - S.ya + 1;
         if (M == 3) {
            S.za = plane + (int)floor(lo[2]);
            S.nzs = plane + (int)floor(hi[2]) + 1 - S.za + 1;
         }
         S.stride = ((ch.w + 2 * kPad) + 3) & ~3;
         use = (int64_t)S.nzs * S.nrows * S.stride <= lds_cap;
      }
   }
   const int cx = ch.x0, cw = ch.w;  // columns of the item (cx = 0, cw = nx: whole rows)
   if (use) stage_frame<M, 6, FT>(R, F, S, lds, cx, cw);
   const int plane_step = S.nrows * S.stride;  // tile floats between staged planes
   const int stride4 = 4 * S.stride;            // tile row in bytes
   const unsigned nx16 = 16u * (unsigned)nx, nx4 = 4u * (unsigned)nx;  // a frame row of records, of node fractions
   const float4* __restrict__ rg4 = uniform_ptr(R.rg4);
   const float* __restrict__ dfp = uniform_ptr(R.df);

   const FT* __restrict__ src = uniform_ptr((const FT*)F.data);
   const float off = F.offset;

   // SSD, then per component sum(g e) and sum((g e) f)
   double acc[NV];
#pragma unroll
   for (int v = 0; v < NV; v++) acc[v] = 0.0;

   constexpr int kUnroll = 4;
   const int rpar = cw >= kBlock ? 1 : kBlock / cw;
   const int ro = cw >= kBlock ? 0 : (int)threadIdx.x / cw;
   // the loop is compiled once per sampler (staged tile or global loads), so the choice, which
   // is uniform over the work item, costs no instructions per pixel
   // the interval's last pixel (no Jacobian row) as a row and column of this plane (its row
   // is past the plane's rows when it lies in a later plane)
   const int ep = ch.end_p - pb, erow = ep / nx, ecol = ep - erow * nx;
   auto rows_loop = [&](auto use_tile) {
      constexpr bool TILE = decltype(use_tile)::value;
      for (int c = cx + (int)threadIdx.x - ro * cw; c < cx + cw; c += kBlock) {
         const int rf = (pb + r0 * nx + c >= ch.p0 ? r0 : r0 + 1) + ro;
         const int rl = pb + r1 * nx + c < ch.p1 ? r1 : r1 - 1;
         const int yend = c == ecol ? erow : -1;
         // byte offset in the tile of the pixel's column shifted by -K0 (row and plane offsets
         // are relative to row r0 and the first staged plane, where K1 and K2 put them)
         const int colb = 4 * (c - cx + kPad - Ksh[0]);
         const unsigned pb16 = 16u * (unsigned)(pb + c), pb4 = 4u * (unsigned)(pb + c);
         // one pixel of the column: row y, its reference record and node fraction
         auto pixel = [&](const int y, const float4 rec, const float fsv) {
            const double fd = (double)fsv;
            int ti[3];
            float wt[3];
#pragma unroll
            for (int q = 0; q < M; q++) {
               double fr;
               if constexpr (TILE) {
                  // ti = floor(d) + K (d + K >= 0: truncation is the floor)
                  const double d = fma(fd, Dd[q], Dsh[q]);
                  fr = __builtin_amdgcn_fract(d);
                  ti[q] = (int)d;
               } else {
                  // the staged sampler's d + K, then floor and -K
                  const double d = fma(fd, Dd[q], Dsh[q]);
                  const double fl = floor(d);
                  fr = d - fl;
                  ti[q] = (int)fl - Ksh[q];
               }
               // below 1 after the rounding to float too, so a weight of 1 never hides the
               // other tap
               wt[q] = fminf((float)fr, 0x1.fffffep-1f);
            }
            float v = 0.f;
            if constexpr (TILE) {
               // byte offset in the tile with 24-bit products (full rate; tile offsets are far
               // below 2^23): tile row y - r0 + floor(d1) + K1, column + floor(d0), plane
               // floor(d2) + K2
               int off = __mul24(y - r0 + ti[1], stride4) + colb;
               off += ti[0] << 2;
               if (M == 3) off += __mul24(ti[2], 4 * plane_step);
               const float* L = (const float*)((const char*)lds + off);
               const float top = fmaf(wt[0], L[1] - L[0], L[0]);
               const float bot = fmaf(wt[0], L[S.stride + 1] - L[S.stride], L[S.stride]);
               const float v0 = fmaf(wt[1], bot - top, top);
               if (M == 3) {
                  const float* L2 = L + plane_step;
                  const float top2 = fmaf(wt[0], L2[1] - L2[0], L2[0]);
                  const float bot2 = fmaf(wt[0], L2[S.stride + 1] - L2[S.stride], L2[S.stride]);
                  const float v1 = fmaf(wt[1], bot2 - top2, top2);
                  v = fmaf(wt[2], v1 - v0, v0);
               } else {
                  v = v0;
               }
            } else {
               // the taps are read from clamped, in-bounds addresses and the ones outside the
               // frame replaced by 0 afterwards, so all of them are loads without branches that
               // issue together (non-finite displacements sample nothing)
               const int x0 = c + ti[0], y0 = y + ti[1];
               const int z0 = M == 3 ? plane + ti[2] : 0;
               const int xa = min(max(x0, 0), nx - 1), xb = min(max(x0 + 1, 0), nx - 1);
               const int ya = min(max(y0, 0), ny - 1), yb = min(max(y0 + 1, 0), ny - 1);
               const bool vxa = (unsigned)x0 < (unsigned)nx, vxb = (unsigned)(x0 + 1) < (unsigned)nx;
               const bool vya = (unsigned)y0 < (unsigned)ny && finite;
               const bool vyb = (unsigned)(y0 + 1) < (unsigned)ny && finite;
               auto bilinear = [&](int zz, bool vz) -> float {
                  const int zc = M == 3 ? min(max(zz, 0), nz - 1) : 0;
                  const unsigned ra = (unsigned)((zc * ny + ya) * nx), rb = (unsigned)((zc * ny + yb) * nx);
                  const float a0 = (float)gld(src, ra + xa), a1 = (float)gld(src, ra + xb);
                  const float b0 = (float)gld(src, rb + xa), b1 = (float)gld(src, rb + xb);
                  const float t00 = vz && vya && vxa ? a0 + off : 0.f;
                  const float t01 = vz && vya && vxb ? a1 + off : 0.f;
                  const float t10 = vz && vyb && vxa ? b0 + off : 0.f;
                  const float t11 = vz && vyb && vxb ? b1 + off : 0.f;
                  const float top = fmaf(wt[0], t01 - t00, t00);
                  const float bot = fmaf(wt[0], t11 - t10, t10);
                  return fmaf(wt[1], bot - top, top);
               };
               const float v0 = bilinear(z0, M != 3 || (unsigned)z0 < (unsigned)nz);
               if (M == 3) {
                  const float v1 = bilinear(z0 + 1, (unsigned)(z0 + 1) < (unsigned)nz);
                  v = fmaf(wt[2], v1 - v0, v0);
               } else {
                  v = v0;
               }
            }
            if (MODE == MODE_WARP) {
               wout[pb + y * nx + c] = (v == 0.f) ? invalid : v;
               return;
            }
            // no sample (v == 0): residual 0, outside the SSD; the interval's last pixel has no
            // Jacobian row
            const float e = v != 0.f ? v - rec.x : 0.f;
            const float ej = y != yend ? e : 0.f;
            const double ed = (double)e;
            acc[0] = fma(ed, ed, acc[0]);
            const float g[3] = {rec.y, rec.z, rec.w};
#pragma unroll
            for (int q = 0; q < M; q++) {
               const double ge = (double)(g[q] * ej);
               acc[1 + q] += ge;
               acc[1 + M + q] = fma(ge, fd, acc[1 + M + q]);
            }
         };
         // tile sampling takes kUnroll rows per step (their records loaded before any is used),
         // then the column's last rows one at a time; global sampling (4 or 8 loads per pixel
         // already in flight) one row per step, which keeps it within the register budget
         constexpr int kU = TILE ? kUnroll : 1;
         int r = rf;
         for (; r + (kU - 1) * rpar <= rl; r += kU * rpar) {
            float4 recs[kU];
            float fs[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) {
               const unsigned row = (unsigned)(r + u * rpar);
               recs[u] = gld_bytes(rg4, __umul24(row, nx16) + pb16);
               fs[u] = (M == 3) ? gld_bytes(dfp, __umul24(row, nx4) + pb4) : recs[u].w;
            }
#pragma unroll
            for (int u = 0; u < kU; u++) pixel(r + u * rpar, recs[u], fs[u]);
         }
         if constexpr (kU > 1) {
            for (; r <= rl; r += rpar) {
               const float4 rec = gld_bytes(rg4, __umul24((unsigned)r, nx16) + pb16);
               const float fsv = (M == 3) ? gld_bytes(dfp, __umul24((unsigned)r, nx4) + pb4) : rec.w;
               pixel(r, rec, fsv);
            }
         }
      }
   };
   if (ro < rpar) {
      if (use) rows_loop(std::true_type{});
      else rows_loop(std::false_type{});
   }
   if (MODE == MODE_PASS) block_reduce_lds<M>(acc, (double*)lds, out);
   else __syncthreads();
}

template <int MODE>
__device__ __forceinline__ void dispatch_chunk(const RefDev& R, const FrameDev& F,
                                               const Chunk ch, const double* D, float* lds,
                                               int lds_cap, bool use_lds, double* out,
                                               float* wout, float invalid)
{
   if (R.n_dim == 2) {
      if (F.dtype == FA_U16)
         pass_chunk<2, uint16_t, MODE>(R, F, ch, D, lds, lds_cap, use_lds, out, wout, invalid);
      else
         pass_chunk<2, float, MODE>(R, F, ch, D, lds, lds_cap, use_lds, out, wout, invalid);
   } else {
      if (F.dtype == FA_U16)
         pass_chunk<3, uint16_t, MODE>(R, F, ch, D, lds, lds_cap, use_lds, out, wout, invalid);
      else
         pass_chunk<3, float, MODE>(R, F, ch, D, lds, lds_cap, use_lds, out, wout, invalid);
   }
}

// Over the active work items of the references [ref_lo, ref_hi) (one kind, one LDS class):
// workgroup b takes a contiguous share of the items, with workgroups renumbered so that each
// XCD (workgroups are dealt to the 8 XCDs round robin) owns one contiguous range, so the
// reference chunks and frame rows it reads stay in its L2: within a reference the items are
// chunk-major, so consecutive items read the same reference chunk for different frames. The
// grid is about one workgroup per item when every frame is active, and the hardware
// dispatcher balances them. k_pass<2> and k_pass<3> are compiled for their own register
// budgets: 6 waves per SIMD in 2-D (80 VGPRs), 5 in 3-D (96 VGPRs); left to itself the compiler
// keeps every unrolled pixel's FP64 temporaries live at once and halves the occupancy.
template <int M>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(M == 2 ? 6 : 5, 8))) void k_pass(PassArgs a)
{
   // staging tile, and the work item's reduction (aliasing it; at least 7 x kBlock doubles)
   extern __shared__ float lds[];
   if (a.counters->n_active_frames <= 0) return;
   const int64_t Wb = a.run_base[a.ref_lo];
   const int64_t W = a.run_base[a.ref_hi] - Wb;
   if (W <= 0) return;
   const int G = gridDim.x;
   int lb = blockIdx.x;
   if (G % 8 == 0) lb = (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8;
   const int64_t w0 = Wb + W * lb / G, w1 = Wb + W * (lb + 1) / G;
   int r = 0, gn = 1, gf = 0;
   int64_t rb = 0, re = -1;
   for (int64_t wi = w0; wi < w1; wi++) {
      if (wi >= re || wi < rb) {
         // the last reference whose items start at or before wi (empty groups share the
         // start of the next one and are skipped)
         int lo = 0, hi = a.n_refs - 1;
         while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.run_base[mid] <= wi) lo = mid;
            else hi = mid - 1;
         }
         r = lo;
         rb = a.run_base[r];
         re = a.run_base[r + 1];
         gn = a.run_count[r];
         gf = a.run_first[r];
      }
      const int64_t local = wi - rb;
      const int c = (int)(local / gn);
      const int j = gf + (int)(local - (int64_t)c * gn);
      const BatchFrame& bf = a.frames[a.active_frames[j]];
      const RefDev& R = bf.r;
      const Chunk ch = R.chunks[c];
      double* out = a.partials + (bf.work_base + c) * kPassStride;
      if (bf.f.dtype == FA_U16)
         pass_chunk<M, uint16_t, MODE_PASS>(R, bf.f, ch, a.Dtrial + bf.d_base, lds, a.lds_cap,
                                            a.use_lds != 0, out, nullptr, 0.f);
      else
         pass_chunk<M, float, MODE_PASS>(R, bf.f, ch, a.Dtrial + bf.d_base, lds, a.lds_cap,
                                         a.use_lds != 0, out, nullptr, 0.f);
   }
}

// Frame blockIdx.y of a launch over several: its samples fstride bytes, its D dstride doubles
// and its output ostride floats after the first's.
__global__ __launch_bounds__(kBlock) void k_warp(RefDev R, FrameDev F, const double* D,
                                                 float* out, float invalid, int lds_cap,
                                                 int use_lds, int64_t fstride, int64_t dstride, int64_t ostride)
{
   extern __shared__ float lds[];
   F.data = (const char*)F.data + blockIdx.y * fstride;
   D += blockIdx.y * dstride;
   out += blockIdx.y * ostride;
   for (int c = blockIdx.x; c < R.n_chunks; c += gridDim.x) {
      dispatch_chunk<MODE_WARP>(R, F, R.chunks[c], D, lds, lds_cap, use_lds != 0, nullptr, out,
                                invalid);
   }
}

// ---- forward splat (AbstractFrameWarper::warpImage, FrameWarper.cpp:71-136) -------------
// Each source pixel p is pushed to p - d(p) with trilinear weights into wimg (weight x value)
// and coverage (weight). loc1 stays at loc0 on the last row/column/plane, and only loc0 must
// lie in the volume (z checked only with more than one plane). The clamp of loc at
// FrameWarper.cpp:87-94 runs after loc0 and the fractions are taken and has no effect.
// Weights are formed in FP64; the float accumulation uses atomics, so the summation order
// differs from the reference's scan order.
template <int M, typename FT>
__device__ void splat_chunk(const RefDev& R, const FrameDev& F, const Chunk ch, const double* D,
                            float* wimg, float* cov, float offset)
{
   const int nx = R.nx, ny = R.ny, nz = R.nz;
   const int64_t plane_sz = (int64_t)nx * ny;
   const int k = ch.k;
   const int plane = (int)(ch.p0 / plane_sz);
   const int q0 = (int)(ch.p0 - plane * plane_sz);
   PixelWalk w;
   w.init(q0 + threadIdx.x, nx);
   for (int p = ch.p0 + threadIdx.x; p < ch.p1; p += kBlock, w.step(nx)) {
      if (w.x < ch.x0 || w.x >= ch.x0 + ch.w) continue;  // strip item: only its columns
      const double fd = (double)(M == 3 ? R.df[p] : R.rg4[p].w);
      double l[3];
      l[0] = (double)w.x - (fd * D[3 * (k + 1) + 0] + (1.0 - fd) * D[3 * k + 0]);
      l[1] = (double)w.y - (fd * D[3 * (k + 1) + 1] + (1.0 - fd) * D[3 * k + 1]);
      l[2] = M == 3 ? (double)plane - (fd * D[3 * (k + 1) + 2] + (1.0 - fd) * D[3 * k + 2]) : 0.0;
      if (!(fabs(l[0]) < 1e9 && fabs(l[1]) < 1e9 && fabs(l[2]) < 1e9)) continue;
      const double fl[3] = {floor(l[0]), floor(l[1]), floor(l[2])};
      const int x0 = (int)fl[0], y0 = (int)fl[1], z0 = (int)fl[2];
      const double fx = l[0] - fl[0], fy = l[1] - fl[1], fz = l[2] - fl[2];
      if (x0 < 0 || y0 < 0 || x0 > nx - 1 || y0 > ny - 1) continue;
      if (nz > 1 && (z0 < 0 || z0 > nz - 1)) continue;
      const int x1 = x0 < nx - 1 ? x0 + 1 : x0, y1 = y0 < ny - 1 ? y0 + 1 : y0;
      const int z1 = z0 < nz - 1 ? z0 + 1 : z0;
      const double v = (double)((float)((const FT*)F.data)[p] + offset);
      const int zz[2] = {z0, z1}, yy[2] = {y0, y1}, xx[2] = {x0, x1};
      const double wz[2] = {1 - fz, fz}, wy[2] = {1 - fy, fy}, wx[2] = {1 - fx, fx};
#pragma unroll
      for (int a = 0; a < (M == 3 ? 2 : 1); a++)
#pragma unroll
         for (int b = 0; b < 2; b++)
#pragma unroll
            for (int c = 0; c < 2; c++) {
               const double wt = wy[b] * wx[c] * wz[a];
               const int64_t o = ((int64_t)zz[a] * ny + yy[b]) * nx + xx[c];
               atomicAdd(&wimg[o], (float)(wt * v));
               atomicAdd(&cov[o], (float)wt);
            }
   }
}

__global__ __launch_bounds__(kBlock) void k_splat(RefDev R, FrameDev F, const double* D, float* wimg,
                                                  float* cov, float offset, int64_t fstride, int64_t dstride,
                                                  int64_t ostride)
{
   F.data = (const char*)F.data + blockIdx.y * fstride;
   D += blockIdx.y * dstride;
   wimg += blockIdx.y * ostride;
   cov += blockIdx.y * ostride;
   for (int c = blockIdx.x; c < R.n_chunks; c += gridDim.x) {
      const Chunk ch = R.chunks[c];
      if (R.n_dim == 2) {
         if (F.dtype == FA_U16) splat_chunk<2, uint16_t>(R, F, ch, D, wimg, cov, offset);
         else splat_chunk<2, float>(R, F, ch, D, wimg, cov, offset);
      } else {
         if (F.dtype == FA_U16) splat_chunk<3, uint16_t>(R, F, ch, D, wimg, cov, offset);
         else splat_chunk<3, float>(R, F, ch, D, wimg, cov, offset);
      }
   }
}

// ---- scheduling: compact the active frames and their work items -----------------------
// The active frames are grouped by reference (a counting sort), so the items of a range of
// references are contiguous and each pass launch takes those of its class. A reference's
// work items are chunk-major: the active frames of one reference process a chunk one after
// the other, so its reference record is read from HBM once per group rather than per frame.
// The order within a group does not affect any result (each item's sums go to the partials
// of its own frame and chunk). account = 0 only refreshes the counters the host polls.
__global__ __launch_bounds__(1024) void k_schedule(BatchDev B, int account)
{
   __shared__ int32_t s_cnt[kMaxRefs];
   __shared__ int32_t s_fill[kMaxRefs];
   __shared__ int32_t s_wf[1024 / 64];
   __shared__ int64_t s_ww[1024 / 64], s_ww2[1024 / 64];
   __shared__ unsigned long long s_bytes, s_px;
   const int t = threadIdx.x, T = blockDim.x;
   const int n = B.n_frames, R = B.n_refs;
   for (int r = t; r < R; r += T) {
      s_cnt[r] = 0;
      s_fill[r] = 0;
   }
   if (t == 0) {
      s_bytes = 0;
      s_px = 0;
   }
   __syncthreads();
   unsigned long long bytes = 0, px = 0;
   for (int f = t; f < n; f += T) {
      if (B.state[f].done) continue;
      const int r = B.frames[f].ref;
      atomicAdd(&s_cnt[r], 1);
      const RefDev& RD = B.refs[r];
      bytes += (unsigned long long)RD.n_px * (unsigned long long)pass_bytes_per_px(B.frames[f].f.dtype, RD.n_dim);
      px += (unsigned long long)RD.n_px;
   }
   if (bytes) {
      atomicAdd(&s_bytes, bytes);
      atomicAdd(&s_px, px);
   }
   __syncthreads();
   // exclusive scans over the references (each thread a contiguous range of them)
   const int r0 = (int)((int64_t)R * t / T), r1 = (int)((int64_t)R * (t + 1) / T);
   int32_t fs = 0;
   int64_t ws = 0, w2s = 0;
   for (int r = r0; r < r1; r++) {
      const int64_t w = (int64_t)s_cnt[r] * B.refs[r].n_chunks;
      fs += s_cnt[r];
      ws += w;
      if (B.refs[r].n_dim == 2) w2s += w;
   }
   // inclusive scans of the threads' sums within each wave (lane shuffles), then the waves'
   // totals before this wave
   const int lane = t & 63, wv = t >> 6;
   int32_t f_in = fs;
   int64_t w_in = ws, w2_in = w2s;
#pragma unroll
   for (int off = 1; off < 64; off <<= 1) {
      const int32_t a = __shfl_up(f_in, off, 64);
      const int64_t b = __shfl_up(w_in, off, 64), c = __shfl_up(w2_in, off, 64);
      if (lane >= off) {
         f_in += a;
         w_in += b;
         w2_in += c;
      }
   }
   if (lane == 63) {
      s_wf[wv] = f_in;
      s_ww[wv] = w_in;
      s_ww2[wv] = w2_in;
   }
   __syncthreads();
   for (int q = 0; q < wv; q++) {
      f_in += s_wf[q];
      w_in += s_ww[q];
      w2_in += s_ww2[q];
   }
   int32_t fb = f_in - fs;
   int64_t wb = w_in - ws;
   for (int r = r0; r < r1; r++) {
      const int32_t c = s_cnt[r];
      B.run_first[r] = fb;
      B.run_count[r] = c;
      B.run_base[r] = wb;
      s_cnt[r] = fb;  // now the group's first position
      fb += c;
      wb += (int64_t)c * B.refs[r].n_chunks;
   }
   if (t == T - 1) {
      B.run_base[R] = w_in;
      B.counters->n_active_frames = f_in;
      B.counters->n_active_work = w_in;
      B.counters->n_active_work2 = w2_in;
      if (account) {
         B.counters->bytes += s_bytes;
         B.counters->px += s_px;
         B.counters->rounds += 1;
      }
   }
   __syncthreads();
   for (int f = t; f < n; f += T) {
      if (B.state[f].done) continue;
      const int r = B.frames[f].ref;
      B.active_frames[s_cnt[r] + atomicAdd(&s_fill[r], 1)] = f;
   }
}

// ---- device trust region ---------------------------------------------------------------
// dlib find_min_trust_region + objective_delta_stop_strategy, as the reference calls it
// (FrameWarpAligner.cpp:205-209), over f = SSD, g = Jtr, B = H (OptimisationModel.cpp:45-63).
// The subproblem is Nocedal & Wright Alg. 4.3 on the block-tridiagonal H.
//
// One workgroup per frame: one wave (64 lanes), or 256 threads for long node chains. Every
// thread runs the same scalar logic on workgroup-reduced values, so the control flow is
// uniform; vector work is spread over the threads, with a barrier wherever one thread reads
// what another wrote (implicit within one wave). The factorisations of
// H + lambda I and the triangular solves use block cyclic reduction (cr_factor, cr_fwd,
// cr_bwd): ceil(log2 nD) + 1 parallel levels instead of a sequential sweep over the nodes.

struct TrScratch {
   double *x, *p, *g;
   double *y;             // solve vector of the subproblem
   double *fa, *fc, *fw;  // factor blocks A, C, W of H + lambda I
};

// Sum over the wave, the same value in every lane: lane pairs, quads, halves of rows and rows
// of 16 by DPP moves, then rows 0+1 and 2+3 (row_bcast:15) and the two halves (row_bcast:31),
// which leave the total in lane 63, read back as a scalar. A fixed order, and every step a
// DPP move instead of an LDS permute.
__device__ __forceinline__ double wave_sum(double v)
{
   v = dpp_add<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
   v = dpp_add<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
   v = dpp_add<0x141, 0xf>(v);  // row_half_mirror
   v = dpp_add<0x140, 0xf>(v);  // row_mirror
   v = dpp_add<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
   v = dpp_add<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
   const unsigned long long b = (unsigned long long)__double_as_longlong(v);
   const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
   const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
   return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the NT threads of the workgroup, the same value in every thread: the waves' sums in
// wave order. Every thread calls it at the same point (the update's control flow is uniform).
template <int NT>
__device__ __forceinline__ double block_sum(double v)
{
   v = wave_sum(v);
   if constexpr (NT == 64) {
      return v;
   } else {
      __shared__ double red[NT / 64];
      if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
      __syncthreads();
      double s = red[0];
#pragma unroll
      for (int w = 1; w < NT / 64; w++) s += red[w];
      __syncthreads();
      return s;
   }
}

template <int NT>
__device__ __forceinline__ double block_dot(int n, const double* a, const double* b)
{
   double s = 0;
   for (int i = threadIdx.x; i < n; i += NT) s += a[i] * b[i];
   return block_sum<NT>(s);
}

// p' H p over the workgroup.
template <int M, int NT>
__device__ __forceinline__ double block_quad(const RefDev& R, const double* p)
{
   constexpr int MM = M * M;
   const int nb = R.nD;
   double s = 0;
   for (int k = threadIdx.x; k < nb; k += NT) {
#pragma unroll
      for (int a = 0; a < M; a++)
#pragma unroll
         for (int b = 0; b < M; b++) s += p[k * M + a] * R.Hdiag[k * MM + a * M + b] * p[k * M + b];
      if (k + 1 < nb)
#pragma unroll
         for (int a = 0; a < M; a++)
#pragma unroll
            for (int b = 0; b < M; b++)
               s += 2.0 * p[(k + 1) * M + a] * R.Hsub[k * MM + a * M + b] * p[k * M + b];
   }
   return block_sum<NT>(s);
}

template <int M, int NT>
__device__ int tr_subproblem(const RefDev& R, const TrScratch& S, const double* g, double radius,
                             double eps, int max_iter, double* p, double* lam_out)
{
   constexpr int MM = M * M;
   const int nb = R.nD, n = nb * M;
   const double gnorm = sqrt(block_dot<NT>(n, g, g));
   for (int i = threadIdx.x; i < n; i += NT) p[i] = 0;
   *lam_out = 0;
   __syncthreads();
   if (gnorm == 0) return 1;
   // (H + lam I) p = -g through the factor (A, C, W); ||p|| and, unless the iteration stops
   // here, ||L^-1 p|| and the new bracket. It is called with the factor of H itself (global
   // memory) or the one just formed (LDS or global scratch); each call site then sees one
   // memory space for the factor blocks.
   double lam = 0, lam_lo = 0, lam_hi;
   // fwd_done: y already holds L^-1 (-g) (formed with the factor)
   auto solve = [&](const double* A, const double* C, const double* W, bool fwd_done, double& pn,
                    double& qn) -> bool {
      double* y = S.y;
      if (!fwd_done) {
         for (int i = threadIdx.x; i < n; i += NT) y[i] = -g[i];
         __syncthreads();
         cr_fwd<M, NT>(nb, A, C, W, y);
      }
      cr_bwd<M, NT>(nb, A, C, W, y);
      for (int i = threadIdx.x; i < n; i += NT) p[i] = y[i];
      *lam_out = lam;
      pn = sqrt(block_dot<NT>(n, y, y));
      if (lam == 0 && pn <= radius) return true;
      if (fabs(pn - radius) <= eps * radius) return true;
      if (pn > radius) lam_lo = lam;
      else lam_hi = lam;
      // q = L^-1 p (y holds p)
      cr_fwd<M, NT>(nb, A, C, W, y);
      qn = sqrt(block_dot<NT>(n, y, y));
      return false;
   };
   const double eig_lo = R.eig_lo;
   lam_hi = gnorm / radius + (eig_lo < 0 ? -eig_lo : 0);
   int it;
   for (it = 0; it < max_iter; it++) {
      const bool use_h = lam == 0 && R.h_pd;
      if (!use_h) {
         // factor H + lam I and, in the same levels, the forward solve of -g
         bool ok = lam != 0;
         if (ok) {
            for (int i = threadIdx.x; i < n; i += NT) S.y[i] = -g[i];
            __syncthreads();
            ok = cr_factor<M, NT>(nb, R.Hdiag, R.Hsub, lam, S.fa, S.fc, S.fw, S.y);
         }
         if (!ok) {
            lam_lo = lam;
            lam = 0.5 * (lam_lo + lam_hi);
            continue;
         }
      }
      double pn = 0, qn = 0;
      const bool stop = use_h ? solve(R.L0, R.L0 + (int64_t)nb * MM, R.L0 + 2 * (int64_t)nb * MM, false, pn, qn)
                              : solve(S.fa, S.fc, S.fw, true, pn, qn);
      if (stop) {
         it++;
         break;
      }
      double lam_new = lam + (pn / qn) * (pn / qn) * (pn - radius) / radius;
      if (!(lam_new > lam_lo && lam_new < lam_hi)) lam_new = 0.5 * (lam_lo + lam_hi);
      lam = lam_new;
   }
   __syncthreads();
   return it;
}

// The same subproblem iteration through H = Q diag(ev) Q^T: with qv = Q^T g,
// ||p(lam)||^2 = sum qv_i^2 / (ev_i + lam)^2 and ||L^-1 p||^2 = sum qv_i^2 / (ev_i + lam)^3 for
// the Cholesky factor L of H + lam I, which exists when ev_min + lam > 0 (lam = 0: when H's own
// factorisation succeeded). Each iteration is two wave reductions instead of a factorisation
// and three block-triangular solves; p = -Q (qv / (ev + lam)) is formed once, for the last
// lambda at which the factor exists.
template <int NT>
__device__ int tr_subproblem_eig(const RefDev& R, const TrScratch& S, const double* qs, const double* g,
                                 double radius, double eps, int max_iter, double* p, double* lam_out)
{
   const int n = R.nD * R.n_dim;
   double* qv = S.fa;  // the factor scratch is free on this path
   const double gnorm = sqrt(block_dot<NT>(n, g, g));
   for (int i = threadIdx.x; i < n; i += NT) p[i] = 0;
   *lam_out = 0;
   __syncthreads();
   if (gnorm == 0) return 1;
   // qv = Q^T g: lane i walks column i (lanes read consecutive doubles); eight loads are
   // issued before their products, so the column's loads overlap
   const double* __restrict__ Qm = R.Q;
   const double* __restrict__ Qt = R.Q + (int64_t)n * n;
   if (qs) {
      // Q staged in LDS: lane i reads column i (consecutive addresses)
      for (int i = threadIdx.x; i < n; i += NT) {
         double s = 0;
         int r = 0;
         for (; r + 8 <= n; r += 8) {
            double q[8], h[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
               q[u] = qs[(r + u) * n + i];
               h[u] = g[r + u];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) s += q[u] * h[u];
         }
         for (; r < n; r++) s += qs[r * n + i] * g[r];
         qv[i] = s;
      }
   } else
   for (int i = threadIdx.x; i < n; i += NT) {
      double s = 0;
      int r = 0;
      for (; r + 8 <= n; r += 8) {
         double q[8];
#pragma unroll
         for (int u = 0; u < 8; u++) q[u] = Qm[(int64_t)(r + u) * n + i];
#pragma unroll
         for (int u = 0; u < 8; u++) s += q[u] * g[r + u];
      }
      for (; r < n; r++) s += Qm[(int64_t)r * n + i] * g[r];
      qv[i] = s;
   }
   __syncthreads();
   const double eig_lo = R.eig_lo;
   double lam = 0, lam_lo = 0, lam_hi = gnorm / radius + (eig_lo < 0 ? -eig_lo : 0);
   double lam_sol = -1;  // lambda of the last solution (none yet)
   int it;
   for (it = 0; it < max_iter; it++) {
      const bool ok = lam == 0 ? R.h_pd != 0 : R.ev_min + lam > 0;
      if (!ok) {
         lam_lo = lam;
         lam = 0.5 * (lam_lo + lam_hi);
         continue;
      }
      double s2 = 0, s3 = 0;
      for (int i = threadIdx.x; i < n; i += NT) {
         const double d = 1.0 / (R.ev[i] + lam);
         const double a = qv[i] * d;
         s2 += a * a;
         s3 += a * a * d;
      }
      const double pn = sqrt(block_sum<NT>(s2));
      lam_sol = lam;
      if (lam == 0 && pn <= radius) {
         it++;
         break;
      }
      if (fabs(pn - radius) <= eps * radius) {
         it++;
         break;
      }
      if (pn > radius) lam_lo = lam;
      else lam_hi = lam;
      const double qn = sqrt(block_sum<NT>(s3));
      double lam_new = lam + (pn / qn) * (pn / qn) * (pn - radius) / radius;
      if (!(lam_new > lam_lo && lam_new < lam_hi)) lam_new = 0.5 * (lam_lo + lam_hi);
      lam = lam_new;
   }
   if (lam_sol >= 0) {
      double* w = S.fa + n;
      for (int i = threadIdx.x; i < n; i += NT) w[i] = qv[i] / (R.ev[i] + lam_sol);
      __syncthreads();
      // p = -Q w: lane r walks row r of Q (in LDS), or column r of Q^T (global)
      if (qs) {
         for (int r = threadIdx.x; r < n; r += NT) {
            double s = 0;
            int i = 0;
            for (; i + 8 <= n; i += 8) {
               double q[8], h[8];
#pragma unroll
               for (int u = 0; u < 8; u++) {
                  q[u] = qs[r * n + i + u];
                  h[u] = w[i + u];
               }
#pragma unroll
               for (int u = 0; u < 8; u++) s += q[u] * h[u];
            }
            for (; i < n; i++) s += qs[r * n + i] * w[i];
            p[r] = -s;
         }
      } else
      for (int r = threadIdx.x; r < n; r += NT) {
         double s = 0;
         int i = 0;
         for (; i + 8 <= n; i += 8) {
            double q[8];
#pragma unroll
            for (int u = 0; u < 8; u++) q[u] = Qt[(int64_t)(i + u) * n + r];
#pragma unroll
            for (int u = 0; u < 8; u++) s += q[u] * w[i + u];
         }
         for (; i < n; i++) s += Qt[(int64_t)i * n + r] * w[i];
         p[r] = -s;
      }
      *lam_out = lam_sol;
   }
   __syncthreads();
   return it;
}

// dst[0, n) = src[0, n) (doubles) by direct global-to-LDS loads: each wave instruction moves
// 128 doubles (16 bytes per lane) into LDS without passing through registers, so all of them
// are in flight at once. src and dst are 16-byte aligned; for odd n the last lane reads and
// writes one double past n (the caller's buffers have it). The caller waits (vmcnt) and
// synchronises before reading dst.
template <int NT>
__device__ __forceinline__ void dma_to_lds(double* dst, const double* __restrict__ src, int n)
{
   const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
   for (int b = 128 * wv; b < n; b += 128 * (NT / 64))
      if (b + 2 * ln < n)
         __builtin_amdgcn_global_load_lds((const void*)(src + b + 2 * ln),
                                          (__attribute__((address_space(3))) void*)(dst + b), 16, 0, 0);
}

// dst[0, n) = src[0, n) by the NT threads, eight loads per thread issued before their stores
// (a loop of load-store pairs would wait out one load latency per element).
template <int NT, typename T>
__device__ __forceinline__ void copy_to_lds(T* dst, const T* __restrict__ src, int n)
{
   for (int i0 = threadIdx.x; i0 < n; i0 += 8 * NT) {
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = i0 + u * NT < n ? src[i0 + u * NT] : T(0);
#pragma unroll
      for (int u = 0; u < 8; u++)
         if (i0 + u * NT < n) dst[i0 + u * NT] = v[u];
   }
}

// dst[0, n) = src[0, n) for doubles or 32-bit values: direct loads when both ends allow them
// (16-byte aligned doubles: the even part by dma_to_lds and an odd last double by one thread;
// 32-bit values one per lane), else through registers. The caller waits and synchronises.
template <int NT>
__device__ __forceinline__ void stage_to_lds(double* dst, const double* __restrict__ src, int n)
{
   if ((((uintptr_t)dst | (uintptr_t)src) & 15) == 0) {
      dma_to_lds<NT>(dst, src, n & ~1);
      if ((n & 1) && threadIdx.x == 0) dst[n - 1] = src[n - 1];
   } else {
      copy_to_lds<NT>(dst, src, n);
   }
}

template <int NT>
__device__ __forceinline__ void stage_to_lds(int32_t* dst, const int32_t* __restrict__ src, int n)
{
   const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
   for (int b = 64 * wv; b < n; b += 64 * (NT / 64))
      if (b + ln < n)
         __builtin_amdgcn_global_load_lds((const void*)(src + b + ln),
                                          (__attribute__((address_space(3))) void*)(dst + b), 4, 0, 0);
}

__device__ __forceinline__ TrScratch scratch_of(const BatchDev& B, const BatchFrame& bf)
{
   TrScratch S;
   double* v = B.vec + bf.vec_base;
   const int64_t n = bf.n;
   S.x = B.x + bf.x_base;
   S.p = v;
   S.g = v + n;
   S.y = v + 2 * n;
   const RefDev& R = bf.r;
   const int64_t nbmm = (int64_t)R.nD * R.n_dim * R.n_dim;
   S.fa = B.fac + bf.fac_base;
   S.fc = S.fa + nbmm;
   S.fw = S.fa + 2 * nbmm;
   return S;
}

// ---- per-frame reductions of the pass partials ------------------------------------------
// SSD: lanes take chunks strided, then a fixed butterfly; Jtr: one node per lane, chunks of
// interval k-1 (weight f) and interval k (weight 1-f) in order. Fixed order, so the result
// does not depend on which other frames share the launch.
template <int NT>
__device__ __forceinline__ double reduce_ssd(const RefDev& R, const double* P)
{
   double s = 0;
   for (int c = threadIdx.x; c < R.n_chunks; c += NT) s += P[(int64_t)c * kPassStride];
   return block_sum<NT>(s);
}

template <int NT>
__device__ __forceinline__ void reduce_jtr(const RefDev& R, const double* P, double* g)
{
   const int m = R.n_dim;
   for (int k = threadIdx.x; k < R.nD; k += NT) {
      double acc[3] = {0, 0, 0};
      if (k >= 1)
         for (int c = R.ichunk[k - 1]; c < R.ichunk[k]; c++)
            for (int d = 0; d < m; d++) acc[d] += P[(int64_t)c * kPassStride + 1 + m + d];
      if (k <= R.nD - 2)
         for (int c = R.ichunk[k]; c < R.ichunk[k + 1]; c++)
            for (int d = 0; d < m; d++) acc[d] += P[(int64_t)c * kPassStride + 1 + d];
      for (int d = 0; d < m; d++) g[k * m + d] = acc[d];
   }
}

__global__ __launch_bounds__(64) void k_reduce_eval(BatchDev B)
{
   const int j = blockIdx.x;
   const BatchFrame& bf = B.frames[j];
   const RefDev& R = bf.r;
   const double* P = B.partials + bf.work_base * kPassStride;
   const double f = reduce_ssd<64>(R, P);
   double* g = B.vec + bf.vec_base + (int64_t)bf.n;
   reduce_jtr<64>(R, P, g);
   if (threadIdx.x == 0) B.state[j].f = f;
}

// Starts frame j: x (already uploaded) is the first evaluation point.
__device__ __forceinline__ void init_frame(const BatchDev& B, const int j, const int eval_only)
{
   const BatchFrame& bf = B.frames[j];
   const RefDev& R = bf.r;
   const double* x = B.x + bf.x_base;
   double* Dt = B.Dtrial + bf.d_base;
   const int m = R.n_dim;
   if (!eval_only)
      for (int i = threadIdx.x; i < R.nD; i += blockDim.x)
         for (int c = 0; c < 3; c++) Dt[3 * i + c] = c < m ? x[i * m + c] : 0.0;
   if (threadIdx.x == 0) {
      TrState st = {};
      st.radius = B.prm.radius;
      st.radius0 = B.prm.radius;
      st.phase = 0;
      B.state[j] = st;
   }
}

__global__ void k_tr_init(BatchDev B, int eval_only) { init_frame(B, blockIdx.x, eval_only); }

// Starts slots of a serving batch from pinned host staging, read directly over the host link
// (no copy engine, whose queues other streams share): slot hslots[i] takes the frame record
// hbf[i] and the start vector hx[i * nmax ...]; its trial displacement and trust-region state
// are then reset as in init_frame.
__global__ void k_tr_start_slots(BatchDev B, const BatchFrame* hbf, const double* hx, const int32_t* hslots,
                                 int nmax)
{
   const int i = blockIdx.x;
   const int j = hslots[i];
   const BatchFrame* src = hbf + i;
   const int m = src->r.n_dim, nD = src->r.nD;
   const double* xs = hx + (size_t)i * nmax;
   double* x = B.x + src->x_base;
   double* Dt = B.Dtrial + src->d_base;
   for (int k = threadIdx.x; k < nD; k += blockDim.x)
      for (int c = 0; c < 3; c++) {
         const double v = c < m ? xs[k * m + c] : 0.0;
         if (c < m) x[k * m + c] = v;
         Dt[3 * k + c] = v;
      }
   // the record word by word (it is a multiple of 8 bytes)
   const uint64_t* ws = (const uint64_t*)src;
   uint64_t* wd = (uint64_t*)(B.frames + j);
   for (int k = threadIdx.x; k < (int)(sizeof(BatchFrame) / 8); k += blockDim.x) wd[k] = ws[k];
   if (threadIdx.x == 0) {
      TrState st = {};
      st.radius = B.prm.radius;
      st.radius0 = B.prm.radius;
      st.phase = 0;
      B.state[j] = st;
   }
}

// Copies a serving batch's trust-region states and parameter vectors into pinned host memory,
// written directly over the host link (no copy engine).
__global__ void k_publish(BatchDev B, TrState* hstate, double* hx, int64_t total_x)
{
   const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, T = (int64_t)gridDim.x * blockDim.x;
   const int64_t ns = (int64_t)B.n_frames * (int64_t)(sizeof(TrState) / 8);
   const uint64_t* ws = (const uint64_t*)B.state;
   uint64_t* wd = (uint64_t*)hstate;
   for (int64_t k = t; k < ns; k += T) wd[k] = ws[k];
   for (int64_t k = t; k < total_x; k += T) hx[k] = B.x[k];
   __threadfence_system();
}

// LDS use of the update (full = 1): the frame's pass partials, H, ichunk, x, p, g, the factor
// blocks and the solve vector are all staged with one burst of loads, so the update's many
// small dependent steps run at LDS latency; x, p, g are written back at the end. full = 0:
// only the factor blocks and the solve vector (read and written at every level of the cyclic
// reduction) are in LDS.
template <int M, bool LDS, int NT, bool FULL>
__device__ void tr_update_frame(const BatchDev& B, const int j, double* smem)
{
   constexpr int MM = M * M;
   const BatchFrame& bf = B.frames[j];
   RefDev R = bf.r;
   const TrParamsDev& prm = B.prm;
   const TrScratch G = scratch_of(B, bf);
   TrScratch S = G;
   const double* P = B.partials + bf.work_base * kPassStride;
   const int n = bf.n, nb = R.nD;
   const double* qs = nullptr;  // Q in LDS (eigendecomposition path, full staging)
   if constexpr (LDS) {
      double* q = smem;
      S.fa = q;
      S.fc = q + nb * MM;
      S.fw = q + 2 * nb * MM;
      S.y = q + 3 * nb * MM;
      q += 3 * nb * MM + n;
      if constexpr (FULL) {
         const int np = R.n_chunks * kPassStride;
         double* pl = q;
         double* hd = pl + np;
         double* hs = hd + nb * MM;
         S.x = hs + nb * MM;
         S.p = S.x + n;
         S.g = S.p + n;
         int32_t* ich = (int32_t*)(S.g + n);
         // every array is requested before any is waited for: direct global-to-LDS loads
         // wherever the addresses allow them, so the whole working set arrives in about one
         // memory latency
         stage_to_lds<NT>(pl, P, np);
         if (R.q_lds) {
            // Q after the interval table, at the next 16-byte boundary
            const uintptr_t a = (uintptr_t)((double*)ich + (nb + 1) / 2);
            double* qd = (double*)((a + 15) & ~(uintptr_t)15);
            stage_to_lds<NT>(qd, R.Q, n * n);
            qs = qd;
            double* evd = qd + ((n * n + 1) & ~1);
            stage_to_lds<NT>(evd, R.ev, n);
            R.ev = evd;
         }
         stage_to_lds<NT>(hd, R.Hdiag, nb * MM);
         stage_to_lds<NT>(hs, R.Hsub, nb * MM);
         stage_to_lds<NT>(S.x, G.x, n);
         stage_to_lds<NT>(S.p, G.p, n);
         stage_to_lds<NT>(S.g, G.g, n);
         stage_to_lds<NT>(ich, R.ichunk, nb);
         __builtin_amdgcn_s_waitcnt(0);  // the direct loads have landed in LDS
         __syncthreads();
         P = pl;
         R.Hdiag = hd;
         R.Hsub = hs;
         R.ichunk = ich;
      }
   }
   const double f_new = reduce_ssd<NT>(R, P);
   TrState st = B.state[j];
   int accept = 0;
   double pred = 0, meas = 0, rho = 0;  // of the step this evaluation tried
   if (st.phase == 0) {
      st.f = f_new;
      st.evaluations = 1;
      st.stale = 0;
      accept = 1;  // first point: take f and g, no step bookkeeping
   } else {
      st.evaluations++;
      const double gp = block_dot<NT>(n, S.g, S.p);
      pred = -(0.5 * block_quad<M, NT>(R, S.p) + gp);
      meas = st.f - f_new;
      if (fabs(pred) <= fabs(meas) * DBL_EPSILON) {
         st.status = FA_STOP_NO_PROGRESS;
         st.done = 1;
      } else {
         rho = meas / fabs(pred);
         if (!isfinite(rho)) {
            st.status = FA_STOP_NONFINITE;
            st.done = 1;
            rho = 0;
         } else {
            int collapse = 0;
            if (rho < 0.25) {
               st.radius *= 0.25;
               if (st.radius <= st.radius0 * DBL_EPSILON) collapse = 1;
            } else if (rho > 0.75 && st.sub_iters > 1) {
               st.radius = 2 * st.radius < 1000 ? 2 * st.radius : 1000;
            }
            if (collapse) {
               // the radius has collapsed: stop where x is, without taking the step
            } else if (rho > 0) {
               for (int i = threadIdx.x; i < n; i += NT) S.x[i] = S.x[i] + S.p[i];
               st.f = f_new;
               st.stale = 0;
               accept = 1;
            } else {
               st.stale = 1;
               st.rejected++;
               if (prm.skip_repeat && st.lambda == 0 && !collapse) {
                  // the unchanged interior step would be proposed, evaluated and rejected
                  // again while it fits the shrunken region
                  const double pn = sqrt(block_dot<NT>(n, S.p, S.p));
                  while (pn <= st.radius) {
                     st.radius *= 0.25;
                     st.rejected++;
                     if (st.radius <= st.radius0 * DBL_EPSILON) {
                        collapse = 1;
                        break;
                     }
                  }
               }
            }
            if (collapse) {
               st.status = FA_STOP_RADIUS;
               st.done = 1;
            }
         }
      }
   }
   if (accept) reduce_jtr<NT>(R, P, S.g);
   __syncthreads();
   if (!st.done && !st.stale) {
      st.iterations++;
      if (st.been_used) {
         const double df = fabs(st.f - st.prev_f);
         if (prm.max_iter && st.iterations > prm.max_iter) {
            st.status = FA_STOP_MAX_ITER;
            st.done = 1;
         } else if (df < prm.min_delta) {
            st.status = FA_STOP_ABS_DELTA;
            st.done = 1;
         } else if (prm.rel_delta > 0 && df < prm.rel_delta * fabs(st.prev_f)) {
            st.status = FA_STOP_REL_DELTA;
            st.done = 1;
         }
      }
      st.been_used = 1;
      st.prev_f = st.f;
   }
   if (!st.done && prm.max_evals && st.evaluations >= prm.max_evals) {
      st.status = FA_STOP_MAX_EVALS;
      st.done = 1;
   }
   const bool batched_eig = R.eig == 2 && B.split_eig;
   if (st.done && st.phase > 1) st.phase = 1;  // a phase the batched subproblem kernels left
   if (!st.done && batched_eig) {
      // the subproblem runs in the batched eigen-space kernels that follow the update (g is
      // written back to the frame's vectors below when it was staged in LDS)
      st.phase = kPhaseEigPending;
   } else if (!st.done) {
      double lam;
      st.sub_iters = R.eig == 1
                        ? tr_subproblem_eig<NT>(R, S, qs, S.g, st.radius, prm.sub_eps, prm.sub_max_iter, S.p, &lam)
                        : tr_subproblem<M, NT>(R, S, S.g, st.radius, prm.sub_eps, prm.sub_max_iter, S.p, &lam);
      st.lambda = lam;
      st.phase = 1;
      double* Dt = B.Dtrial + bf.d_base;
      for (int i = threadIdx.x; i < nb; i += NT)
#pragma unroll
         for (int c = 0; c < 3; c++) Dt[3 * i + c] = c < M ? S.x[i * M + c] + S.p[i * M + c] : 0.0;
   }
   if (B.trace && st.evaluations - 1 < B.trace_cap) {
      // record of this evaluation (layout of kTraceFields), then x, g, p; a subproblem left to
      // the batched kernels gets its fields and p from k_eig_trace
      __syncthreads();
      const bool later = st.phase == kPhaseEigPending;
      const bool ran = !st.done && !later;
      const double pn = ran ? sqrt(block_dot<NT>(n, S.p, S.p)) : 0.0;
      const double gn = sqrt(block_dot<NT>(n, S.g, S.g));
      const double gp = ran ? block_dot<NT>(n, S.g, S.p) : 0.0;
      double* o = B.trace + (int64_t)j * B.trace_cap * B.trace_rec + (int64_t)(st.evaluations - 1) * B.trace_rec;
      if (threadIdx.x == 0) {
         const double v[kTraceFields] = {(double)st.evaluations, f_new, st.f, pred, meas, rho, (double)accept,
                                         st.radius, (double)st.iterations, st.done ? (double)st.status : 0.0,
                                         ran ? st.lambda : 0.0, ran ? (double)st.sub_iters : 0.0, pn, gn, gp,
                                         (double)st.rejected};
         for (int i = 0; i < kTraceFields; i++) o[i] = v[i];
      }
      for (int i = threadIdx.x; i < n; i += NT) {
         o[kTraceFields + i] = S.x[i];
         o[kTraceFields + n + i] = S.g[i];
         if (!later) o[kTraceFields + 2 * n + i] = S.p[i];
      }
   }
   if (threadIdx.x == 0) B.state[j] = st;
   if constexpr (LDS && FULL) {
      __syncthreads();
      for (int i = threadIdx.x; i < n; i += NT) {
         G.x[i] = S.x[i];
         G.p[i] = S.p[i];
         G.g[i] = S.g[i];
      }
   }
}

// One launch per update class (frames of one workgroup width whose factor blocks need a
// similar amount of LDS): the frames of other classes return at once. Long node chains use 256
// threads per frame (the first levels of the cyclic reduction have nD/2 independent blocks),
// short ones one wave; the width is a property of the reference, so a frame's result does
// not depend on the batch it runs in.
// The schedule of the next round for a batch of one reference, run by the update's last
// workgroup (B.fold): every workgroup publishes its frame's done flag with a device-scope
// atomic store, waits for it to complete and counts itself done with an atomic; the one that
// completes the count reads the flags with device-scope atomic loads (no cache-wide fences)
// and lists the frames still running in frame order, with the counters k_schedule keeps. The
// first wave does it with ballots; the count is reset for the next round.
__device__ void fold_schedule(const BatchDev& B)
{
   __shared__ int s_last;
   __syncthreads();  // the frame's update (its state is written by thread 0) is complete
   if (threadIdx.x == 0) {
      int32_t* done = &B.state[blockIdx.x].done;
      __hip_atomic_store(done, *done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __builtin_amdgcn_s_waitcnt(0);  // the flag is visible device-wide before the count
      s_last = __hip_atomic_fetch_add(&B.counters->upd_done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
               (int)gridDim.x - 1;
   }
   __syncthreads();
   if (!s_last || threadIdx.x >= 64) return;
   const int lane = threadIdx.x, n = B.n_frames;
   const RefDev& R = B.refs[0];
   int total = 0;
   unsigned long long bytes = 0;
   for (int base = 0; base < n; base += 64) {
      const int f = base + lane;
      const bool act = f < n && !__hip_atomic_load(&B.state[f].done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned long long m = __ballot(act);
      if (act) {
         B.active_frames[total + __popcll(m & ((1ull << lane) - 1))] = f;
         bytes += (unsigned long long)R.n_px * (unsigned long long)pass_bytes_per_px(B.frames[f].f.dtype, R.n_dim);
      }
      total += __popcll(m);
   }
#pragma unroll
   for (int off = 32; off >= 1; off >>= 1) bytes += __shfl_xor(bytes, off, 64);
   if (lane == 0) {
      const int64_t w = (int64_t)total * R.n_chunks;
      B.run_first[0] = 0;
      B.run_count[0] = total;
      B.run_base[0] = 0;
      B.run_base[1] = w;
      B.counters->n_active_frames = total;
      B.counters->n_active_work = w;
      B.counters->n_active_work2 = R.n_dim == 2 ? w : 0;
      if (total > 0) {
         B.counters->bytes += bytes;
         B.counters->px += (unsigned long long)total * (unsigned long long)R.n_px;
         B.counters->rounds += 1;
      }
      __hip_atomic_store(&B.counters->upd_done, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   }
}

template <bool LDS, int NT, bool FULL>
__global__ __launch_bounds__(NT) void k_tr_update(BatchDev B, int cls)
{
   extern __shared__ double smem[];
   const int j = blockIdx.x;
   if (!B.state[j].done && B.frames[j].trc == cls) {
      if (B.frames[j].r.n_dim == 2) tr_update_frame<2, LDS, NT, FULL>(B, j, smem);
      else tr_update_frame<3, LDS, NT, FULL>(B, j, smem);
   }
   if (B.fold) fold_schedule(B);
}

// One frame's update in the variant its class launches (NT threads, FULL staging).
template <bool LDS, int NT>
__device__ __forceinline__ void tr_update_variant(const BatchDev& B, const int j, const bool full, double* smem)
{
   const bool three = B.frames[j].r.n_dim == 3;
   if constexpr (LDS) {
      if (full) {
         if (three) tr_update_frame<3, true, NT, true>(B, j, smem);
         else tr_update_frame<2, true, NT, true>(B, j, smem);
         return;
      }
   }
   if (three) tr_update_frame<3, LDS, NT, false>(B, j, smem);
   else tr_update_frame<2, LDS, NT, false>(B, j, smem);
}

// Every update class in one launch, for batches whose workgroups are all resident at once (the
// update is latency-bound, so class launches one after another would add their latencies).
// Bit c of wide / full: class c runs 256 threads / stages its whole working set. A frame of a
// one-wave class runs in the first wave and the other three waves end at once (a workgroup
// barrier waits only for the waves that have not ended), so every frame runs the same code as
// in its class's own launch.
template <bool LDS>
__global__ __launch_bounds__(256) void k_tr_update_all(BatchDev B, uint32_t wide, uint32_t full)
{
   extern __shared__ double smem[];
   const int j = blockIdx.x;
   if (B.state[j].done) return;
   const int c = B.frames[j].trc;
   const bool f = (full >> c) & 1;
   if ((wide >> c) & 1) {
      tr_update_variant<LDS, 256>(B, j, f, smem);
   } else {
      if (__builtin_amdgcn_readfirstlane(threadIdx.x) >= 64) return;
      tr_update_variant<LDS, 64>(B, j, f, smem);
   }
}

// ---- launchers ----------------------------------------------------------------------------

// ---- batched subproblems through H = Q diag(ev) Q^T (references with RefDev::eig == 2) ----
// A frame whose update left its subproblem pending (kPhaseEigPending) takes it in three steps
// over the whole batch: qv = Q^T g for all of a reference's pending frames at once
// (k_eig_gemm<false>: Q is read once per round rather than once per frame), the secular
// iteration of tr_subproblem_eig on qv (k_eig_secular), then p = -Q w and the trial
// displacement x + p (k_eig_gemm<true>, k_eig_apply); k_eig_trace completes the trace
// records when the batch keeps one. The frame's next update takes phases kPhaseEigW and
// kPhaseEigZero as 1 (a step was proposed). Here Q is the matrix of eigenvectors (column i for
// ev[i]); RefDev::Q stores Q row-major, then Q^T (as for the in-update path).

typedef double dbl4 __attribute__((ext_vector_type(4)));
// 16-row blocks of the operator; 64 frames per workgroup; k split over kGemmSplit waves of a
// workgroup and kGemmGroups workgroups. A frame's factor scratch holds the kGemmGroups
// partial products (n each), then w (n): 5 n doubles, within its 3 nD n_dim^2 (>= 6 n).
constexpr int kGemmRows = 16, kGemmCols = 64, kGemmSplit = 4, kGemmGroups = 4;

// Partial products of tile blockIdx.x of B.eig_tiles (without a table: the tiles of reference
// 0 for one frame, in eig_tiles_of's order): over k group kg, for output rows
// [row0, row0 + 16) of reference r (eig == 2), for the frames among entries [base, base + 64)
// of its active list that are in phase kPhaseEigPending
// (TRANS false: qv = Q^T g) or kPhaseEigW (TRANS true: Q w, to be negated into p). Partial kg goes to the frame's factor scratch at kg n. FP64
// MFMA 16x16x4 steps: wave w takes columns [16 (w & 3), +16) and the k range w >> 2 of the
// group's four, whose sums are added in k order. An output element depends only on its row of
// the operator and its frame's vector, summed in a fixed order, so a frame's result does not
// depend on the other frames. Lanes of rows past n or columns past the frames read row n - 1 or
// the first frame's vector; their outputs are not stored.
template <bool TRANS>
__global__ __launch_bounds__(256 * kGemmSplit) void k_eig_gemm(BatchDev B)
{
   int r = 0, row0 = (blockIdx.x / kGemmGroups) * kGemmRows, base = 0, kg = blockIdx.x % kGemmGroups;
   if (B.eig_tiles) {
      const int32_t* tl = B.eig_tiles + 4 * (int64_t)blockIdx.x;
      r = tl[0];
      row0 = tl[1];
      base = tl[2];
      kg = tl[3];
   }
   const RefDev& R = B.refs[r];
   const int n = R.nD * R.n_dim;
   const int first = B.run_first[r], count = B.run_count[r];
   if (base >= count) return;
   __shared__ int cols[kGemmCols];
   __shared__ int s_nc;
   __shared__ dbl4 part[kGemmSplit - 1][4][64];
   const int t = threadIdx.x, lane = t & 63, w = t >> 6;
   const int want = TRANS ? kPhaseEigW : kPhaseEigPending;
   if (t < 64) {
      const int i = base + t;
      const int j = i < count ? B.active_frames[first + i] : -1;
      const bool take = j >= 0 && B.state[j].phase == want;
      const unsigned long long m = __ballot(take);
      if (take) cols[__popcll(m & ((1ull << t) - 1))] = j;
      if (t == 0) s_nc = __popcll(m);
   }
   __syncthreads();
   const int nc = s_nc;
   const int cb = w & 3, ks = kg * kGemmSplit + (w >> 2);  // k range ks of 16
   const bool busy = cb * 16 < nc;  // uniform over the wave
   dbl4 acc = {0.0, 0.0, 0.0, 0.0};
   if (busy) {
      const int li = lane & 15, lk = lane >> 4;
      // A operand: lane (li, lk) holds A[li][lk] of each 16x4 step, B operand: B[lk][li]
      const int row = min(row0 + li, n - 1);
      const int c = min(cb * 16 + li, nc - 1);
      const BatchFrame& bf = B.frames[cols[c]];
      const double* __restrict__ v = TRANS ? B.fac + bf.fac_base + kGemmGroups * n : B.vec + bf.vec_base + n;
      // GEMM1 reads V (V[k n + i] = eigenvector i at k), GEMM2 V^T: either way the lanes of
      // one k read 16 consecutive rows
      const double* __restrict__ qa = R.Q + (TRANS ? (int64_t)n * n : 0) + row;
      const int64_t qs = n;
      constexpr int KS = kGemmSplit * kGemmGroups;
      const int span = ((n + 4 * KS - 1) / (4 * KS)) * 4;
      const int k_lo = min(n, ks * span), k_hi = min(n, k_lo + span);
      int k = k_lo;
      dbl4 acc1 = {0.0, 0.0, 0.0, 0.0};
      // eight steps' operands are loaded before their products, so their loads overlap
      constexpr int U = 8;
      for (; k + 4 * U <= k_hi; k += 4 * U) {
         double a[U], b[U];
#pragma unroll
         for (int u = 0; u < U; u++) {
            a[u] = qa[(k + 4 * u + lk) * qs];
            b[u] = v[k + 4 * u + lk];
         }
#pragma unroll
         for (int u = 0; u < U; u += 2) {
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u + 1], b[u + 1], acc1, 0, 0, 0);
         }
      }
      for (; k < k_hi; k += 4) {
         const bool in = k + lk < k_hi;
         const double a0 = in ? qa[(k + lk) * qs] : 0.0, b0 = in ? v[k + lk] : 0.0;
         acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc, 0, 0, 0);
      }
      acc += acc1;
   }
   if (w >= 4) part[(w >> 2) - 1][cb][lane] = acc;
   __syncthreads();
   if (w < 4 && busy) {
#pragma unroll
      for (int q = 0; q < kGemmSplit - 1; q++) acc += part[q][cb][lane];
      // lane holds D[(lane >> 4) + 4 e][lane & 15], e = 0..3 (the FP64 16x16x4 output layout)
      const int c = cb * 16 + (lane & 15);
      if (c < nc) {
         double* o = B.fac + B.frames[cols[c]].fac_base + (int64_t)kg * n;
#pragma unroll
         for (int e = 0; e < 4; e++) {
            const int row = row0 + (lane >> 4) + 4 * e;
            if (row < n) o[row] = acc[e];
         }
      }
   }
}

// p = -(sum of the kGemmGroups partial products of Q w, in order) and the trial displacement
// x + p of frame eig_frames[blockIdx.x] when it is in phase kPhaseEigW.
__global__ __launch_bounds__(1024) void k_eig_apply(BatchDev B)
{
   const int j = B.eig_frames ? B.eig_frames[blockIdx.x] : blockIdx.x;
   if (B.state[j].phase != kPhaseEigW) return;
   const BatchFrame& bf = B.frames[j];
   const int n = bf.n, m = bf.r.n_dim;
   const double* P = B.fac + bf.fac_base;
   for (int i = threadIdx.x; i < n; i += 1024) {
      double s = P[i];
#pragma unroll
      for (int q = 1; q < kGemmGroups; q++) s += P[(int64_t)q * n + i];
      const double pv = -s;
      B.vec[bf.vec_base + i] = pv;
      const int node = i / m, comp = i - node * m;
      double* Dt = B.Dtrial + bf.d_base + 3 * node;
      Dt[comp] = B.x[bf.x_base + i] + pv;
      if (m == 2 && comp == 1) Dt[2] = 0.0;
   }
}

// The secular iteration of tr_subproblem_eig for a frame in phase kPhaseEigPending, from
// qv = Q^T g (the sum of k_eig_gemm<false>'s partial products), held in registers with the
// eigenvalues (at most kEigRegs per thread): lambda and the iteration count into its state,
// and either w = qv / (ev + lambda) after the partial products (phase kPhaseEigW; then
// k_eig_gemm<true> and k_eig_apply form p and x + p) or,
// when the iteration found no lambda at which H + lambda I is positive definite or g = 0,
// p = 0 and the trial displacement x (phase kPhaseEigZero).
constexpr int kEigRegs = (kEigBatchMax + 255) / 256;

__global__ __launch_bounds__(256) void k_eig_secular(BatchDev B)
{
   constexpr int NT = 256;
   const int j = B.eig_frames ? B.eig_frames[blockIdx.x] : blockIdx.x;
   TrState st = B.state[j];
   if (st.phase != kPhaseEigPending) return;
   const BatchFrame& bf = B.frames[j];
   const RefDev& R = bf.r;
   const int n = bf.n, t = threadIdx.x;
   const double* g = B.vec + bf.vec_base + n;
   const double* P = B.fac + bf.fac_base;  // the partial products of qv, in order
   double qr[kEigRegs], er[kEigRegs];
   double gg = 0;  // g'g in block_dot's per-thread order
#pragma unroll
   for (int u = 0; u < kEigRegs; u++) {
      const int i = t + NT * u;
      double q = 0.0;
      if (i < n) {
         q = P[i];
#pragma unroll
         for (int k = 1; k < kGemmGroups; k++) q += P[(int64_t)k * n + i];
         gg += g[i] * g[i];
      }
      qr[u] = q;
      er[u] = i < n ? R.ev[i] : 1.0;
   }
   const double radius = st.radius, eps = B.prm.sub_eps;
   const int max_iter = B.prm.sub_max_iter;
   const double gnorm = sqrt(block_sum<NT>(gg));
   double lam = 0, lam_lo = 0, lam_hi = 0, lam_sol = -1;
   int it = 1;
   if (gnorm != 0) {
      lam_hi = gnorm / radius + (R.eig_lo < 0 ? -R.eig_lo : 0);
      for (it = 0; it < max_iter; it++) {
         const bool ok = lam == 0 ? R.h_pd != 0 : R.ev_min + lam > 0;
         if (!ok) {
            lam_lo = lam;
            lam = 0.5 * (lam_lo + lam_hi);
            continue;
         }
         // the sums of tr_subproblem_eig, in the same per-thread order (i = t, t + NT, ...)
         double s2 = 0, s3 = 0;
#pragma unroll
         for (int u = 0; u < kEigRegs; u++) {
            if (t + NT * u < n) {
               const double d = 1.0 / (er[u] + lam);
               const double a = qr[u] * d;
               s2 += a * a;
               s3 += a * a * d;
            }
         }
         const double pn = sqrt(block_sum<NT>(s2));
         lam_sol = lam;
         if (lam == 0 && pn <= radius) {
            it++;
            break;
         }
         if (fabs(pn - radius) <= eps * radius) {
            it++;
            break;
         }
         if (pn > radius) lam_lo = lam;
         else lam_hi = lam;
         const double qn = sqrt(block_sum<NT>(s3));
         double lam_new = lam + (pn / qn) * (pn / qn) * (pn - radius) / radius;
         if (!(lam_new > lam_lo && lam_new < lam_hi)) lam_new = 0.5 * (lam_lo + lam_hi);
         lam = lam_new;
      }
   }
   if (lam_sol >= 0) {
      double* w = B.fac + bf.fac_base + kGemmGroups * n;
#pragma unroll
      for (int u = 0; u < kEigRegs; u++) {
         const int i = t + NT * u;
         if (i < n) w[i] = qr[u] / (er[u] + lam_sol);
      }
      st.lambda = lam_sol;
      st.phase = kPhaseEigW;
   } else {
      double* p = B.vec + bf.vec_base;
      const double* x = B.x + bf.x_base;
      double* Dt = B.Dtrial + bf.d_base;
      const int m = R.n_dim;
      for (int i = t; i < n; i += NT) p[i] = 0.0;
      for (int i = t; i < R.nD; i += NT)
         for (int c = 0; c < 3; c++) Dt[3 * i + c] = c < m ? x[i * m + c] + 0.0 : 0.0;
      st.lambda = 0;
      st.phase = kPhaseEigZero;
   }
   st.sub_iters = it;
   if (t == 0) B.state[j] = st;
}

// Trace records of the frames (eig_frames) whose subproblem the batched kernels solved in this round
// (phase kPhaseEigW or kPhaseEigZero): lambda, iterations, |p|, g'p and p (the update wrote the
// rest of the record). Launched only when the batch records a trace.
__global__ __launch_bounds__(256) void k_eig_trace(BatchDev B)
{
   constexpr int NT = 256;
   const int j = B.eig_frames ? B.eig_frames[blockIdx.x] : blockIdx.x;
   const TrState st = B.state[j];
   if (st.phase != kPhaseEigW && st.phase != kPhaseEigZero) return;
   if (st.evaluations - 1 >= B.trace_cap) return;
   const BatchFrame& bf = B.frames[j];
   const int n = bf.n;
   const double* p = B.vec + bf.vec_base;
   const double* g = p + n;
   const double pn = sqrt(block_dot<NT>(n, p, p));
   const double gp = block_dot<NT>(n, g, p);
   double* o = B.trace + (int64_t)j * B.trace_cap * B.trace_rec + (int64_t)(st.evaluations - 1) * B.trace_rec;
   if (threadIdx.x == 0) {
      o[10] = st.lambda;
      o[11] = (double)st.sub_iters;
      o[12] = pn;
      o[14] = gp;
   }
   for (int i = threadIdx.x; i < n; i += NT) o[kTraceFields + 2 * n + i] = p[i];
}

// dst = src^T for an n x n row-major matrix, through 32 x 32 tiles in LDS.
__global__ __launch_bounds__(256) void k_transpose(const double* __restrict__ src, double* __restrict__ dst, int n)
{
   __shared__ double tile[32][33];
   const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
   const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
   for (int r = ty; r < 32; r += 8)
      if (by + r < n && bx + tx < n) tile[r][tx] = src[(int64_t)(by + r) * n + bx + tx];
   __syncthreads();
   for (int r = ty; r < 32; r += 8)
      if (bx + r < n && by + tx < n) dst[(int64_t)(bx + r) * n + by + tx] = tile[tx][r];
}

hipError_t launch_transpose(const double* src, double* dst, int n, hipStream_t s)
{
   const dim3 g((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32));
   hipLaunchKernelGGL(k_transpose, g, dim3(256), 0, s, src, dst, n);
   return hipGetLastError();
}

void eig_tiles_of(int ref, int n, int frames, std::vector<int32_t>& tiles)
{
   for (int base = 0; base < frames; base += kGemmCols)
      for (int row0 = 0; row0 < n; row0 += kGemmRows)
         for (int kg = 0; kg < kGemmGroups; kg++) {
            tiles.push_back(ref);
            tiles.push_back(row0);
            tiles.push_back(base);
            tiles.push_back(kg);
         }
}

hipError_t launch_eig_subproblems(const BatchDev& B, hipStream_t s)
{
   if (B.n_eig_tiles <= 0 || B.n_eig_frames <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_eig_gemm<false>, dim3(B.n_eig_tiles), dim3(256 * kGemmSplit), 0, s, B);
   hipLaunchKernelGGL(k_eig_secular, dim3(B.n_eig_frames), dim3(256), 0, s, B);
   hipLaunchKernelGGL(k_eig_gemm<true>, dim3(B.n_eig_tiles), dim3(256 * kGemmSplit), 0, s, B);
   hipLaunchKernelGGL(k_eig_apply, dim3(B.n_eig_frames), dim3(1024), 0, s, B);
   if (B.trace) hipLaunchKernelGGL(k_eig_trace, dim3(B.n_eig_frames), dim3(256), 0, s, B);
   return hipGetLastError();
}

static inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_node_map(const ScanDev& sc, int nx, int ny, int64_t n, uint16_t* Di, float* Df,
                           hipStream_t s)
{
   hipLaunchKernelGGL(k_node_map, dim3(blocks_for(n, 256)), dim3(256), 0, s, sc, nx, ny, n, Di, Df);
   return hipGetLastError();
}

hipError_t launch_interval_begin(const uint16_t* Di, int64_t n, int nD, int32_t* begin,
                                 int32_t* flags, hipStream_t s)
{
   hipLaunchKernelGGL(k_interval_begin, dim3(blocks_for(n, 256)), dim3(256), 0, s, Di, n, nD, begin,
                      flags);
   return hipGetLastError();
}

hipError_t launch_reference_record(const float* T, int nz, int ny, int nx, int n_dim,
                                   const ScanDev& sc, float4* rg4, float* df, hipStream_t s)
{
   int64_t n = (int64_t)nz * ny * nx;
   hipLaunchKernelGGL(k_reference_record, dim3(blocks_for(n, 256)), dim3(256), 0, s, T, nz, ny, nx,
                      n_dim, sc, rg4, df);
   return hipGetLastError();
}

hipError_t launch_hessian(const RefDev& R, double* partials, double* Hdiag, double* Hsub,
                          double* L0, double* info, hipStream_t s)
{
   unsigned g = (unsigned)std::min<int64_t>(R.n_chunks, 65535);
   hipLaunchKernelGGL(k_hess_partials, dim3(g), dim3(kBlock), 0, s, R, partials);
   hipError_t e = hipGetLastError();
   if (e != hipSuccess) return e;
   hipLaunchKernelGGL(k_hess_assemble, dim3(1), dim3(kHessThreads), 0, s, R, (const double*)partials,
                      Hdiag, Hsub, L0, info);
   return hipGetLastError();
}

hipError_t launch_eigen(const RefDev& R, double* Q, double* ev, double* info, hipStream_t s)
{
   const int n0 = R.nD * R.n_dim, m = n0 + (n0 & 1);
   const size_t bytes = (2 * (size_t)m * m + m) * sizeof(double);
   hipLaunchKernelGGL(k_eigen, dim3(1), dim3(kEigThreads), bytes, s, R, Q, ev, info);
   return hipGetLastError();
}

size_t pass_lds_bytes(int lds_cap, int n_dim)
{
   return std::max((size_t)lds_cap * sizeof(float), (size_t)(1 + 2 * n_dim) * kBlock * sizeof(double));
}

hipError_t launch_pass(const PassArgs& a, int n_dim, int grid, hipStream_t s)
{
   const size_t lds = pass_lds_bytes(a.lds_cap, n_dim);
   if (n_dim == 2)
      hipLaunchKernelGGL(k_pass<2>, dim3(grid), dim3(kBlock), lds, s, a);
   else
      hipLaunchKernelGGL(k_pass<3>, dim3(grid), dim3(kBlock), lds, s, a);
   return hipGetLastError();
}

hipError_t launch_warp(const RefDev& R, const FrameDev& F, const double* D, float* out,
                       float invalid, int lds_cap, int use_lds, int grid, hipStream_t s, int count)
{
   // frames of one launch lie one after the other: samples, nD x 3 displacements, outputs
   const int64_t fstride = R.n_px * (F.dtype == FA_U16 ? 2 : 4);
   hipLaunchKernelGGL(k_warp, dim3(grid, count), dim3(kBlock), (size_t)lds_cap * sizeof(float), s, R, F, D,
                      out, invalid, lds_cap, use_lds, fstride, (int64_t)3 * R.nD, R.n_px);
   return hipGetLastError();
}

hipError_t launch_splat(const RefDev& R, const FrameDev& F, const double* D, float* wimg, float* cov,
                        float offset, int grid, hipStream_t s, int count)
{
   const int64_t fstride = R.n_px * (F.dtype == FA_U16 ? 2 : 4);
   hipLaunchKernelGGL(k_splat, dim3(grid, count), dim3(kBlock), 0, s, R, F, D, wimg, cov, offset, fstride,
                      (int64_t)3 * R.nD, R.n_px);
   return hipGetLastError();
}

hipError_t launch_schedule(const BatchDev& B, int account, hipStream_t s)
{
   hipLaunchKernelGGL(k_schedule, dim3(1), dim3(1024), 0, s, B, account);
   return hipGetLastError();
}

hipError_t launch_tr_init(const BatchDev& B, int eval_only, hipStream_t s)
{
   hipLaunchKernelGGL(k_tr_init, dim3(B.n_frames), dim3(64), 0, s, B, eval_only);
   return hipGetLastError();
}

hipError_t launch_tr_start_slots(const BatchDev& B, const BatchFrame* hbf, const double* hx, const int32_t* hslots,
                                 int n, int nmax, hipStream_t s)
{
   hipLaunchKernelGGL(k_tr_start_slots, dim3(n), dim3(64), 0, s, B, hbf, hx, hslots, nmax);
   return hipGetLastError();
}

hipError_t launch_publish(const BatchDev& B, TrState* hstate, double* hx, int64_t total_x, hipStream_t s)
{
   const int64_t words = (int64_t)B.n_frames * (int64_t)(sizeof(TrState) / 8) + total_x;
   const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(64, (words + 255) / 256));
   hipLaunchKernelGGL(k_publish, dim3(grid), dim3(256), 0, s, B, hstate, hx, total_x);
   return hipGetLastError();
}

// Allows the LDS variant of the trust-region update up to max_lds_bytes of dynamic LDS.
hipError_t configure_kernels(int max_lds_bytes)
{
   (void)hipFuncSetAttribute((const void*)k_eigen, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds_bytes);
   const void* fns[] = {(const void*)k_tr_update<true, 64, false>, (const void*)k_tr_update<true, 64, true>,
                        (const void*)k_tr_update<true, 256, false>, (const void*)k_tr_update<true, 256, true>,
                        (const void*)k_tr_update_all<true>};
   hipError_t e = hipSuccess;
   for (const void* f : fns) {
      const hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds_bytes);
      if (r != hipSuccess) e = r;
   }
   return e;
}

// The update's variants: factor blocks in global memory (lds_doubles = 0), in LDS, or the whole
// working set in LDS (full); 64 or 256 threads per frame.
hipError_t launch_tr_update(const BatchDev& B, int cls, int lds_doubles, int full, int threads, hipStream_t s)
{
   const dim3 grid(B.n_frames);
   const size_t lds = (size_t)lds_doubles * sizeof(double);
   if (threads == 256) {
      if (lds_doubles <= 0) hipLaunchKernelGGL((k_tr_update<false, 256, false>), grid, dim3(256), 0, s, B, cls);
      else if (full) hipLaunchKernelGGL((k_tr_update<true, 256, true>), grid, dim3(256), lds, s, B, cls);
      else hipLaunchKernelGGL((k_tr_update<true, 256, false>), grid, dim3(256), lds, s, B, cls);
   } else {
      if (lds_doubles <= 0) hipLaunchKernelGGL((k_tr_update<false, 64, false>), grid, dim3(64), 0, s, B, cls);
      else if (full) hipLaunchKernelGGL((k_tr_update<true, 64, true>), grid, dim3(64), lds, s, B, cls);
      else hipLaunchKernelGGL((k_tr_update<true, 64, false>), grid, dim3(64), lds, s, B, cls);
   }
   return hipGetLastError();
}

// Workgroups of k_tr_update_all one compute unit holds at once with this much dynamic LDS
// (registers and LDS both count), at least 1.
int tr_update_all_per_cu(int lds_doubles)
{
   int n = 0;
   const void* f = lds_doubles > 0 ? (const void*)k_tr_update_all<true> : (const void*)k_tr_update_all<false>;
   const size_t lds = lds_doubles > 0 ? (size_t)lds_doubles * sizeof(double) : 0;
   if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, 256, lds) != hipSuccess) {
      (void)hipGetLastError();
      return 1;
   }
   return n > 1 ? n : 1;
}

// All classes in one launch (k_tr_update_all): LDS of the largest class, or global (0).
hipError_t launch_tr_update_all(const BatchDev& B, int lds_doubles, uint32_t wide, uint32_t full, hipStream_t s)
{
   const dim3 grid(B.n_frames);
   if (lds_doubles <= 0)
      hipLaunchKernelGGL((k_tr_update_all<false>), grid, dim3(256), 0, s, B, wide, 0u);
   else
      hipLaunchKernelGGL((k_tr_update_all<true>), grid, dim3(256), (size_t)lds_doubles * sizeof(double), s, B,
                         wide, full);
   return hipGetLastError();
}

hipError_t launch_reduce_eval(const BatchDev& B, hipStream_t s)
{
   hipLaunchKernelGGL(k_reduce_eval, dim3(B.n_frames), dim3(64), 0, s, B);
   return hipGetLastError();
}

}  // namespace fa
