/*
 * CPU oracle for the frame-aligner warp-registration hot path.
 * TEST INFRASTRUCTURE ONLY: the checker, never the product. See fa_oracle.h.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the flimfit/frame-aligner tree). The code is written
 * from the behaviour, not copied.
 */
#include "fa_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* Axis convention of the reference: volumes are (z, y, x), x fastest
   (Cv3dUtils.cpp:6-8: X=2, Y=1, Z=0). */
enum { AZ = 0, AY = 1, AX = 2 };

static __thread char g_err[256];
const char* orc_last_error(void) { return g_err; }
static int fail(const char* msg)
{
   snprintf(g_err, sizeof g_err, "%s", msg);
   return -1;
}

typedef struct {
   double line, interline, pixel, frame, interframe, stack;
   int nx, ny, nz;
} scan_t;

/* OffsetVector<Point3f> (FrameWarper.h:58-90): values for indices [i0, i1),
   size i1-i0, or 0 when i0 > i1. at(i) returns zero outside [i0, i1]. */
typedef struct {
   int64_t i0, i1, n;
   double* v; /* 3 components per element */
} offvec_t;

struct orc_warper {
   int fp64;
   int dims[3];
   int n_dim, nD;
   int64_t n_px;
   scan_t scan;      /* after processScanParameters */
   scan_t scan_in;   /* as constructed (ImageScanParameters ctor) */
   float* reference;
   uint16_t* Di;
   double* Df;
   int32_t* D_range; /* nD x (begin, end) */
   double *gx, *gy, *gz;
   offvec_t* vi;     /* nD */
   double* H;        /* dense (nD*n_dim)^2 */
   double* Hdiag;    /* block-tridiagonal copy of H: nD blocks of n_dim^2 */
   double* Hsub;     /* nD-1 blocks: block (k+1, k) */
};

/* Rounds to float where the reference stores or computes a float. */
static inline double rs(const orc_warper* w, double v) { return w->fp64 ? v : (double)(float)v; }

/* ImageScanParameters ctor (ImageScanParameters.h:7-20). */
static scan_t scan_ctor(double line, double interline, double interframe, int nx, int ny, int nz)
{
   scan_t s;
   s.line = line;
   s.interline = interline;
   s.interframe = interframe;
   s.nx = nx < 1 ? 1 : nx;
   s.ny = ny < 1 ? 1 : ny;
   s.nz = nz;
   if (s.interline < s.line) s.interline = 1.01 * s.line;
   s.pixel = s.line / s.nx;
   s.frame = s.ny * s.interline;
   s.stack = (s.nz - 1) * s.interframe + s.frame;
   return s;
}

/* AbstractFrameWarper::processScanParameters (FrameWarper.cpp:223-236): caps the
   gap between planes at one node spacing. */
static scan_t process_scan(const scan_t* in, const int* dims, int nD)
{
   scan_t p = *in;
   double frame_interval_max = dims[AZ] * in->frame / (nD - dims[AZ]);
   if ((in->interframe - in->frame) > frame_interval_max) {
      p.interframe = p.frame + frame_interval_max;
      p.stack = (dims[AZ] - 1) * p.interframe + p.frame;
   }
   return p;
}

orc_warper* orc_create(int fp64)
{
   orc_warper* w = (orc_warper*)calloc(1, sizeof(orc_warper));
   w->fp64 = fp64 ? 1 : 0;
   return w;
}

static void free_ref(orc_warper* w)
{
   free(w->reference);
   free(w->Di);
   free(w->Df);
   free(w->D_range);
   free(w->gx);
   free(w->gy);
   free(w->gz);
   if (w->vi)
      for (int i = 0; i < w->nD; i++) free(w->vi[i].v);
   free(w->vi);
   free(w->H);
   free(w->Hdiag);
   free(w->Hsub);
   w->reference = NULL;
   w->Di = NULL;
   w->Df = NULL;
   w->D_range = NULL;
   w->gx = w->gy = w->gz = NULL;
   w->vi = NULL;
   w->H = w->Hdiag = w->Hsub = NULL;
}

void orc_destroy(orc_warper* w)
{
   if (!w) return;
   free_ref(w);
   free(w);
}

/* AbstractFrameWarper::precomputeInterp (FrameWarper.cpp:239-293). Scan-order time of each
   pixel -> node coordinate u = t/stack*(nD-1); Di = integer part, Df = fractional part.
   D_range[i].begin is the first pixel of interval i, D_range[i].end the last one
   (inclusive); consumers iterate [begin, end), which drops each interval's last pixel. */
static int precompute_interp(orc_warper* w)
{
   const scan_t p = w->scan;
   const int nD = w->nD;
   const int nx = w->dims[AX], ny = w->dims[AY], nz = w->dims[AZ];
   for (int i = 0; i < nD; i++) w->D_range[2 * i] = w->D_range[2 * i + 1] = 0;
   int i = 0, last_i = -1;
   for (int z = 0; z < nz; z++)
      for (int y = 0; y < ny; y++)
         for (int x = 0; x < nx; x++) {
            double t = z * p.interframe + y * p.interline + x * p.pixel;
            double ip;
            double f = modf(t / p.stack * (nD - 1), &ip);
            i = (int)ip;
            int64_t idx = ((int64_t)z * ny + y) * nx + x;
            w->Di[idx] = (uint16_t)i;
            w->Df[idx] = rs(w, f);
            if (i > last_i) {
               w->D_range[2 * i] = (int32_t)idx;
               if (last_i >= 0) w->D_range[2 * last_i + 1] = w->D_range[2 * i] - 1;
               last_i = i;
            }
         }
   if (i != nD - 2) return fail("Something went wrong");
   w->D_range[2 * i + 1] = (int32_t)(w->n_px - 1);

   for (int k = 0; k < nD; k++) {
      int a = k - 1 < 0 ? 0 : k - 1;
      int b = k < nD - 2 ? k : nD - 2;
      offvec_t* v = &w->vi[k];
      v->i0 = w->D_range[2 * a];
      v->i1 = w->D_range[2 * b + 1];
      v->n = v->i1 - v->i0;
      if (v->i0 > v->i1) v->n = 0;
      v->v = (double*)calloc((size_t)(v->n > 0 ? v->n : 1) * 3, sizeof(double));
   }
   return 0;
}

/* Sobel3d.cpp:8-58: one 3-tap pass along an axis, with two-tap edge rules, into a separate
   buffer. The Z pass copies its input when there is a single plane (Sobel3d.cpp:40-44). */
static void filt(const orc_warper* w, const double* m, double* out, int axis, const float* c3,
                 const float* e1, const float* e2)
{
   const int nz = w->dims[AZ], ny = w->dims[AY], nx = w->dims[AX];
   const int64_t sx = 1, sy = nx, sz = (int64_t)nx * ny;
   int n;
   int64_t st;
   if (axis == AX) { n = nx; st = sx; }
   else if (axis == AY) { n = ny; st = sy; }
   else { n = nz; st = sz; }
   if (axis == AZ && nz == 1) {
      memcpy(out, m, sizeof(double) * (size_t)w->n_px);
      return;
   }
   for (int z = 0; z < nz; z++)
      for (int y = 0; y < ny; y++)
         for (int x = 0; x < nx; x++) {
            int64_t idx = (int64_t)z * sz + (int64_t)y * sy + x;
            int pos = axis == AX ? x : (axis == AY ? y : z);
            double r;
            if (pos > 0 && pos < n - 1) {
               double a = rs(w, (double)c3[0] * m[idx - st]);
               double b = rs(w, (double)c3[1] * m[idx]);
               double c = rs(w, (double)c3[2] * m[idx + st]);
               r = rs(w, rs(w, a + b) + c);
            } else if (pos == 0) {
               double a = rs(w, (double)e1[0] * m[idx]);
               double b = rs(w, (double)e1[1] * m[idx + st]);
               r = rs(w, a + b);
            } else {
               double a = rs(w, (double)e2[0] * m[idx]);
               double b = rs(w, (double)e2[1] * m[idx - st]);
               r = rs(w, a + b);
            }
            out[idx] = r;
         }
}

/* sobel3d (Sobel3d.cpp:60-84): gx = Dx(Sz(Sy(T))), gy = Dy(Sz(Sx(T))), gz = Dz(Sy(Sx(T)))
   with S = [1/4, 1/2, 1/4] (edge taps 2/3, 2/3) and D = [-1/2, 0, 1/2] (edges -1/2,1/2 / 1/2,-1/2).
   n == 1 along x or y is rejected by the caller (the reference reads past the edge there). */
static void sobel3d(orc_warper* w)
{
   const size_t n = (size_t)w->n_px;
   const float a1[3] = {1.0f / 4.0f, 2.0f / 4.0f, 1.0f / 4.0f};
   const float a2[2] = {2.0f / 3.0f, 2.0f / 3.0f};
   const float s1[3] = {-0.5f, 0.0f, 0.5f};
   const float s2[2] = {-0.5f, 0.5f};
   const float s3[2] = {0.5f, -0.5f};
   double* m = (double*)malloc(n * sizeof(double));
   double* a = (double*)malloc(n * sizeof(double));
   double* b = (double*)malloc(n * sizeof(double));
   double* c = (double*)malloc(n * sizeof(double));
   for (size_t i = 0; i < n; i++) m[i] = w->reference[i];
   filt(w, m, a, AX, a1, a2, a2);
   filt(w, a, b, AY, a1, a2, a2);
   filt(w, a, c, AZ, a1, a2, a2);
   filt(w, b, w->gz, AZ, s1, s2, s3);
   filt(w, c, w->gy, AY, s1, s2, s3);
   filt(w, m, a, AY, a1, a2, a2);
   filt(w, a, b, AZ, a1, a2, a2);
   filt(w, b, w->gx, AX, s1, s2, s3);
   free(m);
   free(a);
   free(b);
   free(c);
}

/* computeSteepestDecentImages (FrameWarper.cpp:296-345): node i+1 gets Df*grad over interval
   i, node i gets (1-Df)*grad over interval i; both over [begin, end). The product is formed
   in double and stored as float (Point3f); z only in 3-D. */
static void steepest_descent(orc_warper* w)
{
   const int nD = w->nD;
   for (int i = 1; i < nD; i++) {
      int64_t p0 = w->D_range[2 * (i - 1)], p1 = w->D_range[2 * (i - 1) + 1];
      offvec_t* v = &w->vi[i];
      if (v->n > 0)
         for (int64_t p = p0; p < p1; p++) {
            if (p - v->i0 >= v->n) break;
            double jac = w->Df[p];
            double* dp = &v->v[3 * (p - v->i0)];
            dp[0] = rs(w, w->gx[p] * jac);
            dp[1] = rs(w, w->gy[p] * jac);
            if (w->n_dim == 3) dp[2] = rs(w, w->gz[p] * jac);
         }
   }
   for (int i = 0; i < nD - 1; i++) {
      int64_t p0 = w->D_range[2 * i], p1 = w->D_range[2 * i + 1];
      offvec_t* v = &w->vi[i];
      for (int64_t p = p0; p < p1; p++) {
         if (p - v->i0 < 0 || p - v->i0 >= v->n) break;
         double jac = 1 - w->Df[p];
         double* dp = &v->v[3 * (p - v->i0)];
         dp[0] = rs(w, w->gx[p] * jac);
         dp[1] = rs(w, w->gy[p] * jac);
         if (w->n_dim == 3) dp[2] = rs(w, w->gz[p] * jac);
      }
   }
}

static double vi_at(const offvec_t* v, int64_t p, int d)
{
   if (p < v->i0 || p > v->i1) return 0.0;
   return v->v[3 * (p - v->i0) + d];
}

/* computeHessianEntry / computeHessian (FrameWarper.cpp:347-395): H[(i,a),(j,b)] sums
   VI_i.a*VI_j.b (float products, double sum) over the overlap of the two ranges; zero when
   the nodes are more than one apart. */
static double hessian_entry(const orc_warper* w, int pi, int pj)
{
   int i = pi / w->n_dim, j = pj / w->n_dim;
   int di = pi % w->n_dim, dj = pj % w->n_dim;
   if (j > i) return hessian_entry(w, pj, pi);
   if ((i - j) > 1) return 0;
   const offvec_t* vi = &w->vi[i];
   const offvec_t* vj = &w->vi[j];
   int64_t p0 = vi->i0 > vj->i0 ? vi->i0 : vj->i0;
   int64_t p1 = vi->i1 < vj->i1 ? vi->i1 : vj->i1;
   double h = 0;
   for (int64_t p = p0; p < p1; p++) h += rs(w, vi_at(vi, p, di) * vi_at(vj, p, dj));
   return h;
}

static void compute_hessian(orc_warper* w)
{
   const int n = w->nD * w->n_dim, m = w->n_dim;
   for (int i = 0; i < n; i++)
      for (int j = 0; j <= i; j++) {
         double h = hessian_entry(w, i, j);
         w->H[(size_t)i * n + j] = h;
         w->H[(size_t)j * n + i] = h;
      }
   for (int k = 0; k < w->nD; k++)
      for (int a = 0; a < m; a++)
         for (int b = 0; b < m; b++) {
            w->Hdiag[(size_t)k * m * m + a * m + b] = w->H[(size_t)(k * m + a) * n + (k * m + b)];
            if (k + 1 < w->nD)
               w->Hsub[(size_t)k * m * m + a * m + b] = w->H[(size_t)((k + 1) * m + a) * n + (k * m + b)];
         }
}

/* AbstractFrameWarper::setReference (FrameWarper.cpp:186-205). */
int orc_set_reference(orc_warper* w, const float* ref, int nz, int ny, int nx, int nD,
                      double line, double interline, double interframe)
{
   free_ref(w);
   if (nz < 1 || ny < 2 || nx < 2) return fail("Reference must be three dimensional");
   if (nD < 2 || nD - 1 > 65535) return fail("Invalid number of nodes");
   if (nD <= nz) return fail("Invalid number of nodes");
   w->dims[AZ] = nz;
   w->dims[AY] = ny;
   w->dims[AX] = nx;
   w->nD = nD;
   w->n_dim = nz > 1 ? 3 : 2;
   w->n_px = (int64_t)nx * ny * nz;
   w->scan_in = scan_ctor(line, interline, interframe, nx, ny, nz);
   w->scan = process_scan(&w->scan_in, w->dims, nD);
   size_t n = (size_t)w->n_px;
   w->reference = (float*)malloc(n * sizeof(float));
   memcpy(w->reference, ref, n * sizeof(float));
   w->Di = (uint16_t*)malloc(n * sizeof(uint16_t));
   w->Df = (double*)malloc(n * sizeof(double));
   w->D_range = (int32_t*)calloc((size_t)nD * 2, sizeof(int32_t));
   w->gx = (double*)calloc(n, sizeof(double));
   w->gy = (double*)calloc(n, sizeof(double));
   w->gz = (double*)calloc(n, sizeof(double));
   w->vi = (offvec_t*)calloc((size_t)nD, sizeof(offvec_t));
   size_t nn = (size_t)nD * w->n_dim;
   w->H = (double*)calloc(nn * nn, sizeof(double));
   w->Hdiag = (double*)calloc((size_t)nD * w->n_dim * w->n_dim, sizeof(double));
   w->Hsub = (double*)calloc((size_t)nD * w->n_dim * w->n_dim, sizeof(double));
   if (precompute_interp(w) != 0) {
      free_ref(w);
      return -1;
   }
   sobel3d(w);
   steepest_descent(w);
   compute_hessian(w);
   return 0;
}

int orc_n_dim(const orc_warper* w) { return w->n_dim; }

void orc_get_scan(const orc_warper* w, double* o)
{
   o[0] = w->scan.line;
   o[1] = w->scan.interline;
   o[2] = w->scan.pixel;
   o[3] = w->scan.frame;
   o[4] = w->scan.interframe;
   o[5] = w->scan.stack;
}

void orc_get_node_map(const orc_warper* w, uint16_t* Di, double* Df, int32_t* D_range)
{
   if (Di) memcpy(Di, w->Di, (size_t)w->n_px * sizeof(uint16_t));
   if (Df) memcpy(Df, w->Df, (size_t)w->n_px * sizeof(double));
   if (D_range) memcpy(D_range, w->D_range, (size_t)w->nD * 2 * sizeof(int32_t));
}

void orc_get_gradient(const orc_warper* w, double* gx, double* gy, double* gz)
{
   size_t n = (size_t)w->n_px * sizeof(double);
   if (gx) memcpy(gx, w->gx, n);
   if (gy) memcpy(gy, w->gy, n);
   if (gz) memcpy(gz, w->gz, n);
}

void orc_get_vi_ranges(const orc_warper* w, int64_t* fl)
{
   for (int i = 0; i < w->nD; i++) {
      fl[2 * i] = w->vi[i].i0;
      fl[2 * i + 1] = w->vi[i].i1;
   }
}

void orc_get_hessian(const orc_warper* w, double* H)
{
   size_t n = (size_t)w->nD * w->n_dim;
   memcpy(H, w->H, n * n * sizeof(double));
}

/* AbstractFrameWarper::warpPoint (FrameWarper.cpp:397-419): node-interpolated displacement of
   pixel (x,y,z) in double from the float Df; zero outside the volume. */
void orc_warp_point(const orc_warper* w, const double* D, int x, int y, int z, int binning,
                    double* out)
{
   double factor = 1.0 / binning;
   int xs = (int)(x / factor);
   int ys = (int)(y / factor);
   out[0] = out[1] = out[2] = 0;
   if (xs < 0 || xs >= w->dims[AX] || ys < 0 || ys >= w->dims[AY] || z < 0 || z >= w->dims[AZ])
      return;
   int64_t loc = ((int64_t)z * w->dims[AY] + ys) * w->dims[AX] + xs;
   double f = w->Df[loc];
   int i = w->Di[loc];
   for (int c = 0; c < 3; c++) out[c] = (f * D[3 * (i + 1) + c] + (1 - f) * D[3 * i + c]) * factor;
}

/* getZeroPadded (FrameWarper.h:39-47): z is only range-checked when there is more than one
   plane. */
static inline double zpad(const orc_warper* w, const float* img, int z, int y, int x)
{
   int nz = w->dims[AZ], ny = w->dims[AY], nx = w->dims[AX];
   int ok = x >= 0 && x < nx && y >= 0 && y < ny && (nz == 1 || (z >= 0 && z < nz));
   return ok ? (double)img[((int64_t)z * ny + y) * nx + x] : 0.0;
}

/* Displacement of pixel idx in double (warpPoint with binning 1). In 2-D D.z is taken as 0,
   as col2D leaves it (OptimisationModel.cpp:29-42). */
static inline void disp(const orc_warper* w, const double* D, int64_t idx, double* d)
{
   double f = w->Df[idx];
   int i = w->Di[idx];
   d[0] = f * D[3 * (i + 1) + 0] + (1 - f) * D[3 * i + 0];
   d[1] = f * D[3 * (i + 1) + 1] + (1 - f) * D[3 * i + 1];
   d[2] = w->n_dim == 3 ? f * D[3 * (i + 1) + 2] + (1 - f) * D[3 * i + 2] : 0.0;
}

/* CpuFrameWarper::warpImageInterpolated (FrameWarper.cpp:31-69): trilinear (bilinear in 2-D)
   zero-padded pull-sample at p + d(p); float taps times double weights, stored as float; an
   exactly zero sample becomes invalid_value. */
void orc_warp_interpolated(const orc_warper* w, const float* img, const double* D,
                           double invalid_value, double* out)
{
   const int nz = w->dims[AZ], ny = w->dims[AY], nx = w->dims[AX];
   for (int z = 0; z < nz; z++)
      for (int y = 0; y < ny; y++)
         for (int x = 0; x < nx; x++) {
            int64_t idx = ((int64_t)z * ny + y) * nx + x;
            double d[3];
            disp(w, D, idx, d);
            double lx = d[0] + x, ly = d[1] + y, lz = d[2] + z;
            int x0 = (int)floor(lx), y0 = (int)floor(ly), z0 = (int)floor(lz);
            double fx = lx - x0, fy = ly - y0, fz = lz - z0;
            int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
            double s0 = zpad(w, img, z0, y0, x0) * (1 - fy) * (1 - fx) +
                        zpad(w, img, z0, y1, x0) * (fy) * (1 - fx) +
                        zpad(w, img, z0, y0, x1) * (1 - fy) * (fx) +
                        zpad(w, img, z0, y1, x1) * (fy) * (fx);
            double px = rs(w, (1 - fz) * s0);
            if (w->n_dim == 3) {
               double s1 = zpad(w, img, z1, y0, x0) * (1 - fy) * (1 - fx) +
                           zpad(w, img, z1, y1, x0) * (fy) * (1 - fx) +
                           zpad(w, img, z1, y0, x1) * (1 - fy) * (fx) +
                           zpad(w, img, z1, y1, x1) * (fy) * (fx);
               px = rs(w, px + fz * s1);
            }
            if (px == 0) px = invalid_value;
            out[idx] = px;
         }
}

/* CpuFrameWarper::computeErrorImage (FrameWarper.cpp:4-28): e = wimg - reference (float);
   pixels with a negative (invalid) warped value get e = 0 and wimg = 0; SSD sums e*e
   (float product) in double in scan order. */
static double error_image(const orc_warper* w, double* wimg, double* err)
{
   double ms = 0;
   for (int64_t p = 0; p < w->n_px; p++) {
      double e = rs(w, wimg[p] - (double)w->reference[p]);
      if (wimg[p] >= 0) {
         ms += rs(w, e * e);
      } else {
         e = 0;
         wimg[p] = 0;
      }
      err[p] = e;
   }
   return ms;
}

/* AbstractFrameWarper::computeJacobian (FrameWarper.cpp:147-167). */
static void jacobian(const orc_warper* w, const double* err, double* jac)
{
   const int m = w->n_dim;
   memset(jac, 0, sizeof(double) * (size_t)w->nD * m);
   for (int i = 0; i < w->nD; i++) {
      const offvec_t* v = &w->vi[i];
      for (int64_t p = v->i0; p < v->i1; p++) {
         const double* dp = &v->v[3 * (p - v->i0)];
         jac[i * m + 0] += rs(w, dp[0] * err[p]);
         jac[i * m + 1] += rs(w, dp[1] * err[p]);
         if (m == 3) jac[i * m + 2] += rs(w, dp[2] * err[p]);
      }
   }
}

/* CpuFrameWarper::getError (FrameWarper.cpp:169-176): warp with invalid = -1, then SSD. */
double orc_get_error(const orc_warper* w, const float* frame, const double* D, double* err_img,
                     double* wimg_out)
{
   size_t n = (size_t)w->n_px;
   double* wimg = (double*)malloc(n * sizeof(double));
   double* err = err_img ? err_img : (double*)malloc(n * sizeof(double));
   orc_warp_interpolated(w, frame, D, -1.0, wimg);
   double e = error_image(w, wimg, err);
   if (wimg_out) memcpy(wimg_out, wimg, n * sizeof(double));
   free(wimg);
   if (!err_img) free(err);
   return e;
}

double orc_get_jacobian(const orc_warper* w, const float* frame, const double* D, double* jac)
{
   double* err = (double*)malloc((size_t)w->n_px * sizeof(double));
   double e = orc_get_error(w, frame, D, err, NULL);
   jacobian(w, err, jac);
   free(err);
   return e;
}

/* AbstractFrameWarper::warpImage (FrameWarper.cpp:71-136): forward splat of each source pixel
   to p - d(p) with trilinear weights into wimg and coverage. The clamp of loc at :87-94 runs
   after loc0/locf are taken and never feeds back, so it has no effect and is not restated.
   Accumulation is float += double in scan order. */
void orc_warp_image(const orc_warper* w, const float* img, const double* D, double* wimg,
                    double* cov)
{
   const int nz = w->dims[AZ], ny = w->dims[AY], nx = w->dims[AX];
   memset(wimg, 0, sizeof(double) * (size_t)w->n_px);
   memset(cov, 0, sizeof(double) * (size_t)w->n_px);
   for (int z = 0; z < nz; z++)
      for (int y = 0; y < ny; y++)
         for (int x = 0; x < nx; x++) {
            int64_t idx = ((int64_t)z * ny + y) * nx + x;
            double d[3];
            disp(w, D, idx, d);
            double lx = x - d[0], ly = y - d[1], lz = z - d[2];
            int x0 = (int)floor(lx), y0 = (int)floor(ly), z0 = (int)floor(lz);
            double fx = lx - x0, fy = ly - y0, fz = lz - z0;
            int x1 = x0, y1 = y0, z1 = z0;
            if (x1 < nx - 1) x1++;
            if (y1 < ny - 1) y1++;
            if (z1 < nz - 1) z1++;
            int valid = x0 >= 0 && y0 >= 0 && x0 <= nx - 1 && y0 <= ny - 1;
            if (nz > 1) valid = valid && z0 >= 0 && z0 <= nz - 1;
            if (!valid) continue;
            double v = img[idx];
            const int zz[8] = {z0, z0, z0, z0, z1, z1, z1, z1};
            const int yy[8] = {y0, y1, y0, y1, y0, y1, y0, y1};
            const int xx[8] = {x0, x0, x1, x1, x0, x0, x1, x1};
            const double wy[8] = {1 - fy, fy, 1 - fy, fy, 1 - fy, fy, 1 - fy, fy};
            const double wx[8] = {1 - fx, 1 - fx, fx, fx, 1 - fx, 1 - fx, fx, fx};
            const double wz[8] = {1 - fz, 1 - fz, 1 - fz, 1 - fz, fz, fz, fz, fz};
            for (int t = 0; t < 8; t++) {
               int64_t o = ((int64_t)zz[t] * ny + yy[t]) * nx + xx[t];
               wimg[o] = rs(w, wimg[o] + wy[t] * wx[t] * wz[t] * v);
            }
            for (int t = 0; t < 8; t++) {
               int64_t o = ((int64_t)zz[t] * ny + yy[t]) * nx + xx[t];
               cov[o] = rs(w, cov[o] + wy[t] * wx[t] * wz[t]);
            }
         }
}

/* ---------------------------------------------------------------------------------------
   Linear algebra for the trust region: block-tridiagonal Cholesky of H + lambda*I. */

static int chol_small(int m, double* A)
{
   /* In-place lower Cholesky of an m x m row-major block; 0 on failure. */
   for (int j = 0; j < m; j++) {
      double s = A[j * m + j];
      for (int k = 0; k < j; k++) s -= A[j * m + k] * A[j * m + k];
      if (!(s > 0) || !isfinite(s)) return 0;
      double l = sqrt(s);
      double il = 1.0 / l;
      A[j * m + j] = l;
      for (int i = j + 1; i < m; i++) {
         double t = A[i * m + j];
         for (int k = 0; k < j; k++) t -= A[i * m + k] * A[j * m + k];
         A[i * m + j] = t * il;
      }
      for (int i = 0; i < j; i++) A[i * m + j] = 0;
   }
   return 1;
}

int orc_bt_cholesky(int nb, int m, const double* diag, const double* sub, double lambda,
                    double* Ld, double* Ls)
{
   const int mm = m * m;
   for (int k = 0; k < nb; k++) {
      double A[9];
      for (int i = 0; i < mm; i++) A[i] = diag[k * mm + i];
      for (int i = 0; i < m; i++) A[i * m + i] += lambda;
      if (k > 0) {
         const double* C = &Ls[(k - 1) * mm];
         for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) {
               double s = 0;
               for (int t = 0; t < m; t++) s += C[i * m + t] * C[j * m + t];
               A[i * m + j] -= s;
            }
      }
      if (!chol_small(m, A)) return 0;
      for (int i = 0; i < mm; i++) Ld[k * mm + i] = A[i];
      if (k + 1 < nb) {
         /* L(k+1,k) = B(k+1,k) * L(k,k)^-T: solve X * L^T = B row by row. */
         const double* B = &sub[k * mm];
         double* X = &Ls[k * mm];
         for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) {
               double t = B[i * m + j];
               for (int q = 0; q < j; q++) t -= X[i * m + q] * A[j * m + q];
               X[i * m + j] = t * (1.0 / A[j * m + j]);
            }
      }
   }
   return 1;
}

static void bt_forward(int nb, int m, const double* Ld, const double* Ls, const double* b, double* y)
{
   const int mm = m * m;
   for (int k = 0; k < nb; k++) {
      double r[3] = {0, 0, 0};
      for (int i = 0; i < m; i++) {
         r[i] = b[k * m + i];
         if (k > 0)
            for (int t = 0; t < m; t++) r[i] -= Ls[(k - 1) * mm + i * m + t] * y[(k - 1) * m + t];
      }
      const double* L = &Ld[k * mm];
      for (int i = 0; i < m; i++) {
         double t = r[i];
         for (int q = 0; q < i; q++) t -= L[i * m + q] * y[k * m + q];
         y[k * m + i] = t * (1.0 / L[i * m + i]);
      }
   }
}

static void bt_backward(int nb, int m, const double* Ld, const double* Ls, const double* y, double* x)
{
   const int mm = m * m;
   for (int k = nb - 1; k >= 0; k--) {
      double r[3] = {0, 0, 0};
      for (int i = 0; i < m; i++) {
         r[i] = y[k * m + i];
         if (k + 1 < nb)
            for (int t = 0; t < m; t++) r[i] -= Ls[k * mm + t * m + i] * x[(k + 1) * m + t];
      }
      const double* L = &Ld[k * mm];
      for (int i = m - 1; i >= 0; i--) {
         double t = r[i];
         for (int q = i + 1; q < m; q++) t -= L[q * m + i] * x[k * m + q];
         x[k * m + i] = t * (1.0 / L[i * m + i]);
      }
   }
}

void orc_bt_solve(int nb, int m, const double* Ld, const double* Ls, const double* b, double* x)
{
   double* y = (double*)malloc(sizeof(double) * (size_t)nb * m);
   bt_forward(nb, m, Ld, Ls, b, y);
   bt_backward(nb, m, Ld, Ls, y, x);
   free(y);
}

/* Matrix access used by the subproblem, for either storage. */
typedef struct {
   int n, nb, m;
   const double* dense;
   const double *diag, *sub;
   double *Ld, *Ls, *Ldense;
} mat_t;

static double mat_get(const mat_t* A, int i, int j)
{
   if (A->dense) return A->dense[(size_t)i * A->n + j];
   int bi = i / A->m, bj = j / A->m, a = i % A->m, b = j % A->m;
   if (bi == bj) return A->diag[(size_t)bi * A->m * A->m + a * A->m + b];
   if (bi == bj + 1) return A->sub[(size_t)bj * A->m * A->m + a * A->m + b];
   if (bj == bi + 1) return A->sub[(size_t)bi * A->m * A->m + b * A->m + a];
   return 0;
}

static int mat_chol(mat_t* A, double lambda)
{
   if (!A->dense) return orc_bt_cholesky(A->nb, A->m, A->diag, A->sub, lambda, A->Ld, A->Ls);
   int n = A->n;
   double* L = A->Ldense;
   for (int i = 0; i < n * n; i++) L[i] = A->dense[i];
   for (int i = 0; i < n; i++) L[(size_t)i * n + i] += lambda;
   for (int j = 0; j < n; j++) {
      double s = L[(size_t)j * n + j];
      for (int k = 0; k < j; k++) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
      if (!(s > 0) || !isfinite(s)) return 0;
      double l = sqrt(s);
      L[(size_t)j * n + j] = l;
      for (int i = j + 1; i < n; i++) {
         double t = L[(size_t)i * n + j];
         for (int k = 0; k < j; k++) t -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
         L[(size_t)i * n + j] = t / l;
      }
   }
   return 1;
}

/* y = L^-1 b (forward) and x = L^-T y (backward) for the last factorisation. */
static void mat_fwd(const mat_t* A, const double* b, double* y)
{
   if (!A->dense) {
      bt_forward(A->nb, A->m, A->Ld, A->Ls, b, y);
      return;
   }
   int n = A->n;
   for (int i = 0; i < n; i++) {
      double t = b[i];
      for (int k = 0; k < i; k++) t -= A->Ldense[(size_t)i * n + k] * y[k];
      y[i] = t / A->Ldense[(size_t)i * n + i];
   }
}

static void mat_bwd(const mat_t* A, const double* y, double* x)
{
   if (!A->dense) {
      bt_backward(A->nb, A->m, A->Ld, A->Ls, y, x);
      return;
   }
   int n = A->n;
   for (int i = n - 1; i >= 0; i--) {
      double t = y[i];
      for (int k = i + 1; k < n; k++) t -= A->Ldense[(size_t)k * n + i] * x[k];
      x[i] = t / A->Ldense[(size_t)i * n + i];
   }
}

static double norm2(int n, const double* v)
{
   double s = 0;
   for (int i = 0; i < n; i++) s += v[i] * v[i];
   return sqrt(s);
}

/* Trust-region subproblem min g'p + p'Bp/2 s.t. |p| <= radius: Nocedal & Wright Alg. 4.3
   (More & Sorensen 1983), in the shape dlib's solve_trust_region_subproblem gives it:
   lambda starts at 0; the unconstrained step is taken when B is positive definite and the
   step lies inside the region; otherwise Newton's method on 1/|p(lambda)| - 1/radius with a
   [lambda_lo, lambda_hi] bracket (bisection when Newton leaves it or the factorisation fails)
   until ||p| - radius| <= eps*radius. Returns the number of iterations used. */
static int tr_subproblem(mat_t* A, const double* g, double radius, double eps, int max_iter,
                         double* p, double* lambda_out)
{
   const int n = A->n;
   double gnorm = norm2(n, g);
   double* y = (double*)malloc(sizeof(double) * (size_t)n);
   double* tmp = (double*)malloc(sizeof(double) * (size_t)n);
   memset(p, 0, sizeof(double) * (size_t)n);
   *lambda_out = 0;
   if (gnorm == 0) {
      free(y);
      free(tmp);
      return 1;
   }
   /* Gershgorin lower bound on the smallest eigenvalue gives a safe upper bracket. */
   double eig_lo = INFINITY;
   for (int i = 0; i < n; i++) {
      double c = mat_get(A, i, i), r = 0;
      int j0 = A->dense ? 0 : ((i / A->m) - 1) * A->m;
      int j1 = A->dense ? n : ((i / A->m) + 2) * A->m;
      if (j0 < 0) j0 = 0;
      if (j1 > n) j1 = n;
      for (int j = j0; j < j1; j++)
         if (j != i) r += fabs(mat_get(A, i, j));
      if (c - r < eig_lo) eig_lo = c - r;
   }
   double lam = 0, lam_lo = 0, lam_hi = gnorm / radius + (eig_lo < 0 ? -eig_lo : 0);
   int it;
   for (it = 0; it < max_iter; it++) {
      if (!mat_chol(A, lam)) {
         lam_lo = lam;
         lam = 0.5 * (lam_lo + lam_hi);
         continue;
      }
      for (int i = 0; i < n; i++) tmp[i] = -g[i];
      mat_fwd(A, tmp, y);
      mat_bwd(A, y, p);
      *lambda_out = lam;
      double pn = norm2(n, p);
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
      mat_fwd(A, p, y); /* q = L^-1 p */
      double qn = norm2(n, y);
      double lam_new = lam + (pn / qn) * (pn / qn) * (pn - radius) / radius;
      if (!(lam_new > lam_lo && lam_new < lam_hi)) lam_new = 0.5 * (lam_lo + lam_hi);
      lam = lam_new;
   }
   free(y);
   free(tmp);
   return it;
}

int orc_tr_subproblem_bt(int nb, int m, const double* diag, const double* sub, const double* g,
                         double radius, double eps, int max_iter, double* p, double* lambda_out)
{
   mat_t A = {0};
   A.n = nb * m;
   A.nb = nb;
   A.m = m;
   A.diag = diag;
   A.sub = sub;
   A.Ld = (double*)malloc(sizeof(double) * (size_t)nb * m * m);
   A.Ls = (double*)malloc(sizeof(double) * (size_t)nb * m * m);
   int it = tr_subproblem(&A, g, radius, eps, max_iter, p, lambda_out);
   free(A.Ld);
   free(A.Ls);
   return it;
}

int orc_tr_subproblem_dense(int n, const double* B, const double* g, double radius, double eps,
                            int max_iter, double* p, double* lambda_out)
{
   mat_t A = {0};
   A.n = n;
   A.m = 1;
   A.nb = n;
   A.dense = B;
   A.Ldense = (double*)malloc(sizeof(double) * (size_t)n * n);
   int it = tr_subproblem(&A, g, radius, eps, max_iter, p, lambda_out);
   free(A.Ldense);
   return it;
}

/* p'Bp for a block-tridiagonal B (nb diagonal blocks of m x m, sub = blocks (k+1, k)). */
static double quad_bt(int nb, int m, const double* diag, const double* sub, const double* p)
{
   const int mm = m * m;
   double s = 0;
   for (int k = 0; k < nb; k++) {
      for (int a = 0; a < m; a++)
         for (int b = 0; b < m; b++) s += p[k * m + a] * diag[k * mm + a * m + b] * p[k * m + b];
      if (k + 1 < nb)
         for (int a = 0; a < m; a++)
            for (int b = 0; b < m; b++) s += 2.0 * p[(k + 1) * m + a] * sub[k * mm + a * m + b] * p[k * m + b];
   }
   return s;
}

void orc_tr_default(orc_tr_params* p)
{
   p->radius = 40;
   p->min_delta = 1e-8;
   p->rel_delta = 0;
   p->max_iter = 0;
   p->max_evals = 0;
   p->sub_eps = 0.1;
   p->sub_max_iter = 20;
   p->skip_repeat = 0;
   p->dense = 0;
}

/* One trace record per objective evaluation (layout in fa_oracle.h), then x, g, p. */
static void put_rec(double* trace, int cap, int n, int r, const double* v, const double* x,
                    const double* g, const double* p)
{
   if (!trace || r < 0 || r >= cap) return;
   double* o = trace + (size_t)r * (ORC_TRACE_FIELDS + 3 * (size_t)n);
   memcpy(o, v, sizeof(double) * ORC_TRACE_FIELDS);
   memcpy(o + ORC_TRACE_FIELDS, x, sizeof(double) * (size_t)n);
   memcpy(o + ORC_TRACE_FIELDS + n, g, sizeof(double) * (size_t)n);
   memcpy(o + ORC_TRACE_FIELDS + 2 * (size_t)n, p, sizeof(double) * (size_t)n);
}

/* dlib find_min_trust_region with objective_delta_stop_strategy, as called at
   FrameWarpAligner.cpp:205-209 with the OptimisationModel concept (OptimisationModel.h:6-27):
   a model gives f(x) and the derivative at the last evaluated point, and a constant B.
   Restated from dlib's published algorithm (Nocedal & Wright Alg. 4.1; dlib is not part of
   the reference tree):
     - the stop strategy runs at the top of every iteration whose x is new: iteration count
       +1; from the second call on, stop when the count exceeds max_iter (if set) or
       |f - f_prev| < min_delta; then f_prev = f;
     - p from the subproblem at the current radius, f_new = f(x + p);
     - stop when |pred| <= |meas| * DBL_EPSILON (no progress) or rho = meas / |pred| is not
       finite;
     - rho < 0.25: radius /= 4, and stop without moving x once the radius has collapsed;
       otherwise rho > 0.75 with a subproblem that iterated more than once (the step was
       held by the radius): radius = min(2 radius, 1000);
     - rho > 0: x += p, new f and g; else x is stale (no stop check next time).
   Deliberate differences, none of which changes the returned x beyond 1e-13 of the radius:
   the radius counts as collapsed at radius0 * DBL_EPSILON (dlib: DBL_MIN, which costs up to
   ~500 more rejected evaluations before the same stop), and the relative stop, evaluation
   cap and skip_repeat are additions for the benchmark configurations. */
int orc_tr_run(const orc_tr_model* M, const orc_tr_params* prm, double* x, orc_tr_stats* st,
               double* trace, int cap)
{
   const int n = M->nb * M->m;
   orc_tr_stats s = {0};
   double* g = (double*)malloc(sizeof(double) * (size_t)n);
   double* p = (double*)calloc((size_t)n, sizeof(double));
   double* xt = (double*)malloc(sizeof(double) * (size_t)n);
   mat_t A = {0};
   A.n = n;
   A.nb = M->nb;
   A.m = M->m;
   A.diag = M->diag;
   A.sub = M->sub;
   A.Ld = (double*)malloc(sizeof(double) * (size_t)n * M->m);
   A.Ls = (double*)malloc(sizeof(double) * (size_t)n * M->m);
   double* Bd = NULL; /* dense B (prm->dense): the n x n matrix dlib receives */
   if (prm->dense) {
      const int m = M->m, mm = m * m;
      Bd = (double*)calloc((size_t)n * n, sizeof(double));
      A.Ldense = (double*)malloc(sizeof(double) * (size_t)n * n);
      for (int k = 0; k < M->nb; k++)
         for (int a = 0; a < m; a++)
            for (int b = 0; b < m; b++) {
               Bd[(size_t)(k * m + a) * n + k * m + b] = M->diag[k * mm + a * m + b];
               if (k + 1 < M->nb) {
                  const double v = M->sub[k * mm + a * m + b];
                  Bd[(size_t)((k + 1) * m + a) * n + k * m + b] = v;
                  Bd[(size_t)(k * m + b) * n + (k + 1) * m + a] = v;
               }
            }
      A.dense = Bd;
   }

   double f = M->value(M->user, x);
   s.evaluations = 1;
   M->gradient(M->user, g);
   double radius = prm->radius;
   const double r0 = radius;
   int stale = 0, been_used = 0;
   double prev_f = 0;
   /* the evaluation the next trace record describes */
   double rv[ORC_TRACE_FIELDS] = {0};
   rv[ORC_T_FNEW] = f;
   rv[ORC_T_ACCEPT] = 1;
   for (;;) {
      int status = 0;
      if (!stale) {
         s.iterations++;
         if (been_used) {
            double df = fabs(f - prev_f);
            if (prm->max_iter && s.iterations > prm->max_iter) status = ORC_STOP_MAX_ITER;
            else if (df < prm->min_delta) status = ORC_STOP_ABS_DELTA;
            else if (prm->rel_delta > 0 && df < prm->rel_delta * fabs(prev_f)) status = ORC_STOP_REL_DELTA;
         }
         been_used = 1;
         prev_f = f;
      }
      if (!status && prm->max_evals && s.evaluations >= prm->max_evals) status = ORC_STOP_MAX_EVALS;
      double lam = 0;
      int it = 0;
      if (!status) it = tr_subproblem(&A, g, radius, prm->sub_eps, prm->sub_max_iter, p, &lam);
      rv[ORC_T_EVALS] = s.evaluations;
      rv[ORC_T_F] = f;
      rv[ORC_T_RADIUS] = radius;
      rv[ORC_T_ITER] = s.iterations;
      rv[ORC_T_STATUS] = status;
      rv[ORC_T_LAMBDA] = lam;
      rv[ORC_T_SUBIT] = it;
      rv[ORC_T_PNORM] = status ? 0 : norm2(n, p);
      rv[ORC_T_GNORM] = norm2(n, g);
      double gp = 0;
      for (int i = 0; i < n; i++) gp += g[i] * p[i];
      rv[ORC_T_GP] = status ? 0 : gp;
      rv[ORC_T_REJ] = s.rejected;
      put_rec(trace, cap, n, s.evaluations - 1, rv, x, g, p);
      if (status) {
         s.status = status;
         break;
      }
      for (int i = 0; i < n; i++) xt[i] = x[i] + p[i];
      double fn = M->value(M->user, xt);
      s.evaluations++;
      double pBp = 0;
      if (Bd) {
         for (int i = 0; i < n; i++) {
            double t = 0;
            for (int j = 0; j < n; j++) t += Bd[(size_t)i * n + j] * p[j];
            pBp += p[i] * t;
         }
      } else {
         pBp = quad_bt(M->nb, M->m, M->diag, M->sub, p);
      }
      double pred = -(0.5 * pBp + gp);
      double meas = f - fn;
      memset(rv, 0, sizeof rv);
      rv[ORC_T_FNEW] = fn;
      rv[ORC_T_PRED] = pred;
      rv[ORC_T_MEAS] = meas;
      if (fabs(pred) <= fabs(meas) * DBL_EPSILON) status = ORC_STOP_NO_PROGRESS;
      double rho = status ? 0 : meas / fabs(pred);
      if (!status && !isfinite(rho)) status = ORC_STOP_NONFINITE;
      if (!status) {
         rv[ORC_T_RHO] = rho;
         if (rho < 0.25) {
            radius *= 0.25;
            if (radius <= r0 * DBL_EPSILON) status = ORC_STOP_RADIUS; /* x stays */
         } else if (rho > 0.75 && it > 1) {
            radius = 2 * radius < 1000 ? 2 * radius : 1000;
         }
      }
      if (!status) {
         if (rho > 0) {
            memcpy(x, xt, sizeof(double) * (size_t)n);
            f = fn;
            M->gradient(M->user, g);
            stale = 0;
            rv[ORC_T_ACCEPT] = 1;
         } else {
            stale = 1;
            s.rejected++;
            if (prm->skip_repeat && lam == 0) {
               /* The unchanged interior step would be proposed, evaluated and rejected again
                  while it fits the shrunken region. */
               double pn = norm2(n, p);
               while (pn <= radius) {
                  radius *= 0.25;
                  s.rejected++;
                  if (radius <= r0 * DBL_EPSILON) {
                     status = ORC_STOP_RADIUS;
                     break;
                  }
               }
            }
         }
      }
      if (status) {
         rv[ORC_T_EVALS] = s.evaluations;
         rv[ORC_T_F] = f;
         rv[ORC_T_RADIUS] = radius;
         rv[ORC_T_ITER] = s.iterations;
         rv[ORC_T_STATUS] = status;
         rv[ORC_T_GNORM] = norm2(n, g);
         rv[ORC_T_REJ] = s.rejected;
         put_rec(trace, cap, n, s.evaluations - 1, rv, x, g, p);
         s.status = status;
         break;
      }
   }
   s.f_final = f;
   s.radius_final = radius;
   if (st) *st = s;
   free(g);
   free(p);
   free(xt);
   free(A.Ld);
   free(A.Ls);
   free(Bd);
   free(A.Ldense);
   return 0;
}

/* col2D (OptimisationModel.cpp:29-42): x = [x0, y0, (z0), x1, ...] -> D (nD x 3). */
static void col2D(const orc_warper* w, const double* x, double* D)
{
   for (int i = 0; i < w->nD; i++) {
      D[3 * i + 0] = x[i * w->n_dim + 0];
      D[3 * i + 1] = x[i * w->n_dim + 1];
      D[3 * i + 2] = w->n_dim == 3 ? x[i * w->n_dim + 2] : 0.0;
   }
}

/* OptimisationModel over one frame (OptimisationModel.cpp:45-63): operator() = getError at
   col2D(x), keeping the error image; get_derivative_and_hessian = getJacobian from that error
   image (the last getError on the frame) and the warper's H. */
typedef struct {
   const orc_warper* w;
   const float* frame;
   double* D;
   double* err;
} warp_model_t;

static double warp_model_value(void* u, const double* x)
{
   warp_model_t* m = (warp_model_t*)u;
   col2D(m->w, x, m->D);
   return orc_get_error(m->w, m->frame, m->D, m->err, NULL);
}

static void warp_model_gradient(void* u, double* g)
{
   warp_model_t* m = (warp_model_t*)u;
   jacobian(m->w, m->err, g);
}

int orc_align_trace(const orc_warper* w, const float* frame, const orc_tr_params* prm, double* x,
                    orc_tr_stats* st, double* trace, int cap)
{
   warp_model_t wm;
   wm.w = w;
   wm.frame = frame;
   wm.D = (double*)malloc(sizeof(double) * 3 * (size_t)w->nD);
   wm.err = (double*)malloc(sizeof(double) * (size_t)w->n_px);
   orc_tr_model M;
   M.nb = w->nD;
   M.m = w->n_dim;
   M.diag = w->Hdiag;
   M.sub = w->Hsub;
   M.value = warp_model_value;
   M.gradient = warp_model_gradient;
   M.user = &wm;
   int rc = orc_tr_run(&M, prm, x, st, trace, cap);
   free(wm.D);
   free(wm.err);
   return rc;
}

int orc_align(const orc_warper* w, const float* frame, const orc_tr_params* prm, double* x,
              orc_tr_stats* st)
{
   return orc_align_trace(w, frame, prm, x, st, NULL, 0);
}

typedef struct {
   const orc_warper* w;
   const float* frames;
   int n_frames;
   const orc_tr_params* prm;
   double* x;
   orc_tr_stats* st;
   double* secs;
   int next;
   pthread_mutex_t mu;
} pool_t;

static double now_s(void)
{
   struct timespec ts;
   clock_gettime(CLOCK_MONOTONIC, &ts);
   return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void* worker(void* arg)
{
   pool_t* P = (pool_t*)arg;
   const int n = P->w->nD * P->w->n_dim;
   for (;;) {
      pthread_mutex_lock(&P->mu);
      int i = P->next++;
      pthread_mutex_unlock(&P->mu);
      if (i >= P->n_frames) break;
      double t0 = now_s();
      orc_align(P->w, P->frames + (size_t)i * P->w->n_px, P->prm, P->x + (size_t)i * n,
                P->st ? &P->st[i] : NULL);
      if (P->secs) P->secs[i] = now_s() - t0;
   }
   return NULL;
}

/* One frame per task over a fixed pool, as the MEX front end spreads addFrame tasks over a
   hardware_concurrency() pool (FrameAlignerMex/FrameAlignerMex.cpp:24-25,120). */
int orc_align_many(const orc_warper* w, const float* frames, int n_frames, const orc_tr_params* prm,
                   double* x, orc_tr_stats* st, double* secs, int n_threads)
{
   pool_t P;
   P.w = w;
   P.frames = frames;
   P.n_frames = n_frames;
   P.prm = prm;
   P.x = x;
   P.st = st;
   P.secs = secs;
   P.next = 0;
   pthread_mutex_init(&P.mu, NULL);
   if (n_threads < 1) n_threads = 1;
   pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)n_threads);
   for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, worker, &P);
   for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
   free(th);
   pthread_mutex_destroy(&P.mu);
   return 0;
}
