"""CPU restatement of FrameWarpAligner::addFrame's front end (numpy, float64 FFTs).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device front end.

* downsample            - AbstractFrameAligner.cpp:37-56 (x, y summed in bins, z kept)
* hann / window         - VolumePhaseCorrelator.cpp:134-158 (square root of the Hann window in
                          y and x, 1 in z)
* PhaseCorrelator       - VolumePhaseCorrelator.cpp:10-131: reference spectrum conj(FFT(ref * w)),
                          cross-power spectrum normalised by |.| + DBL_MIN, inverse FFT, first
                          maximum, indices above half the size wrapped negative; the sub-pixel
                          centroid is computed by the reference and discarded, so the shift is
                          integer
* gaussian_kernel       - OpenCV's getGaussianKernel for cv::GaussianBlur(.., Size(0, 0), sx, sy)
                          on float images: ksize = round(8 sigma + 1) | 1, weights
                          exp(-(i - (ksize-1)/2)^2 / (2 sigma^2)) normalised to sum 1 (float
                          weights). OpenCV is not in this image and its version is not pinned:
                          the blur is parity unpinned against OpenCV itself at the level of
                          rounding (its SIMD row/column filters may sum in another order).
* smooth_stack          - FrameWarpAligner.cpp:46-64: per z-slice GaussianBlur with sigma x =
                          smoothing and sigma y = 1 (the reference's second argument), border
                          BORDER_REFLECT_101, row pass then column pass in float32
* rigid start, starts   - FrameWarpAligner.cpp:150-198: starts {0, 0, rigid}, the first with the
                          smallest SSD
"""
import numpy as np

DBL_MIN = np.finfo(np.float64).tiny


def downsample(vol, factor):
    """im2[z, y // f, x // f] += im1[z, y, x] in float32, y outer and x inner."""
    vol = np.asarray(vol, np.float32)
    nz, h, w = vol.shape
    if factor == 1:
        return vol.copy()
    nh, nw = -(-h // factor), -(-w // factor)
    out = np.zeros((nz, nh, nw), np.float32)
    for y in range(h):
        for x in range(w):
            out[:, y // factor, x // factor] += vol[:, y, x]
    return out


def hann(n):
    if n <= 1:
        return np.ones(max(n, 1))
    f = 2.0 * np.pi / (n - 1)
    i = np.arange(n, dtype=np.float64)
    return np.sqrt(0.5 * (1.0 - np.cos(f * i)))


def window(dims):
    nz, ny, nx = dims
    return np.ones(nz)[:, None, None] * hann(ny)[None, :, None] * hann(nx)[None, None, :]


class PhaseCorrelator:
    def __init__(self, dims):
        self.dims = tuple(int(d) for d in dims)
        self.w = window(self.dims)
        self.ref = None

    def _spec(self, vol):
        v = np.asarray(vol, np.float32).astype(np.float64) * self.w
        return np.fft.rfftn(v, axes=(0, 1, 2))

    def set_reference(self, vol):
        self.ref = np.conj(self._spec(vol))
        return self

    def correlation(self, vol):
        p = self._spec(vol) * self.ref
        p = p / (np.abs(p) + DBL_MIN)
        return np.fft.irfftn(p, s=self.dims, axes=(0, 1, 2)) * np.prod(self.dims)  # unnormalised

    def compute_shift(self, vol):
        c = self.correlation(vol)
        idx = int(np.argmax(c.ravel()))  # first maximum
        nz, ny, nx = self.dims
        x, y, z = idx % nx, (idx // nx) % ny, idx // (nx * ny)
        if x > nx // 2:
            x -= nx
        if y > ny // 2:
            y -= ny
        if z > nz // 2:
            z -= nz
        return float(x), float(y), float(z)


def rigid_shift(correlator, raw_frame, binning):
    """(x, y, z) of the rigid start: phase correlation of the binned raw frame, x and y scaled
    by the binning (FrameWarpAligner.cpp:186-189)."""
    x, y, z = correlator.compute_shift(downsample(raw_frame, binning))
    return x * binning, y * binning, z


def gaussian_kernel(sigma):
    """Float weights: each exp() rounded to float, normalised by 1 / (their double sum) and
    rounded again, as getGaussianKernel does for a float kernel."""
    ksize = int(np.rint(sigma * 8 + 1)) | 1
    c = (ksize - 1) * 0.5
    i = np.arange(ksize, dtype=np.float64)
    t = np.exp(-0.5 / (sigma * sigma) * (i - c) ** 2).astype(np.float32)
    inv = 1.0 / float(np.sum(t.astype(np.float64)))
    return (t.astype(np.float64) * inv).astype(np.float32)


def _reflect101(i, n):
    if n == 1:
        return 0
    period = 2 * n - 2
    i = abs(i) % period
    return period - i if i >= n else i


def _filter_axis(img, k, axis):
    """Symmetric filter along axis in float32: k[r] x[i] + sum_{j=1..r} k[r+j] (x[i-j] + x[i+j]),
    j ascending, rows outside the image reflected (BORDER_REFLECT_101)."""
    img = np.asarray(img, np.float32)
    n = img.shape[axis]
    r = len(k) // 2
    take = lambda d: np.take(img, np.array([_reflect101(i + d, n) for i in range(n)]), axis=axis)
    out = np.float32(k[r]) * img
    for j in range(1, r + 1):
        out = (out + np.float32(k[r + j]) * (take(-j) + take(j))).astype(np.float32)
    return out.astype(np.float32)


def smooth_stack(vol, smoothing):
    """Per-slice blur (sigma x = smoothing, sigma y = 1) of a (nz, ny, nx) float volume; the
    volume converted to float when smoothing is 0."""
    vol = np.asarray(vol, np.float32)
    if not smoothing > 0:
        return vol.copy()
    kx, ky = gaussian_kernel(smoothing), gaussian_kernel(1.0)
    out = np.empty_like(vol)
    for z in range(vol.shape[0]):
        out[z] = _filter_axis(_filter_axis(vol[z], kx, 1), ky, 0)
    return out


def choose_start(errors):
    """Index of the start addFrame takes: the first smallest SSD (std::min_element)."""
    return int(np.argmin(np.asarray(errors)))


def to_u16(v):
    """cv::Mat::convertTo(CV_16U) of float values: round half to even, saturate to [0, 65535];
    NaN, infinities and values beyond the int range give 0 (x86's out-of-range conversion)."""
    v = np.asarray(v, np.float32)
    out = np.zeros(v.shape, np.uint16)
    ok = np.abs(v) < np.float32(2147483648.0)
    out[ok] = np.clip(np.rint(v[ok]), 0, 65535).astype(np.uint16)
    return out


def to_u8(v):
    """cv::Mat::convertTo(CV_8U) of float values: the same rounding, saturated to [0, 255]."""
    v = np.asarray(v, np.float32)
    out = np.zeros(v.shape, np.uint8)
    ok = np.abs(v) < np.float32(2147483648.0)
    out[ok] = np.clip(np.rint(v[ok]), 0, 255).astype(np.uint8)
    return out


def realign_as_frame(orc, frame, D, interpolate_missing=False, cv_type=0):
    """FrameWarpAligner::realignAsFrame (FrameWarpAligner.cpp:270-285) with the oracle's warper:
    the interpolated warp (FrameWarper.cpp:31-69, invalid 0, float32) or warpImageNormalised
    (FrameWarper.cpp:138-144: the splat's weighted values, not divided by the coverage,
    converted to cv_type 0 = CV_8U, 2 = CV_16U, 5 = CV_32F)."""
    f = np.asarray(frame, np.float32)
    if interpolate_missing:
        return orc.warp_interpolated(f, D, 0.0).astype(np.float32)
    wimg, _ = orc.warp_image(f, D)
    wimg = np.asarray(wimg, np.float32)
    return {0: to_u8, 2: to_u16, 5: lambda v: v}[cv_type](wimg)


def preserving(wimg, cov):
    """getWarpedFrame (FrameWarpAligner.cpp:317-325): intensity_preserving /= mask, to uint16."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return to_u16(np.asarray(wimg, np.float32) / np.asarray(cov, np.float32))


def correlation(a, b, mask):
    """correlation() of Cv3dUtils.cpp:64-88: mean and standard deviation (population, double
    sums) of each image over the mask, then the mean over the mask of the float product of the
    float-centred images, over the product of the deviations."""
    a = np.asarray(a, np.float32).ravel()
    b = np.asarray(b, np.float32).ravel()
    m = np.asarray(mask).ravel() != 0
    n = int(m.sum())
    if n == 0:
        return float("nan")
    av, bv = a[m].astype(np.float64), b[m].astype(np.float64)
    ma, mb = av.sum() / n, bv.sum() / n
    sa = np.sqrt(max((av * av).sum() / n - ma * ma, 0.0))
    sb = np.sqrt(max((bv * bv).sum() / n - mb * mb, 0.0))
    prod = (a[m] - np.float32(ma)) * (b[m] - np.float32(mb))
    return float(prod.astype(np.float64).sum() / n / (sa * sb))


def realignment_csv(scan, rows):
    """Text of writeRealignmentInfo (FrameWarpAligner.cpp:288-314). scan: dict of the
    ImageScanParameters constructor arguments; rows: per frame (unaligned correlation,
    correlation, coverage, D (nD x 3)). Doubles print as std::ostream does by default (%g)."""
    line, interline = scan["line"], scan["interline"]
    ny = max(scan["n_y"], 1)
    if interline < line:
        interline = 1.01 * line
    out = ["Realignment Information Version, 2",
           "Line Duration,%g" % line, "Interline Duration,%g" % interline,
           "Frame Duration,%g" % (ny * interline), "Interframe Duration,%g" % scan["interframe"],
           "Lines,%d" % ny, "Frames Per Stack,%d" % scan["n_z"]]
    nD = len(rows[0][3])
    out.append("Frame, UnalignedCorrelation, Correlation, Coverage" +
               "".join(", dx_%d, dy_%d, dz_%d" % (j, j, j) for j in range(nD)))
    for i, (u, c, cov, D) in enumerate(rows):
        out.append("%d,%g,%g, %g" % (i, u, c, cov) + "".join(", %g, %g, %g" % tuple(d) for d in D))
    return "\n".join(out) + "\n"
