"""Seeded synthetic acquisitions for the warp-registration path.

A stack is a Gaussian-blob phantom (the sample), a reference image of it, and
frames in which every pixel was acquired at its scan time t(p) while the
sample sat displaced by d(t), a per-scan-line displacement that is linear
between node times. Frames follow F(q) = P(q - d(t(q))), so the warper, which
pulls F at p + D(p) to match the reference at p, recovers D ~ d.

Everything is drawn from numpy's PCG64 generator seeded by the caller, so the
same seed gives the same arrays on every machine running this numpy.
"""
import numpy as np
from scipy import ndimage


def scan_timing(shape, line=1.0, interline=1.0, interframe=None):
    """Scan durations in the reference's units (ImageScanParameters.h:7-20)."""
    nz, ny, nx = shape
    if interframe is None:
        interframe = ny * max(interline, line)
    return dict(line=float(line), interline=float(interline), interframe=float(interframe))


def node_coordinate(shape, nD, line=1.0, interline=1.0, interframe=None):
    """u(p) = t(p)/stack*(nD-1) for every pixel, in double, following the node map of
    AbstractFrameWarper::precomputeInterp (FrameWarper.cpp:239-293) including the
    interframe cap of processScanParameters (FrameWarper.cpp:223-236)."""
    nz, ny, nx = shape
    st = scan_timing(shape, line, interline, interframe)
    line, interline, interframe = st["line"], st["interline"], st["interframe"]
    if interline < line:
        interline = 1.01 * line
    pixel = line / nx
    frame = ny * interline
    stack = (nz - 1) * interframe + frame
    cap = nz * frame / (nD - nz)
    if interframe - frame > cap:
        interframe = frame + cap
        stack = (nz - 1) * interframe + frame
    z = np.arange(nz, dtype=np.float64)[:, None, None]
    y = np.arange(ny, dtype=np.float64)[None, :, None]
    x = np.arange(nx, dtype=np.float64)[None, None, :]
    t = z * interframe + y * interline + x * pixel
    return t / stack * (nD - 1)


def blob_phantom(shape, n_blobs, sigma=(2.0, 6.0), amplitude=(20.0, 200.0), background=5.0,
                 rng=None, mean=None):
    """Sum of isotropic Gaussian blobs on a constant background (float64)."""
    rng = np.random.default_rng(rng)
    nz, ny, nx = shape
    img = np.zeros(shape, np.float64)
    cz = rng.uniform(0, nz, n_blobs) if nz > 1 else np.zeros(n_blobs)
    cy = rng.uniform(0, ny, n_blobs)
    cx = rng.uniform(0, nx, n_blobs)
    sg = rng.uniform(sigma[0], sigma[1], n_blobs)
    am = rng.uniform(amplitude[0], amplitude[1], n_blobs)
    for i in range(n_blobs):
        r = int(np.ceil(4 * sg[i]))
        x0, x1 = max(0, int(cx[i]) - r), min(nx, int(cx[i]) + r + 1)
        y0, y1 = max(0, int(cy[i]) - r), min(ny, int(cy[i]) + r + 1)
        gx = np.exp(-0.5 * ((np.arange(x0, x1) - cx[i]) / sg[i]) ** 2)
        gy = np.exp(-0.5 * ((np.arange(y0, y1) - cy[i]) / sg[i]) ** 2)
        if nz > 1:
            z0, z1 = max(0, int(cz[i]) - r), min(nz, int(cz[i]) + r + 1)
            gz = np.exp(-0.5 * ((np.arange(z0, z1) - cz[i]) / sg[i]) ** 2)
            img[z0:z1, y0:y1, x0:x1] += am[i] * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
        else:
            img[0, y0:y1, x0:x1] += am[i] * gy[:, None] * gx[None, :]
    if mean is not None:
        img *= (mean - background) / max(img.mean(), 1e-12)
    return img + background


def node_displacements(nD, n_dim, rng, walk_sigma=0.5, sin_amp=3.0, z_amp=2.0):
    """Node values: a random walk (walk_sigma px per node) plus a sinusoid of amplitude
    sin_amp with a random phase, per component; the z component (3-D) is scaled to stay
    within z_amp. Returns (nD, 3) with z = 0 in 2-D."""
    D = np.zeros((nD, 3))
    k = np.arange(nD)
    for c in range(n_dim):
        walk = np.cumsum(rng.normal(0.0, walk_sigma, nD))
        walk -= walk.mean()
        phase = rng.uniform(0, 2 * np.pi)
        D[:, c] = walk + sin_amp * np.sin(2 * np.pi * k / max(nD - 1, 1) + phase)
    if n_dim == 3:
        m = np.abs(D[:, 2]).max()
        if m > z_amp:
            D[:, 2] *= z_amp / m
    return D


def displacement_field(u, D):
    """d(p) linear between the nodes bracketing u(p)."""
    nD = D.shape[0]
    i = np.minimum(np.floor(u).astype(np.int64), nD - 2)
    f = u - i
    return [f * D[i + 1, c] + (1 - f) * D[i, c] for c in range(3)]


def displaced(phantom, d, order=3):
    """phantom sampled at q - d(q) (cval = edge background)."""
    nz, ny, nx = phantom.shape
    zz, yy, xx = np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(ny, dtype=np.float64),
                             np.arange(nx, dtype=np.float64), indexing="ij")
    bg = float(np.median(phantom))
    if nz == 1:
        out = ndimage.map_coordinates(phantom[0], [yy[0] - d[1][0], xx[0] - d[0][0]], order=order,
                                      mode="constant", cval=bg)
        return out[None]
    return ndimage.map_coordinates(phantom, [zz - d[2], yy - d[1], xx - d[0]], order=order,
                                   mode="constant", cval=bg)


def make_stack(shape, n_frames, nD, seed, n_blobs=200, sigma=(2.0, 6.0), amplitude=(20.0, 200.0),
               background=5.0, mean=None, dtype="float32", walk_sigma=0.5, sin_amp=3.0,
               z_amp=2.0, noise=True, order=3):
    """One acquisition: returns dict(reference, frames, D_true, nD, shape).

    reference and frames are raw photon counts (uint16) or floats, before the +1 offset
    the aligner adds (FrameWarpAligner.cpp:90,142)."""
    rng = np.random.default_rng(seed)
    shape = tuple(shape)
    n_dim = 3 if shape[0] > 1 else 2
    phantom = blob_phantom(shape, n_blobs, sigma, amplitude, background, rng, mean)
    ref = rng.poisson(phantom).astype(np.float64) if noise else phantom.copy()
    u = node_coordinate(shape, nD)
    frames = np.zeros((n_frames,) + shape, np.float64)
    Dt = np.zeros((n_frames, nD, 3))
    for f in range(n_frames):
        D = node_displacements(nD, n_dim, rng, walk_sigma, sin_amp, z_amp)
        Dt[f] = D
        clean = np.maximum(displaced(phantom, displacement_field(u, D), order), 0)
        frames[f] = rng.poisson(clean) if noise else clean
    if dtype == "uint16":
        ref = np.clip(ref, 0, 65535).astype(np.uint16)
        frames = np.clip(frames, 0, 65535).astype(np.uint16)
    else:
        ref = ref.astype(np.float32)
        frames = frames.astype(np.float32)
    return dict(reference=ref, frames=frames, D_true=Dt, nD=nD, shape=shape, n_dim=n_dim)


def with_offset(a):
    """float32(a) + 1.0f, the aligner's preprocessing of both reference and frames
    (FrameWarpAligner.cpp:86-90,123,141-142) with smoothing = 0."""
    return a.astype(np.float32) + np.float32(1.0)


# Workloads of the benchmark configurations (nodes = n_resampling_points * n_z + 1,
# FrameWarpAligner.cpp:100).
CONFIGS = {
    # CPU-warper workload: 256x256 float32, 20 frames, 10 resampling points
    "2d_256_f32": dict(shape=(1, 256, 256), n_frames=20, nD=11, n_blobs=200, dtype="float32"),
    # headline workload: 512x512 uint16 counts (~50 photons/px), 200 frames, 30 points
    "2d_512_u16": dict(shape=(1, 512, 512), n_frames=200, nD=31, n_blobs=600, mean=50.0, dtype="uint16"),
    # 3-D precision workload: 256x256x32 float32, 50 volumes, 16 points per plane
    "3d_256x256x32_f32": dict(shape=(32, 256, 256), n_frames=50, nD=16 * 32 + 1, n_blobs=1500,
                              sigma=(1.5, 4.0), dtype="float32"),
}
