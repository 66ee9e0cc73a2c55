"""Deterministic synthetic workloads for the warp-and-project loop.

Random numbers come from a SplitMix64 stream (uniforms in [0,1) from the top 53 bits)
with Box-Muller normals, so the same seed gives the same phantom, deformation, angles
and noise everywhere (tests, benchmark and the CPU baseline).
Arrays are numpy, planar x-fastest: shape (D, H, W) or (C, D, H, W).
"""
import math

import numpy as np

MASK64 = (1 << 64) - 1
SEED_BASE = 0x5741500000000000


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def normal(self):
        u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(max(u1, 1e-300))) * math.cos(2.0 * math.pi * u2)

    def uniforms(self, n, lo=0.0, hi=1.0):
        return np.array([self.uniform(lo, hi) for _ in range(n)])

    def normals(self, n):
        return np.array([self.normal() for _ in range(n)])


def golden_angles(n, span_deg):
    g = (math.sqrt(5.0) - 1.0) / 2.0
    return np.array([((i * g) % 1.0) * span_deg for i in range(n)], dtype=np.float64)


def blob_phantom(shape, nblobs, rng, centre=(0.1875, 0.8125), sigma=(2.0, 6.0),
                 amp=(0.2, 1.0)):
    """Sum of isotropic Gaussian blobs; centres uniform in the given fraction of each axis."""
    D, H, W = shape
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    out = np.zeros(shape, np.float64)
    for _ in range(nblobs):
        cx = rng.uniform(centre[0] * W, centre[1] * W)
        cy = rng.uniform(centre[0] * H, centre[1] * H)
        cz = rng.uniform(centre[0] * D, centre[1] * D)
        s = rng.uniform(*sigma)
        a = rng.uniform(*amp)
        out += a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (2 * s * s))
    return out.astype(np.float32)


def foam_phantom(L, nvoids, rng, axes=(4.0, 16.0), margin=0.125):
    """Unit-density cube (inside the margin) with ellipsoidal voids of random
    orientation; each void is rasterised over its bounding box only."""
    vol = np.zeros((L, L, L), np.float32)
    m = int(round(margin * L))
    vol[m:L - m, m:L - m, m:L - m] = 1.0
    for _ in range(nvoids):
        c = np.array([rng.uniform(m, L - m) for _ in range(3)])
        ax = np.array([rng.uniform(*axes) for _ in range(3)])
        # random rotation from a normalised quaternion
        q = np.array([rng.normal() for _ in range(4)])
        q /= np.linalg.norm(q)
        w, a, b, d = q
        R = np.array([[1 - 2 * (b * b + d * d), 2 * (a * b - d * w), 2 * (a * d + b * w)],
                      [2 * (a * b + d * w), 1 - 2 * (a * a + d * d), 2 * (b * d - a * w)],
                      [2 * (a * d - b * w), 2 * (b * d + a * w), 1 - 2 * (a * a + b * b)]])
        r = int(math.ceil(ax.max()))
        lo = np.maximum(np.floor(c - r).astype(int), 0)
        hi = np.minimum(np.ceil(c + r).astype(int) + 1, L)
        zz, yy, xx = np.meshgrid(np.arange(lo[2], hi[2]), np.arange(lo[1], hi[1]),
                                 np.arange(lo[0], hi[0]), indexing="ij")
        p = np.stack([xx - c[0], yy - c[1], zz - c[2]], axis=-1) @ R
        inside = ((p / ax) ** 2).sum(-1) <= 1.0
        sub = vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
        sub[inside] = 0.0
    return vol


def smooth_field(shape, nmodes, kmax, max_amp, rng):
    """3-component displacement from random low-frequency Fourier modes, scaled so the
    largest magnitude is max_amp voxels."""
    D, H, W = shape
    z, y, x = np.meshgrid(np.arange(D) / D, np.arange(H) / H, np.arange(W) / W, indexing="ij")
    f = np.zeros((3,) + shape, np.float64)
    for _ in range(nmodes):
        k = [int(rng.uniform(0, kmax + 1)) for _ in range(3)]
        ph = rng.uniform(0, 2 * math.pi)
        amp = [rng.normal() for _ in range(3)]
        wave = np.cos(2 * math.pi * (k[0] * x + k[1] * y + k[2] * z) + ph)
        for c in range(3):
            f[c] += amp[c] * wave
    mag = np.sqrt((f ** 2).sum(0)).max()
    if mag > 0:
        f *= max_amp / mag
    return f.astype(np.float32)
