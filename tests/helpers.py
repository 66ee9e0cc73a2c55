"""Shared test inputs: small seeded volumes, scans and flows."""
import numpy as np

import phantoms


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def blob_volume(shape, seed, nblobs=12):
    rng = phantoms.SplitMix64(phantoms.SEED_BASE + seed)
    return phantoms.blob_phantom(shape, nblobs, rng, sigma=(1.5, 4.0))


def smooth_flow(shape, seed, amp=1.5):
    rng = phantoms.SplitMix64(phantoms.SEED_BASE + 1000 + seed)
    return phantoms.smooth_field(shape, 6, 2, amp, rng)


def cone_geometry(mod, shape, nu, nv, ds=1.0):
    d, h, w = shape
    L = max(shape)
    return mod.make_geometry((w, h, d), nu, nv, ds=ds, sid=4.0 * L, sdd=6.0 * L)


def parallel_geometry(mod, shape, nu, nv, ds=1.0):
    d, h, w = shape
    L = max(shape)
    return mod.make_geometry((w, h, d), nu, nv, ds=ds, sid=2.0 * L, sdd=3.0 * L, parallel=True)


SD = np.float32(np.sqrt(3.0) / 2.0)
