"""Pins for the oracle functions whose reference source cannot be compiled here
(Project/WaP_Tomo.cpp needs GL/Windows headers): the SART correction, the voxel-driven
bicubic backprojector, the view order, frame windows, ROI mask and TV/TP step; plus
invariants of the compiled pieces that the goldens leave open (constant and linear
fields, line integrals of a Gaussian, the projector/scatter adjoint, geometry
consistency between the ray projector and the voxel backprojector).
Each check is chosen so a plausible slip (sign, index, transposed operand, dropped
term) fails it."""
import math

import numpy as np
import pytest

import helpers
import pyoracle
from helpers import SD


def geo(shape, nu, nv, parallel=False, **kw):
    d, h, w = shape
    L = max(shape)
    return pyoracle.make_geometry((w, h, d), nu, nv, sid=kw.get("sid", 4.0 * L),
                                  sdd=kw.get("sdd", 6.0 * L), parallel=parallel,
                                  ds=kw.get("ds", 1.0))


# ---------------------------------------------------------------- SART correction
def test_sart_correct_closed_form(oracle):
    rng = np.random.default_rng(0)
    b, ax, rl, py = [rng.random(50).astype(np.float32) * s for s in (5, 5, 40, 0.3)]
    corr, py2 = oracle.sart_correct(b, ax, rl, py, 0.25)
    want = ((b - ax) - py) / (rl + np.float32(1.0))
    np.testing.assert_array_equal(corr, want.astype(np.float32))
    np.testing.assert_array_equal(py2, (py.astype(np.float64) + 0.25 * corr).astype(np.float32))
    # exact data and zero dual give a zero correction
    c0, _ = oracle.sart_correct(ax, ax, rl, np.zeros_like(py), 0.25)
    assert not c0.any()


# ---------------------------------------------------------------- voxel backprojector
def _project_point_f64(view, x, y, z):
    A = np.array(view.mvv2m, np.float64).reshape(4, 4).T
    det = A @ np.array([x, y, z, 1.0])
    if view.parallel:
        u = (view.half_w + det[0]) / view.ds
        v = (view.half_h + det[1]) / view.ds
    else:
        u = (view.half_w - det[0] / det[2] * view.sdd) / view.ds
        v = (view.half_h - det[1] / det[2] * view.sdd) / view.ds
    return u, v


@pytest.mark.parametrize("parallel", [False, True])
def test_voxel_bp_reproduces_linear_detector_data(oracle, parallel):
    """Catmull-Rom reproduces linear data exactly, so backprojecting corr(u,v) = a*u +
    b*v + c gives, at every voxel whose 4x4 stencil is inside the detector, the linear
    function at the voxel's projected detector position (computed here independently
    in FP64)."""
    shape, nu, nv = (12, 14, 16), 40, 36
    v = oracle.view(geo(shape, nu, nv, parallel), 63.0)
    uu, vv = np.meshgrid(np.arange(nu), np.arange(nv))
    a, b, c = 0.37, -0.21, 1.3
    corr = (a * uu + b * vv + c).astype(np.float32)
    out = oracle.backproject_voxel(np.zeros(shape, np.float32), corr, v, 1.0, 0.0, (0, 0, 0),
                                   (16, 14, 12), 1)
    n = 0
    for z in range(12):
        for y in range(14):
            for x in range(16):
                pu, pv = _project_point_f64(v, x, y, z)
                if 2 <= int(pu) < nu - 2 and 2 <= int(pv) < nv - 2:
                    assert abs(out[z, y, x] - (a * pu + b * pv + c)) < 2e-4
                    n += 1
                else:
                    assert out[z, y, x] == 0.0
    assert n > 500


def test_voxel_bp_prox_clamps_and_respects_box(oracle):
    shape, nu, nv = (10, 10, 10), 24, 24
    v = oracle.view(geo(shape, nu, nv), 10.0)
    corr = np.full((nv, nu), 3.0, np.float32)
    vol = np.full(shape, 0.5, np.float32)
    out = oracle.backproject_voxel(vol, corr, v, 0.5, 1.2, (2, 3, 4), (8, 7, 9), 0)
    inside = np.zeros(shape, bool)
    inside[4:9, 3:7, 2:8] = True
    np.testing.assert_array_equal(out[~inside], 0.5)
    np.testing.assert_array_equal(out[inside], np.float32(1.2))  # 0.5 + 0.5*3 clamped
    out2 = oracle.backproject_voxel(vol, -corr, v, 0.5, 1.2, (2, 3, 4), (8, 7, 9), 0)
    np.testing.assert_array_equal(out2[inside], 0.0)


@pytest.mark.parametrize("parallel", [False, True])
def test_ray_and_voxel_geometry_agree(oracle, parallel):
    """A point on the ray of detector pixel (i, j) backprojects to (i, j)."""
    shape, nu, nv = (20, 22, 24), 48, 44
    v = oracle.view(geo(shape, nu, nv, parallel), 137.0)
    z, u, tn, tf = oracle.produce_rays(shape, v, nu, nv)
    hits = np.nonzero(tn < tf)[0]
    assert len(hits) > 200
    for k in hits[:: max(1, len(hits) // 200)]:
        t = 0.5 * (tn[k] + tf[k])
        p = z[k, :3].astype(np.float64) + u[k, :3] * t
        pu, pv = _project_point_f64(v, *p)
        assert abs(pu - k % nu) < 2e-3 and abs(pv - k // nu) < 2e-3


# ---------------------------------------------------------------- camera chain
@pytest.mark.parametrize("parallel", [False, True])
@pytest.mark.parametrize("angle", [0.0, 33.0, 90.0, 211.7])
def test_view_geometry_closed_form(oracle, parallel, angle):
    """The per-view rays follow from R_y(theta) alone, independent of the matrix chain:
    the view frame is the volume frame rotated by theta about +y and moved back by sid
    (theta = deg * 3.1415926/180, as the reference's RotateVolume); in view coordinates
    the source sits at the origin and pixel (i, j) at ((i - cu) ds, (j - cv) ds, -sdd)
    (cone) or its rays leave ((i - cu) ds, (j - cv) ds, 0) along -z (parallel)."""
    shape, nu, nv, ds, sid, sdd = (20, 22, 24), 31, 27, 1.3, 96.0, 144.0
    d, h, w = shape
    g = pyoracle.make_geometry((w, h, d), nu, nv, sid=sid, sdd=sdd, ds=ds, parallel=parallel)
    v = oracle.view(g, angle)
    th = angle * 3.1415926 / 180.0
    c, s = math.cos(th), math.sin(th)
    Rt = np.array([[c, 0, -s], [0, 1, 0], [s, 0, c]])  # R_y(theta)^T: view -> volume
    centre = np.array([(w - 1) / 2.0, (h - 1) / 2.0, (d - 1) / 2.0])
    to_vol = lambda e: centre + Rt @ (np.asarray(e, np.float64) + np.array([0.0, 0.0, sid]))
    z, u, tn, tf = oracle.produce_rays(shape, v, nu, nv)
    cu, cv = (nu - 1) / 2.0, (nv - 1) / 2.0
    for k in range(0, nu * nv, 7):
        i, j = k % nu, k // nu
        org = z[k, :3].astype(np.float64)
        dirv = u[k, :3].astype(np.float64)
        assert abs(np.linalg.norm(dirv) - 1.0) < 1e-5
        if parallel:
            expect_org = to_vol(((i - cu) * ds, (j - cv) * ds, 0.0))
            expect_dir = Rt @ np.array([0.0, 0.0, -1.0])
            # the origin lies on the ray through the pixel's source-plane point
            off = org - expect_org
            assert np.linalg.norm(off - np.dot(off, expect_dir) * expect_dir) < 2e-3
        else:
            expect_org = to_vol((0.0, 0.0, 0.0))
            pix = to_vol(((i - cu) * ds, (j - cv) * ds, -sdd))
            expect_dir = (pix - expect_org) / np.linalg.norm(pix - expect_org)
            np.testing.assert_allclose(org, expect_org, atol=2e-3)
        np.testing.assert_allclose(dirv, expect_dir, atol=2e-5)


# ---------------------------------------------------------------- projector invariants
@pytest.mark.parametrize("parallel", [False, True])
def test_constant_volume_integral(oracle, parallel):
    shape, nu, nv = (16, 18, 20), 40, 40
    v = oracle.view(geo(shape, nu, nv, parallel), 29.0)
    p, rl, w = oracle.forward_project(np.full(shape, 2.5, np.float32), v, nu, nv, SD, f64=False)
    hit = w > 0
    np.testing.assert_allclose(p[hit], 2.5 * float(SD) * w[hit], rtol=2e-6)
    # every sd along the chord is sampled at least once (segment restarts at the 0.5
    # planes can re-take samples, so this is a lower bound)
    assert np.all(w[hit] >= np.floor(rl[hit] / SD) + 1)


def test_gaussian_line_integral(oracle):
    """Parallel rays through an isotropic Gaussian blob centred on a grid point, along
    grid lines: integral = A*s*sqrt(2pi)*exp(-d^2/2s^2) (FP64 oracle; the trilinear
    samples then integrate as a fine trapezoid rule)."""
    L, s, A = 49, 4.0, 1.0
    c = (L - 1) / 2.0
    zz, yy, xx = np.meshgrid(*[np.arange(L)] * 3, indexing="ij")
    vol = A * np.exp(-((xx - c) ** 2 + (yy - c) ** 2 + (zz - c) ** 2) / (2 * s * s))
    nu = nv = 31
    g = pyoracle.make_geometry((L, L, L), nu, nv, ds=1.0, sid=2.0 * L, sdd=3.0 * L,
                               parallel=True)
    v = oracle.view(g, 0.0)
    p, _, _ = oracle.forward_project(vol, v, nu, nv, 0.05, f64=True)
    peak = A * s * math.sqrt(2 * math.pi)
    for (i, j) in [(15, 15), (10, 15), (15, 22), (5, 9)]:
        du, dv = i - 15, j - 15
        want = peak * math.exp(-(du * du + dv * dv) / (2 * s * s))
        assert abs(p[j, i] - want) < 1e-3 * peak
    # an off-axis pixel reads the blob displaced the other way on the opposite view
    v180 = oracle.view(g, 180.0)
    p180, _, _ = oracle.forward_project(vol, v180, nu, nv, 0.05, f64=True)
    assert abs(p180[15, 10] - p[15, 20]) < 1e-6 * peak


@pytest.mark.parametrize("parallel", [False, True])
def test_projector_adjoint_fp64(oracle, parallel):
    shape, nu, nv = (9, 10, 11), 16, 14
    v = oracle.view(geo(shape, nu, nv, parallel), 71.0)
    rng = np.random.default_rng(3)
    x = rng.random(shape)
    y = rng.standard_normal((nv, nu))
    ax = oracle.forward_project(x, v, nu, nv, float(SD), f64=True)[0]
    aty = oracle.backproject_ray(y, shape, v, float(SD), scale=float(SD), f64=True)[0]
    lhs, rhs = float((ax * y).sum()), float((x * aty).sum())
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_projector_brute_force_matrix(oracle):
    """Explicit system matrix on 6^3 x 8^2 from unit impulses equals the scatter's
    transpose (brute force over every voxel)."""
    shape, nu, nv = (6, 6, 6), 8, 8
    v = oracle.view(geo(shape, nu, nv, sid=24.0, sdd=36.0), 40.0)
    cols = []
    for q in range(216):
        e = np.zeros(216)
        e[q] = 1.0
        cols.append(oracle.forward_project(e.reshape(shape), v, nu, nv, 1.0, f64=True)[0].ravel())
    A = np.stack(cols, axis=1)
    rows = []
    for r in range(nu * nv):
        e = np.zeros(nu * nv)
        e[r] = 1.0
        rows.append(oracle.backproject_ray(e.reshape(nv, nu), shape, v, 1.0, f64=True)[0].ravel())
    np.testing.assert_allclose(A, np.stack(rows), atol=1e-12)
    assert A.sum() > 0


# ---------------------------------------------------------------- advection / image ops
def test_advect_zero_flow_and_translation(oracle):
    shape = (12, 13, 14)
    src = helpers.blob_volume(shape, 4)
    z = np.zeros((3,) + shape, np.float32)
    out = oracle.advect(z, src, 0.7)
    np.testing.assert_array_equal(out[1:-1, 1:-1, 1:-1], src[1:-1, 1:-1, 1:-1])
    # boundary layer takes the clamped neighbour (grid clamped to [1, dim-2])
    np.testing.assert_array_equal(out[0, 1:-1, 1:-1], src[1, 1:-1, 1:-1])
    # uniform velocity (1, 0, 0) over dt = 1 shifts by one voxel in x
    vel = np.zeros((3,) + shape, np.float32)
    vel[0] = 1.0
    out = oracle.advect(vel, src, 1.0)
    np.testing.assert_allclose(out[2:-2, 2:-2, 3:-2], src[2:-2, 2:-2, 2:-3], atol=1e-6)


def test_blur_and_resize_invariants(oracle):
    c = np.full((10, 11, 12), 1.75, np.float32)
    np.testing.assert_allclose(oracle.blur(c, 0.8), 1.75, rtol=1e-5)
    imp = np.zeros((1, 1, 41), np.float32)
    imp[0, 0, 20] = 1.0
    assert abs(oracle.blur(imp, 2.0).sum() - 1.0) < 1e-3
    np.testing.assert_allclose(oracle.resize(c, (5, 5, 6)), 1.75, rtol=1e-6)
    np.testing.assert_allclose(oracle.resize(c, (20, 22, 24)), 1.75, rtol=1e-6)
    # cubic growth of a linear ramp stays linear at the resampled positions away from
    # the end cells (where the stencil replicates the edge sample)
    ramp = np.tile(np.arange(12, dtype=np.float32), (1, 1, 1))
    up = oracle.resize(ramp, (1, 1, 23))
    np.testing.assert_allclose(up[0, 0, 2:-2], (np.arange(23) * 11 / 22)[2:-2], atol=1e-5)
    assert up[0, 0, 1] != 0.5  # the replicated edge bends the first interval


def test_flow_identical_frames_zero_flow(oracle):
    shape = (12, 12, 12)
    I = oracle.blur(helpers.blob_volume(shape, 5), 1.0)
    z3 = np.zeros((3,) + shape, np.float32)
    z1 = np.zeros(shape, np.float32)
    V = oracle.flow_single(I, I, z3, z3, z3, z3, z1, 0.008, 0.05, 30, 2, [0, 12, 0, 12, 0, 12],
                           [2] * 6)[0]
    assert np.abs(V).max() == 0.0


# ---------------------------------------------------------------- loop bookkeeping
def test_view_order(oracle):
    assert list(oracle.view_order(3, 9)) == [3, 9, 4, 8, 5, 7, 6]
    assert list(oracle.view_order(0, 5)) == [0, 5, 1, 4, 2, 3]
    assert list(oracle.view_order(7, 7)) == [7]
    for a, b in [(0, 30), (11, 17), (5, 6)]:
        assert sorted(oracle.view_order(a, b)) == list(range(a, b + 1))


def test_frame_windows(oracle):
    ft, s, e = oracle.frame_windows(8, 64)
    np.testing.assert_array_equal(ft, np.arange(8) * 8.0)
    np.testing.assert_array_equal(s, np.maximum(0, (np.arange(8) * 8 - 3)))
    np.testing.assert_array_equal(e, np.minimum(np.arange(8) * 8 + 3, 63))
    ft, s, e = oracle.frame_windows(3, 10)
    np.testing.assert_array_equal(s, [0, 0, 3])
    np.testing.assert_array_equal(e, [3, 6, 9])


def _vdc(i, p):
    """Exact radical inverse of i in base p (most significant digit last)."""
    from fractions import Fraction
    r, q = Fraction(0), Fraction(1, p)
    while i:
        r += (i % p) * q
        q /= p
        i //= p
    return r


def test_low_discrepancy_closed_forms(oracle):
    from fractions import Fraction
    # base 2, one segment: 360 * (0, 1/2, 1/4, 3/4, 1/8, 5/8, 3/8, 7/8), exact in double
    assert list(oracle.low_discrepancy(8, 1, 2, 0)) == [0, 180, 90, 270, 45, 225, 135, 315]
    # base 3: 0, 1/3, 2/3, 1/9, 4/9, 7/9, 2/9, 5/9, 8/9 of the circle
    np.testing.assert_allclose(oracle.low_discrepancy(9, 1, 3, 0),
                               [0, 120, 240, 40, 160, 280, 80, 200, 320], rtol=0, atol=1e-12)
    # 2^m base-2 points are exactly the 2^m equispaced angles
    a = oracle.low_discrepancy(256, 1, 2, 0)
    assert sorted(a) == [360.0 * j / 256 for j in range(256)]
    # segments (the scanner script's Np = 4000, Ns = 10): row i of the Ns-column
    # reshape is 36 * vdc(i) + 36 k, k = 0..9; every segment holds Np / Ns angles
    for base in (2, 3, 5):
        a = oracle.low_discrepancy(4000, 10, base, 10000)
        want = [float(Fraction(360) * _vdc(i, base) / 10 + Fraction(360 * k, 10))
                for i in range(400) for k in range(10)]
        np.testing.assert_allclose(a, want, rtol=1e-14, atol=1e-12)
        counts = np.bincount((a // 36.0).astype(int), minlength=10)
        assert list(counts) == [400] * 10


def test_low_discrepancy_reference_views(oracle):
    # the demo config (Np = 100, Ns = 1, p = 2, Nz = 50): one 0-degree view at index 50
    a = oracle.low_discrepancy(100, 1, 2, 50)
    assert a[50] == 0.0
    rest = np.delete(a, 50)
    assert list(rest) == [float(360 * _vdc(i, 2)) for i in range(99)]
    # with one segment the inserted views land at every multiple of Nz
    a = oracle.low_discrepancy(400, 1, 2, 50)
    zeros = [int(i) for i in np.flatnonzero(a == 0.0)]
    assert zeros == [0, 50, 100, 150, 200, 250, 300, 350]
    # Nz = 1: every other view is the reference view
    a = oracle.low_discrepancy(9, 1, 2, 1)
    assert list(a) == [0, 0, 180, 0, 90, 0, 270, 0, 45]
    # with Ns segments the check runs between batches only: with Ns = 4, Nz = 6 the
    # count is 4, 8, 12 (insert -> 13), 17, 21, 25, 29 -- one insertion, at index 12
    a = oracle.low_discrepancy(30, 4, 2, 6)
    assert list(a[8:12]) == [22.5, 112.5, 202.5, 292.5]
    assert a[12] == 0.0 and list(a[13:17]) == [67.5, 157.5, 247.5, 337.5]
    assert a[29] == 78.75
    assert list(np.flatnonzero(a == 0.0)) == [0, 12]


def test_roi_mask_brute_force(oracle):
    shape = (13, 11, 15)
    r = [2, 3, 1, 2, 3, 2, 4, 3]
    m = oracle.roi_mask(shape, r)
    D, H, W = shape
    for k in range(D):
        for j in range(H):
            for i in range(W):
                corner = ((i < r[6] or i > W - r[6]) and (k < r[7] or k > D - r[7])) or \
                         ((k < r[6] or k > D - r[6]) and (i < r[7] or i > W - r[7]))
                box = r[0] <= i <= W - r[1] and r[2] <= j <= H - r[3] and r[4] <= k <= D - r[5]
                assert m[k, j, i] == (1.0 if box and not corner else 0.0)


def test_tvtp_step_brute_force(oracle):
    shape = (7, 8, 9)
    rng = np.random.default_rng(6)
    vol = rng.random(shape).astype(np.float32)
    xb = rng.random(shape).astype(np.float32)
    xn = rng.random(shape).astype(np.float32)
    ytv = (rng.standard_normal((3,) + shape) * 0.1).astype(np.float32)
    ytp = (rng.standard_normal(shape) * 0.1).astype(np.float32)
    ytp_prev = (rng.standard_normal(shape) * 0.1).astype(np.float32)
    roi = [1, 1, 1, 1, 1, 1]
    sig, tau, tvw, tpw, gmax = np.float32(0.1125), np.float32(1.0), np.float32(0.05), \
        np.float32(0.3), np.float32(0.9)
    args = [np.ascontiguousarray(a).copy() for a in (vol, np.zeros(shape, np.float32), xb, xn,
                                                      ytv, ytp, ytp_prev)]
    d = pyoracle._dims(shape)
    oracle.lib.orc_tvtp_step_f32(1, d, np.array(roi, np.int32), args[0], args[1], args[2],
                                 args[3], args[4], args[5], args[6], sig, tau, tvw, tpw, gmax)
    # independent float32 restatement
    Ytp = ytp + sig * (xn - xb)
    Ytv = ytv.copy()
    V = vol.copy()
    for k in range(1, 6):
        for j in range(1, 7):
            for i in range(1, 8):
                for c, (di, dj, dk) in enumerate([(1, 0, 0), (0, 1, 0), (0, 0, 1)]):
                    val = Ytv[c, k, j, i] + sig * (xb[k + dk, j + dj, i + di] - xb[k, j, i]) * \
                        np.float32(1.0)
                    Ytv[c, k, j, i] = max(-tvw, min(tvw, val))
                Ytp[k, j, i] = Ytp[k, j, i] * tpw / (tpw + np.float32(2) * sig)
    V = V - tau * (ytp_prev - Ytp)
    for k in range(1, 6):
        for j in range(1, 7):
            for i in range(1, 8):
                y1 = -(Ytv[0, k, j, i] - Ytv[0, k, j, i - 1] + Ytv[1, k, j, i] -
                       Ytv[1, k, j - 1, i] + Ytv[2, k, j, i] - Ytv[2, k - 1, j, i])
                V[k, j, i] = max(min(gmax, V[k, j, i] - tau * y1), np.float32(0))
    np.testing.assert_allclose(args[4], Ytv, atol=1e-7)
    np.testing.assert_allclose(args[5], Ytp, atol=1e-7)
    np.testing.assert_allclose(args[0], V, atol=1e-6)
    np.testing.assert_array_equal(args[1], vol)


def test_degenerate_rays_terminate_with_bounded_samples(oracle):
    """Rays lying exactly in a face plane (parallel beam at 90 degrees puts pixel
    origins on integer coordinates with a zero direction component) and rays with a
    tiny negative component must keep a chord-sized sample count (resolved defects of
    the reference march; see the ray-setup comments in the oracle)."""
    shape, nu, nv = (18, 20, 24), 32, 24
    g = pyoracle.make_geometry((24, 20, 18), nu, nv, sid=48.0, sdd=72.0, parallel=True)
    for a in (90.0, 180.0, 270.0):
        v = oracle.view(g, a)
        _, rl, w = oracle.forward_project(np.ones(shape, np.float32), v, nu, nv, SD)
        hit = w > 0
        assert hit.any()
        assert np.all(w[hit] <= np.floor(rl[hit] / SD) + 8)
        assert rl.max() < 24 * 1.01


def test_xbar_update_brute_force(oracle):
    """X_bar = 2x - x_prev inside the ROI box, untouched outside it."""
    shape = (7, 8, 9)
    rng = np.random.default_rng(8)
    x, xp, xb = (rng.random(shape).astype(np.float32) for _ in range(3))
    roi = np.array([1, 2, 0, 1, 2, 1], np.int32)
    out = xb.copy()
    oracle.lib.orc_xbar_update_f32(pyoracle._dims(shape), roi, x, xp, out)
    expect = xb.copy()
    d, h, w = shape
    for k in range(roi[4], d - roi[5]):
        for j in range(roi[2], h - roi[3]):
            for i in range(roi[0], w - roi[1]):
                expect[k, j, i] = np.float32(2) * x[k, j, i] - xp[k, j, i]
    np.testing.assert_array_equal(out, expect)
