"""Host side of the C ABI and of the drop-in C++ classes (no GPU needed).

* libwap.so loads and exports every entry point include/wap.h declares;
* the host helpers (wap_view_geometry, wap_view_order, wap_frame_windows, wap_roi_mask)
  equal the oracle bit for bit, and wap_view_geometry rebuilds the view matrices stored
  with the projector goldens (tests/golden/g1a.npz, g1b.npz);
* the C++ math classes (Matrix4x4 product and inverse, Quaternion::ToMatrix) reproduce
  the reference's matrices of tests/golden/g8.npz bit for bit;
* CoherentVolume bricking and the Scene camera chain (tests/cpp/test_classes.cpp).
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle
import wap_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
CPP = os.path.join(ROOT, "tests", "cpp")
CLASSES_BIN = os.path.join(CPP, "test_classes")


def header_entry_points():
    text = open(os.path.join(ROOT, "include", "wap.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(wap_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_entry_point():
    lib = wap_amd.load()
    names = header_entry_points()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python mirror binds each of them
    assert sorted(wap_amd.SIGNATURES) == names
    assert lib.wap_version() == 1


def test_struct_layouts_match_oracle():
    assert C.sizeof(wap_amd.WapScanGeometry) == C.sizeof(pyoracle.WapScanGeometry) == 64
    assert C.sizeof(wap_amd.WapView) == C.sizeof(pyoracle.WapView) == 256
    assert C.sizeof(wap_amd.WapLoopConfig) == C.sizeof(pyoracle.OrcLoop)
    for a, b in zip(wap_amd.WapLoopConfig._fields_, pyoracle.OrcLoop._fields_):
        assert a[0] == b[0]
        assert C.sizeof(a[1]) == C.sizeof(b[1])


GEOMETRIES = [
    dict(dims=(16, 16, 16), nu=32, nv=24, sid=48.0, sdd=72.0),
    dict(dims=(37, 21, 29), nu=41, nv=33, sid=90.0, sdd=150.0, ds=0.75, spacing=1.25),
    dict(dims=(24, 24, 24), nu=40, nv=36, sid=96.0, sdd=144.0, translate=(1.5, -2.0, 0.25)),
    dict(dims=(20, 28, 16), nu=30, nv=26, parallel=True, ds=1.1, ccw=False),
]


@pytest.mark.parametrize("gi", range(len(GEOMETRIES)))
def test_view_geometry_equals_oracle_bitwise(oracle, gi):
    kw = dict(GEOMETRIES[gi])
    dims = kw.pop("dims")
    g = wap_amd.make_geometry(dims, **kw)
    go = pyoracle.make_geometry(dims, **kw)
    for a in [0.0, 0.5, 33.3, 89.999, 90.0, 137.5, 180.0, 271.1, 359.9, -45.0]:
        assert bytes(wap_amd.view_geometry(g, a)) == bytes(oracle.view(go, a)), a


@pytest.mark.parametrize("name,dims,nu,nv", [("g1a", (16, 16, 16), 32, 24),
                                              ("g1b", (24, 20, 18), 40, 36)])
def test_view_geometry_rebuilds_golden_views(name, dims, nu, nv):
    """The views stored with the projector goldens (whose projections the reference's
    own projector produced from them) are rebuilt bit for bit."""
    stored = np.load(os.path.join(GOLD, name + ".npz"))["views"]
    g = wap_amd.make_geometry(dims, nu, nv, sid=3.0 * dims[0], sdd=4.5 * dims[0])
    for k, a in enumerate([0.0, 45.0, 90.0, 180.0, 211.7]):
        assert bytes(wap_amd.view_geometry(g, a)) == stored[k].tobytes()


def test_view_geometry_rejects_bad_geometry():
    lib = wap_amd.load()
    g = wap_amd.make_geometry((1, 16, 16), 32, 24)
    v = wap_amd.WapView()
    assert lib.wap_view_geometry(C.byref(g), 0.0, C.byref(v)) == -1
    g = wap_amd.make_geometry((16, 16, 16), 32, 24, ds=0.0)
    assert lib.wap_view_geometry(C.byref(g), 0.0, C.byref(v)) == -1
    assert lib.wap_view_geometry(None, 0.0, C.byref(v)) == -1


@pytest.mark.parametrize("first,last", [(0, 0), (0, 1), (3, 9), (5, 17), (0, 191)])
def test_view_order_equals_oracle(oracle, first, last):
    lib = wap_amd.load()
    out = (C.c_int * (last - first + 1))()
    assert lib.wap_view_order(first, last, out) == 0
    np.testing.assert_array_equal(np.array(out[:]), oracle.view_order(first, last))
    assert sorted(out[:]) == list(range(first, last + 1))
    assert lib.wap_view_order(4, 3, out) == -1


@pytest.mark.parametrize("nf,total", [(1, 7), (5, 192), (12, 192), (130, 192), (7, 1000)])
def test_frame_windows_equal_oracle(oracle, nf, total):
    lib = wap_amd.load()
    ft = (C.c_float * nf)()
    s = (C.c_int * nf)()
    e = (C.c_int * nf)()
    assert lib.wap_frame_windows(nf, total, 3.0, ft, s, e) == 0
    oft, os_, oe = oracle.frame_windows(nf, total, 3.0)
    assert np.array(ft[:], np.float32).tobytes() == oft.tobytes()
    np.testing.assert_array_equal(np.array(s[:]), os_)
    np.testing.assert_array_equal(np.array(e[:]), oe)


@pytest.mark.parametrize("np_,ns,p,nz", [(0, 1, 2, 0), (1, 1, 2, 0), (100, 1, 2, 50),
                                         (4000, 10, 2, 10000), (333, 7, 3, 20), (64, 4, 5, 1),
                                         (1000, 3, 7, 0)])
def test_low_discrepancy_angles_equal_oracle(oracle, np_, ns, p, nz):
    got = wap_amd.low_discrepancy_angles(np_, ns, p, nz)
    want = oracle.low_discrepancy(np_, ns, p, nz)
    assert np.array(got, np.float64).tobytes() == want.tobytes()


def test_low_discrepancy_angles_reject_bad_arguments():
    lib = wap_amd.load()
    out = (C.c_double * 4)()
    for args in [(-1, 1, 2, 0), (4, 0, 2, 0), (4, 1, 1, 0), (4, 1, 2, -1)]:
        assert lib.wap_low_discrepancy_angles(*args, out) == -1
    assert lib.wap_low_discrepancy_angles(4, 1, 2, 0, None) == -1
    assert lib.wap_low_discrepancy_angles(0, 1, 2, 0, None) == 0


@pytest.mark.parametrize("roi", [(0,) * 8, (2,) * 8, (1, 3, 2, 0, 4, 1, 5, 2), (8,) * 8])
def test_roi_mask_equals_oracle(oracle, roi):
    lib = wap_amd.load()
    shape = (13, 17, 19)  # (D, H, W)
    m = np.zeros(shape, np.float32)
    r = lib.wap_roi_mask(wap_amd.ints(19, 17, 13), wap_amd.ints(*roi),
                         m.ctypes.data_as(C.POINTER(C.c_float)))
    assert r == 0
    np.testing.assert_array_equal(m, oracle.roi_mask(shape, roi))


def _classes_bin():
    if not os.path.exists(CLASSES_BIN):
        subprocess.run(["make", "-C", CPP], check=True, capture_output=True)
    return CLASSES_BIN


def test_math_classes_reproduce_reference_matrices(tmp_path):
    g = np.load(os.path.join(GOLD, "g8.npz"))
    inp = np.concatenate([g["A"], g["B"], g["quat_angles"]]).astype(np.float32)
    fi, fo = tmp_path / "in.bin", tmp_path / "out.bin"
    inp.tofile(fi)
    subprocess.run([_classes_bin(), "math", str(fi), str(fo)], check=True)
    out = np.fromfile(fo, np.float32)
    assert out[:16].tobytes() == g["AB"].astype(np.float32).tobytes()
    assert out[16:32].tobytes() == g["invA"].astype(np.float32).tobytes()
    assert out[32:].tobytes() == g["quat"].astype(np.float32).reshape(-1).tobytes()
    # and the inverse is an inverse
    A = g["A"].reshape(4, 4).astype(np.float64)
    iA = out[16:32].reshape(4, 4).astype(np.float64)
    np.testing.assert_allclose(A @ iA, np.eye(4), atol=1e-4)


def test_classes_host_parts():
    r = subprocess.run([_classes_bin(), "cpu"], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout


@pytest.mark.gpu
def test_classes_on_device():
    """PSART views, SART_Proximal_Operator, WarpAndProjct, RenderContent (plain and
    warped) and the free kernels advection::advect and
    optical_flow::L1TVOpticalFlowNonlinear_MultiScale3D equal the C ABI bit for bit;
    Render() runs both temporal levels."""
    r = subprocess.run([_classes_bin(), "gpu"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok ") >= 14
