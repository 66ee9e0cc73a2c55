"""The oracle against golden vectors written by the reference's own compiled code
(tests/golden/make_golden.py). Bit-exact: the oracle restates the reference's float
arithmetic operation by operation."""
import os

import numpy as np
import pytest

import pyoracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, f"{name}.npz"))


def views_from(arr):
    return (pyoracle.WapView * len(arr)).from_buffer_copy(arr.tobytes())


@pytest.mark.parametrize("name", ["g1a", "g1b"])
def test_projector_golden(oracle, name):
    g = load(name)
    views = views_from(g["views"])
    nu, nv, sd = int(g["nu"]), int(g["nv"]), np.float32(g["sd"])
    for k in range(len(views)):
        p, r, w = oracle.forward_project(g["vol"], views[k], nu, nv, sd)
        np.testing.assert_array_equal(p, g["proj"][k, 0])
        np.testing.assert_array_equal(r, g["proj"][k, 1])
        np.testing.assert_array_equal(w, g["proj"][k, 2])
    # some rays miss the volume, some hit
    assert (g["proj"][:, 2] == 0).any() and (g["proj"][:, 2] > 0).any()


def test_geometry_golden_views_rebuild(oracle):
    """The oracle's geometry rebuilds the matrices stored with the projector goldens."""
    g = load("g1a")
    geo = pyoracle.make_geometry((16, 16, 16), 32, 24, sid=48.0, sdd=72.0)
    views = views_from(g["views"])
    for k, a in enumerate([0.0, 45.0, 90.0, 180.0, 211.7]):
        assert bytes(oracle.view(geo, a)) == bytes(views[k])


def test_cast_and_scatter_golden(oracle):
    g = load("g2")
    shape = g["vol"].shape
    geo = pyoracle.make_geometry((16, 15, 14), 24, 20, sid=40.0, sdd=60.0, parallel=True)
    v = oracle.view(geo, 33.0)
    z, u, tn, tf = oracle.produce_rays(shape, v, 24, 20)
    for a, b in [(z, g["zero4"]), (u, g["unit4"]), (tn, g["tn"]), (tf, g["tf"])]:
        np.testing.assert_array_equal(a, b)
    p, _, w = oracle.forward_project(g["vol"], v, 24, 20, np.float32(g["sd"]))
    np.testing.assert_array_equal(p.ravel(), g["val"])
    np.testing.assert_array_equal(w.ravel(), g["weight"])
    c, ws = oracle.backproject_ray(g["corr"].reshape(20, 24), shape, v, np.float32(g["sd"]))
    np.testing.assert_array_equal(c, g["contrib"])
    np.testing.assert_array_equal(ws, g["wsum"])


def test_advect_golden(oracle):
    g = load("g4")
    for i, dt in enumerate(g["dts"]):
        np.testing.assert_array_equal(oracle.advect(g["vel"], g["src"], float(dt)), g[f"out{i}"])
    np.testing.assert_array_equal(oracle.advect(g["vel"], g["src3"], -1.0, list(g["aabb"])),
                                  g["outnc3"])


def test_flow_single_golden(oracle):
    g = load("g5")
    I1, I2 = g["I1"], g["I2"]
    shape = I1.shape
    z3 = np.zeros((3,) + shape, np.float32)
    z1 = np.zeros(shape, np.float32)
    for i in range(4):
        case = g[f"case{i}"]
        V, Y1, Y2, Y3, Y4, _ = oracle.flow_single(I1, I2, z3, z3, z3, z3, z1, 0.008, 0.05,
                                                   int(case[0]), int(case[1]),
                                                   [0, 18, 0, 16, 0, 14], list(case[2:]))
        np.testing.assert_array_equal(V, g[f"V{i}"])
        np.testing.assert_array_equal(np.concatenate([Y1, Y2, Y3, Y4[None]]), g[f"Y{i}"])


def test_flow_multiscale_golden(oracle):
    g = load("g6")
    a = oracle.flow_multiscale(g["I1"], g["I2"], None, 0.008, 0.05, 20, 2, 0.5, 0.8, 3, [2] * 8)
    np.testing.assert_array_equal(a, g["V0"])
    b = oracle.flow_multiscale(g["I1"], g["I2"], g["Vinit"], 0.008, 0.05, 20, 2, 0.5, 0.8, 3,
                               [2] * 8)
    np.testing.assert_array_equal(b, g["V1"])


def test_image_ops_golden(oracle):
    g = load("g7")
    np.testing.assert_array_equal(oracle.blur(g["img"], 0.8), g["blur08"])
    np.testing.assert_array_equal(oracle.blur(g["img"], 2.0), g["blur2"])
    np.testing.assert_array_equal(oracle.gradient(g["img"]), g["grad"])
    for i in range(4):
        np.testing.assert_array_equal(oracle.resize(g["img3"], tuple(g[f"nd{i}"])),
                                      g[f"resize{i}"])
