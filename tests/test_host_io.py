"""File formats and the command line of the reconstruction driver (no GPU needed).

* wap_read_image reads the libtiff test images the reference ships (Shared/tiff-4.0.4/
  test/images, copied under tests/golden/tiff) to the same pixels as their PGM/PPM twins;
  Analyze, NIfTI, MetaImage and raw files built here byte by byte from the published
  layouts read back exactly, in both byte orders and every datatype;
* wap_write_analyze writes the header fields CImg::save_analyze writes
  (Project/CImg.h:51715-51782), checked field by field against the Analyze 7.5 layout;
* wap_read_angles reads the angle list as Project/WaP_Tomo.cpp:324-331 does;
* wap_cli_parse has the reference table's defaults, flags and checks
  (Project/main.cpp:15-176), and the wap_tomo driver prints help and rejects bad lines.
"""
import ctypes as C
import os
import struct
import subprocess

import numpy as np
import pytest

import wap_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TIFF = os.path.join(ROOT, "tests", "golden", "tiff")
DRIVER = os.path.join(ROOT, "warp-and-project-tomography_amd", "wap_tomo")


def pnm(path):
    """Pixels of a binary PGM (P5) or PPM (P6) file as (C, H, W)."""
    b = open(path, "rb").read()
    fields, pos = [], 0
    while len(fields) < 4:
        while b[pos:pos + 1].isspace():
            pos += 1
        end = pos
        while not b[end:end + 1].isspace():
            end += 1
        fields.append(b[pos:end])
        pos = end
    pos += 1
    w, h = int(fields[1]), int(fields[2])
    c = 1 if fields[0] == b"P5" else 3
    a = np.frombuffer(b[pos:pos + w * h * c], np.uint8).reshape(h, w, c)
    return a.transpose(2, 0, 1).astype(np.float32)


# ---------------------------------------------------------------- TIFF
def test_tiff_reads_reference_test_images_like_their_pnm_twins():
    g = wap_amd.read_image(os.path.join(TIFF, "minisblack-1c-8b.tiff"))
    assert g.shape == (1, 1, 151, 157)
    np.testing.assert_array_equal(g[:, 0], pnm(os.path.join(TIFF, "minisblack-1c-8b.pgm")))
    rgb = wap_amd.read_image(os.path.join(TIFF, "rgb-3c-8b.tiff"))
    assert rgb.shape == (3, 1, 151, 157)
    np.testing.assert_array_equal(rgb[:, 0], pnm(os.path.join(TIFF, "rgb-3c-8b.ppm")))
    # the 16-bit version of the same picture (big-endian samples): within one 8-bit step
    g16 = wap_amd.read_image(os.path.join(TIFF, "minisblack-1c-16b.tiff"))
    assert g16.max() > 255
    assert np.abs(g16 / 257.0 - g).max() <= 1.0


@pytest.mark.parametrize("name", ["palette-1c-8b.tiff", "minisblack-2c-8b-alpha.tiff"])
def test_tiff_refuses_palette_and_compressed_images(name):
    with pytest.raises(OSError):
        wap_amd.read_image(os.path.join(TIFF, name))


def write_tiff(path, pages, order="<", fmt=3, rows_per_strip=None):
    """Baseline TIFF writer for tests: pages is (D, H, W) or (D, H, W, S) numpy data."""
    pages = np.asarray(pages)
    if pages.ndim == 3:
        pages = pages[..., None]
    d, h, w, spp = pages.shape
    bits = pages.dtype.itemsize * 8
    rps = rows_per_strip or h
    out = bytearray((b"II" if order == "<" else b"MM") + struct.pack(order + "HI", 42, 0))
    prev_link = 4
    for z in range(d):
        data = pages[z].astype(pages.dtype.newbyteorder(order)).tobytes()
        row = w * spp * pages.dtype.itemsize
        strips = [data[r * row:min(h, r + rps) * row] for r in range(0, h, rps)]
        offs = []
        for s in strips:
            offs.append(len(out))
            out += s
        arr_at = len(out)
        out += struct.pack(order + "%dI" % len(offs), *offs)
        cnt_at = len(out)
        out += struct.pack(order + "%dI" % len(strips), *[len(s) for s in strips])
        if len(out) % 2:
            out += b"\0"
        ifd = len(out)
        struct.pack_into(order + "I", out, prev_link, ifd)
        entries = [(256, 4, 1, w), (257, 4, 1, h), (258, 3, 1, bits), (259, 3, 1, 1),
                   (262, 3, 1, 1 if spp == 1 else 2),
                   (273, 4, len(offs), offs[0] if len(offs) == 1 else arr_at),
                   (277, 3, 1, spp), (278, 4, 1, rps),
                   (279, 4, len(strips), len(strips[0]) if len(strips) == 1 else cnt_at),
                   (339, 3, 1, fmt)]
        out += struct.pack(order + "H", len(entries))
        for tag, typ, n, v in entries:
            if typ == 3 and n == 1:
                out += struct.pack(order + "HHIHH", tag, typ, n, v, 0)
            else:
                out += struct.pack(order + "HHII", tag, typ, n, v)
        prev_link = len(out)
        out += struct.pack(order + "I", 0)
    open(path, "wb").write(bytes(out))


@pytest.mark.parametrize("order", ["<", ">"])
@pytest.mark.parametrize("dtype,fmt", [(np.float32, 3), (np.uint16, 1), (np.int16, 2),
                                       (np.uint8, 1), (np.int32, 2), (np.float64, 3)])
def test_tiff_multipage_stack(tmp_path, order, dtype, fmt):
    rng = np.random.default_rng(3)
    stack = (rng.random((5, 7, 9)) * 100 - (0 if fmt == 1 else 50)).astype(dtype)
    p = str(tmp_path / "s.tif")
    write_tiff(p, stack, order, fmt, rows_per_strip=3)
    got = wap_amd.read_image(p)
    assert got.shape == (1, 5, 7, 9)
    np.testing.assert_array_equal(got[0], stack.astype(np.float32))


def test_tiff_samples_become_channels(tmp_path):
    rng = np.random.default_rng(4)
    px = rng.integers(0, 255, (2, 4, 6, 3)).astype(np.uint8)
    p = str(tmp_path / "rgb.tiff")
    write_tiff(p, px, "<", 1)
    got = wap_amd.read_image(p)
    assert got.shape == (3, 2, 4, 6)
    np.testing.assert_array_equal(got, px.transpose(3, 0, 1, 2).astype(np.float32))


# ---------------------------------------------------------------- Analyze / NIfTI
def analyze_header(dims4, datatype, bitpix, order="<", vox_offset=0.0, scale=1.0, size=348):
    h = bytearray(size)
    struct.pack_into(order + "i", h, 0, 348)
    struct.pack_into(order + "8h", h, 40, 4, *dims4, 0, 0, 0)
    struct.pack_into(order + "hh", h, 70, datatype, bitpix)
    struct.pack_into(order + "4f", h, 76, 0, 1, 1, 1)
    struct.pack_into(order + "ff", h, 108, vox_offset, scale)
    return bytes(h)


def test_write_analyze_header_fields(tmp_path):
    rng = np.random.default_rng(5)
    vol = rng.random((2, 3, 4, 5)).astype(np.float32)  # C, D, H, W
    p = str(tmp_path / "v.hdr")
    wap_amd.write_analyze(p, vol, voxel_size=(0.5, 0.25, 2.0))
    h = open(p, "rb").read()
    assert len(h) == 348
    assert struct.unpack_from("<i", h, 0)[0] == 348        # sizeof_hdr
    assert h[4:14] == b"CImg" + bytes(6)                      # data_type
    assert h[14:32] == b" " + bytes(17)                       # db_name
    assert struct.unpack_from("<h", h, 36)[0] == 4096       # session_error as written
    assert h[38] == ord("r")                                  # regular
    assert struct.unpack_from("<8h", h, 40) == (4, 5, 4, 3, 2, 0, 0, 0)
    assert struct.unpack_from("<hh", h, 70) == (16, 4)        # float; bitpix field = bytes
    assert struct.unpack_from("<4f", h, 76) == (0.0, 0.5, 0.25, 2.0)
    assert struct.unpack_from("<ff", h, 108) == (0.0, 1.0)    # vox_offset, scale
    assert set(h[116:]) <= {0} and set(h[39:40]) == {0}
    img = open(str(tmp_path / "v.img"), "rb").read()
    assert img == vol.tobytes()
    np.testing.assert_array_equal(wap_amd.read_image(p), vol)
    np.testing.assert_array_equal(wap_amd.read_image(str(tmp_path / "v.img")), vol)


def test_write_analyze_name_forms(tmp_path):
    vol = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    wap_amd.write_analyze(str(tmp_path / "plain"), vol)
    assert os.path.exists(tmp_path / "plain.hdr") and os.path.exists(tmp_path / "plain.img")
    wap_amd.write_analyze(str(tmp_path / "b.img"), vol)
    assert os.path.getsize(tmp_path / "b.hdr") == 348
    wap_amd.write_analyze(str(tmp_path / "n.nii"), vol)
    b = open(tmp_path / "n.nii", "rb").read()
    assert len(b) == 352 + vol.nbytes
    assert struct.unpack_from("<f", b, 108)[0] == 352.0 and b[352:] == vol.tobytes()
    np.testing.assert_array_equal(wap_amd.read_image(str(tmp_path / "n.nii"))[0], vol)
    # pixdim defaults to 1 1 1, as the reference's saves (no voxel size passed)
    h = open(tmp_path / "plain.hdr", "rb").read()
    assert struct.unpack_from("<4f", h, 76) == (0.0, 1.0, 1.0, 1.0)
    # an empty image leaves an empty file; unknown extensions and huge dims are refused
    lib = wap_amd.load()
    assert lib.wap_write_analyze(str(tmp_path / "e.hdr").encode(), wap_amd.ints(0, 3, 3, 1),
                                 None, None) == 0
    assert os.path.getsize(tmp_path / "e.hdr") == 0
    assert lib.wap_write_analyze(str(tmp_path / "x.foo").encode(), wap_amd.ints(1, 1, 1, 1),
                                 wap_amd.floats(1.0), None) == -1
    assert lib.wap_write_analyze(str(tmp_path / "x.hdr").encode(), wap_amd.ints(40000, 1, 1, 1),
                                 wap_amd.floats(1.0), None) == -1


@pytest.mark.parametrize("order", ["<", ">"])
@pytest.mark.parametrize("dtype,code", [(np.uint8, 2), (np.int16, 4), (np.int32, 8),
                                        (np.float32, 16), (np.float64, 64)])
def test_read_analyze_datatypes_and_byte_order(tmp_path, order, dtype, code):
    rng = np.random.default_rng(code)
    vol = (rng.random((3, 4, 5)) * 200 - (0 if dtype == np.uint8 else 100)).astype(dtype)
    scale = 0.5
    open(tmp_path / "a.hdr", "wb").write(
        analyze_header((5, 4, 3, 1), code, vol.dtype.itemsize * 8, order, scale=scale))
    open(tmp_path / "a.img", "wb").write(vol.astype(vol.dtype.newbyteorder(order)).tobytes())
    got = wap_amd.read_image(str(tmp_path / "a.hdr"))
    assert got.shape == (1, 3, 4, 5)
    if dtype == np.float64:
        want = (vol * scale).astype(np.float32)
    else:
        want = vol.astype(np.float32) * np.float32(scale)
    np.testing.assert_array_equal(got[0], want)


def test_read_nifti_single_file_and_zero_scale(tmp_path):
    vol = np.arange(60, dtype=np.float32).reshape(3, 4, 5) - 7
    hdr = analyze_header((5, 4, 3, 1), 16, 32, "<", vox_offset=352.0, scale=0.0, size=352)
    open(tmp_path / "v.nii", "wb").write(hdr + vol.tobytes())
    np.testing.assert_array_equal(wap_amd.read_image(str(tmp_path / "v.nii"))[0], vol)


def test_read_analyze_errors(tmp_path):
    with pytest.raises(OSError):
        wap_amd.read_image(str(tmp_path / "missing.hdr"))
    open(tmp_path / "t.hdr", "wb").write(analyze_header((5, 4, 3, 1), 16, 32))
    open(tmp_path / "t.img", "wb").write(bytes(10))
    with pytest.raises(OSError):
        wap_amd.read_image(str(tmp_path / "t.hdr"))
    open(tmp_path / "u.hdr", "wb").write(analyze_header((5, 4, 3, 1), 32, 64))
    with pytest.raises(OSError):
        wap_amd.read_image(str(tmp_path / "u.hdr"))
    # a buffer smaller than the image is refused
    lib = wap_amd.load()
    wap_amd.write_analyze(str(tmp_path / "s.hdr"), np.ones((2, 2, 2), np.float32))
    small = (C.c_float * 7)()
    assert lib.wap_read_image(str(tmp_path / "s.hdr").encode(), wap_amd.ints(0, 0, 0, 0),
                              small, 7) == -1


# ---------------------------------------------------------------- MetaImage, raw
@pytest.mark.parametrize("local", [True, False])
def test_read_metaimage(tmp_path, local):
    rng = np.random.default_rng(9)
    vol = rng.integers(-300, 300, (3, 4, 5)).astype(">i2")
    head = ["ObjectType = Image", "NDims = 3", "DimSize = 5 4 3", "ElementType = MET_SHORT",
            "BinaryDataByteOrderMSB = True", "CompressedData = False"]
    if local:
        head.append("ElementDataFile = LOCAL")
        open(tmp_path / "v.mha", "wb").write(("\n".join(head) + "\n").encode() + vol.tobytes())
        p = tmp_path / "v.mha"
    else:
        head.append("ElementDataFile = v.raw")
        open(tmp_path / "v.mhd", "w").write("\n".join(head) + "\n")
        open(tmp_path / "v.raw", "wb").write(vol.tobytes())
        p = tmp_path / "v.mhd"
    np.testing.assert_array_equal(wap_amd.read_image(str(p))[0], vol.astype(np.float32))


def test_read_raw_float32(tmp_path):
    vol = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)
    vol.tofile(tmp_path / "p.f32")
    got = wap_amd.read_image(str(tmp_path / "p.f32"), dims=(4, 3, 2))
    np.testing.assert_array_equal(got[0], vol)
    with pytest.raises(OSError):
        wap_amd.read_image(str(tmp_path / "p.f32"))            # no dims
    with pytest.raises(OSError):
        wap_amd.read_image(str(tmp_path / "p.f32"), dims=(4, 3, 3))   # wrong size


# ---------------------------------------------------------------- angle list
def test_read_angles(tmp_path):
    p = tmp_path / "ours_foam_compress.txt"
    p.write_text("0 1.5\n  3.25\t-4e1\n7\n")
    got, n = wap_amd.read_angles(str(p), 4)
    assert got == [0.0, 1.5, 3.25, -40.0] and n == 4
    got, n = wap_amd.read_angles(str(p), 8)
    assert got == [0.0, 1.5, 3.25, -40.0, 7.0, 0.0, 0.0, 0.0] and n == 5
    p.write_text("10 20 x 30")
    got, n = wap_amd.read_angles(str(p), 4)
    assert got == [10.0, 20.0, 0.0, 0.0] and n == 2
    with pytest.raises(OSError):
        wap_amd.read_angles(str(tmp_path / "none.txt"), 3)


# ---------------------------------------------------------------- command line
REQUIRED = ["-f", "64", "-u", "96", "-i", "proj.hdr", "--oxyz", "0", "--sdg", "0"]


def test_cli_config_layout_matches_c(tmp_path):
    src = tmp_path / "l.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "wap_types.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(WapCliConfig),'
                   ' offsetof(WapCliConfig, proj_files), offsetof(WapCliConfig, bp),'
                   ' offsetof(WapCliConfig, lds), offsetof(WapCliConfig, device));return 0;}\n')
    exe = tmp_path / "l"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)],
                   check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    T = wap_amd.WapCliConfig
    assert got == [C.sizeof(T), T.proj_files.offset, T.bp.offset, T.lds.offset, T.device.offset]


def test_cli_defaults_are_the_reference_drivers():
    r, c, m = wap_amd.cli_parse(REQUIRED)
    assert r == 0, m
    # Project/main.cpp:95-141 defaults
    assert (c.sigma, c.tau, c.lambda_, c.alpha) == (0.3, 0.3, 0.1, 0.1)
    assert (c.voxel_spacing, c.det_spacing, c.sid, c.sdd) == (1.0, 1.0, 400.0, 200.0)
    assert (c.algo_iters, c.sart_iters, c.nframes, c.nrounds) == (2, 2, 1, 1)
    assert c.output == b"Test/test" and c.prior == b"ATV" and c.bp == b"Voxelbased"
    assert c.total_proj == 192                      # fixed in the reference driver
    assert list(c.vol) == [64, 63, 63] and list(c.img) == [96, 512]
    assert c.n_proj_files == 1 and c.proj_files[0].value == b"proj.hdr"
    assert list(c.oxyz) == [0.0, 0.0, 0.0] and c.n_start_deg == 1
    assert c.save == 1 and c.use_lds == 0 and c.angles == b"ours_foam_compress.txt"


def test_cli_every_flag():
    args = ["-s", "0.5", "-t0.25", "--Lambda=2", "--Alpha", "0.05", "-o", "out/run",
            "-i", "a.tif", "--projsvolume=b.tif", "-f", "256", "--XYZ", "128", "-f", "0x40",
            "-u", "384", "--imgUV=0b1000", "-c", "0.75", "-d", "1.5", "-g", "1024", "-j", "1536",
            "--oxyz", "1", "--oxyz=-2", "--oxyz", "3.5", "--sdg", "10", "--sdg", "20",
            "-v", "12", "-r", "3", "-b", "7", "-p", "9", "--prior", "STV", "--bp=Raybased",
            "--nproj", "360", "--angles", "a.txt", "--lds", "10,3,50", "--save", "all",
            "--device", "0"]
    r, c, m = wap_amd.cli_parse(args)
    assert r == 0, m
    assert (c.sigma, c.tau, c.lambda_, c.alpha) == (0.5, 0.25, 2.0, 0.05)
    assert c.output == b"out/run"
    assert c.n_proj_files == 2 and [c.proj_files[i].value for i in range(2)] == [b"a.tif", b"b.tif"]
    assert list(c.vol) == [256, 128, 64] and list(c.img) == [384, 8]
    assert (c.voxel_spacing, c.det_spacing, c.sid, c.sdd) == (0.75, 1.5, 1024.0, 1536.0)
    assert list(c.oxyz) == [1.0, -2.0, 3.5]
    assert c.n_start_deg == 2 and list(c.start_deg[:2]) == [10.0, 20.0]
    assert (c.nframes, c.nrounds, c.algo_iters, c.sart_iters) == (12, 3, 7, 9)
    assert c.prior == b"STV" and c.bp == b"Raybased"
    assert c.total_proj == 360 and c.angles == b"a.txt"
    assert c.use_lds == 1 and list(c.lds) == [10, 3, 50] and c.save == 2


def test_cli_integer_syntax_and_prefixes():
    r, c, m = wap_amd.cli_parse(REQUIRED + ["-v", "0o17", "-b", "2KB", "--Sig=1.25", "--nfr", "5"])
    assert r != 0 and "excess option" in m          # --nfr resolves to -v (first "nframes")
    r, c, m = wap_amd.cli_parse(REQUIRED + ["-v", "0o17", "-b", "2KB", "--Sig=1.25"])
    assert r == 0, m
    assert (c.nframes, c.algo_iters, c.sigma) == (15, 2048, 1.25)
    r, c, m = wap_amd.cli_parse(REQUIRED + ["--nframes", "4"])
    assert r == 0 and c.nframes == 4 and c.nrounds == 1


@pytest.mark.parametrize("args,needle", [
    ([], "missing option -i|--projsvolume"),
    (["-f", "1", "-u", "1", "--oxyz", "0", "--sdg", "0"], "missing option -i"),
    (REQUIRED + ["-s", "abc"], 'invalid argument "abc" to option -s|--Sigma'),
    (REQUIRED + ["-b", "12x"], 'invalid argument "12x"'),
    (REQUIRED + ["-u", "1", "-u", "2"], "excess option -u"),
    (REQUIRED + ["stray"], 'unexpected argument "stray"'),
    (REQUIRED + ["--nosuch=1"], 'invalid option "--nosuch=1"'),
    (REQUIRED + ["-z"], "invalid option -z"),
    (REQUIRED + ["-s"], "option -s requires an argument"),
    (REQUIRED + ["--save", "some"], "expected none, final or all"),
    (REQUIRED + ["--lds", "10,1,5"], "expected Ns,p,Nz"),
    (REQUIRED + ["--s"], 'ambiguous option "--s"'),
])
def test_cli_errors(args, needle):
    r, c, m = wap_amd.cli_parse(args)
    assert r == -1
    assert needle in m
    assert m.rstrip().endswith("--help' for more information.")


def test_cli_help_and_version_take_precedence():
    r, c, m = wap_amd.cli_parse(["-s", "abc", "--help"])
    assert r == wap_amd.WAP_CLI_EXIT and m.startswith("Usage: wap_tomo")
    for flag in ["--Sigma", "--Tau", "--Lambda", "--Alpha", "-o", "--projsvolume", "--XYZ",
                 "--imgUV", "--voxelspacing", "-d", "--sid", "--sdd", "--oxyz", "--sdg",
                 "--nframes", "-r", "--AlgorithmIter", "--SartIter", "--prior", "--bp"]:
        assert flag in m
    r, c, m = wap_amd.cli_parse(["-x", "--version"])
    assert r == wap_amd.WAP_CLI_EXIT and "version" in m


def test_driver_command_line_without_gpu_work():
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-C", os.path.join(ROOT, "warp-and-project-tomography_amd",
                                                   "csrc")], check=True, capture_output=True)
    p = subprocess.run([DRIVER, "--help"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and "Usage:" in p.stdout
    p = subprocess.run([DRIVER, "-f", "8"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and "missing option --oxyz" in p.stdout
