"""CPU tests of the MetaImage boundary (C++ reader/writer used by the CLI, and
the numpy module) and of the CLI's argument contract (paths that exit before
any GPU work)."""
import os
import subprocess

import numpy as np
import pytest

from ambient_image_processor_amd import _build, metaimage


def _need(path):
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run __graft_entry__.build())")


@pytest.mark.parametrize("etype,msb", [("MET_UCHAR", False), ("MET_SHORT", True),
                                       ("MET_USHORT", False), ("MET_INT", True),
                                       ("MET_FLOAT", False), ("MET_FLOAT", True),
                                       ("MET_DOUBLE", False)])
@pytest.mark.parametrize("compress", [False, True])
@pytest.mark.parametrize("ext", [".mha", ".mhd"])
def test_cpp_reader_roundtrip(tmp_path, etype, msb, compress, ext):
    _need(_build.METAIMAGE_CONVERT)
    rng = np.random.default_rng(0)
    vol = rng.integers(0, 200, size=(4, 5, 6)).astype(np.float64)
    if etype in ("MET_FLOAT", "MET_DOUBLE"):
        vol = vol / 7.0
    src = str(tmp_path / ("in" + ext))
    metaimage.write(src, vol, spacing=[0.5, 0.7, 1.5], origin=[1.0, -2.0, 3.0],
                    compress=compress, element_type=etype, msb=msb)
    dst = str(tmp_path / ("out" + ext))
    r = subprocess.run([_build.METAIMAGE_CONVERT, src, dst], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got, meta = metaimage.read(dst)
    np.testing.assert_array_equal(got, vol.astype(np.float32))
    assert meta["spacing"] == [0.5, 0.7, 1.5] and meta["origin"] == [1.0, -2.0, 3.0]


def test_cpp_reader_extra_keys_and_2d(tmp_path):
    _need(_build.METAIMAGE_CONVERT)
    data = np.arange(12, dtype="<f4").reshape(3, 4)
    hdr = ("ObjectType = Image\nNDims = 2\nBinaryData = True\nBinaryDataByteOrderMSB = False\n"
           "CompressedData = False\nTransformMatrix = 1 0 0 1\nOffset = 0 0\n"
           "CenterOfRotation = 0 0\nAnatomicalOrientation = RA\nElementSpacing = 1 1\n"
           "DimSize = 4 3\nElementType = MET_FLOAT\nElementDataFile = LOCAL\n")
    src = tmp_path / "two_d.mha"
    src.write_bytes(hdr.encode() + data.tobytes())
    dst = str(tmp_path / "o.mha")
    assert subprocess.run([_build.METAIMAGE_CONVERT, str(src), dst]).returncode == 0
    got, _ = metaimage.read(dst)
    np.testing.assert_array_equal(got[0], data)


def test_cpp_reader_rejects_truncated(tmp_path):
    _need(_build.METAIMAGE_CONVERT)
    src = tmp_path / "t.mha"
    metaimage.write(str(src), np.ones((2, 3, 4)), compress=False)
    src.write_bytes(src.read_bytes()[:-8])
    r = subprocess.run([_build.METAIMAGE_CONVERT, str(src), str(tmp_path / "o.mha")],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "truncated" in r.stderr


def test_cli_usage_and_missing_files(tmp_path):
    """terminal_programs/tgv2_deshade_main.cpp:135-181: usage and exit codes 1/2/3."""
    _need(_build.CLI)
    r = subprocess.run([_build.CLI, "a", "1", "2"], capture_output=True, text=True)
    assert r.returncode == 1 and "input_image_path lambda alpha0 alpha1" in r.stdout
    args = ["0.5", "2", "1", "10", "-1"]
    outs = [str(tmp_path / n) for n in ("u.mha", "l.mha", "r.mha")]
    r = subprocess.run([_build.CLI, str(tmp_path / "missing.mha")] + args + outs,
                       capture_output=True, text=True)
    assert r.returncode == 2 and "input image file does not exist" in r.stderr
    src = str(tmp_path / "in.mha")
    metaimage.write(src, np.ones((2, 3, 4)))
    r = subprocess.run([_build.CLI, src] + args + [str(tmp_path / "nomask.mha")] + outs,
                       capture_output=True, text=True)
    assert r.returncode == 3 and "mask image file does not exist" in r.stderr


def test_inflate_in_chunks(tmp_path):
    """The reader's zlib loop feeds and drains at most UINT_MAX bytes per call
    (payloads of 4 GiB and more); compiled here with chunks of 1, 7 and 4096
    bytes, it must inflate exactly and reject truncated or mis-sized data."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "inflate_check")
    src = os.path.join(root, "tests", "compat", "metaimage_inflate_check.cpp")
    inc = os.path.join(root, "ambient_image_processor_amd", "cli")
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-I" + inc, src, "-o", exe, "-lz"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
