"""MetaImage (.mha / .mhd) reading and writing in numpy, the file format of the
deshade CLI (ambient_image_processor_amd/cli/tgv2_deshade_main.cpp; the reference
reads and writes it through ITK, itk/ITKImage.cpp:116-232).  Volumes are
returned as float32 arrays of shape (D, H, W); 2-D images as (1, H, W)."""
from __future__ import annotations

import os
import zlib

import numpy as np

_TYPES = {"MET_UCHAR": "u1", "MET_CHAR": "i1", "MET_USHORT": "u2", "MET_SHORT": "i2",
          "MET_UINT": "u4", "MET_INT": "i4", "MET_ULONG": "u8", "MET_LONG": "i8",
          "MET_ULONG_LONG": "u8", "MET_LONG_LONG": "i8", "MET_FLOAT": "f4", "MET_DOUBLE": "f8"}


def read(path):
    """Returns (volume float32 (D, H, W), dict(spacing, origin))."""
    with open(path, "rb") as fh:
        raw = fh.read()
    header, pos = {}, 0
    while True:
        end = raw.index(b"\n", pos)
        line = raw[pos:end].decode("latin-1")
        pos = end + 1
        if "=" not in line:
            continue
        k, v = (s.strip() for s in line.split("=", 1))
        header[k] = v
        if k == "ElementDataFile":
            break
    dims = [int(x) for x in header["DimSize"].split()]
    nd = int(header["NDims"])
    if nd not in (2, 3) or len(dims) != nd:
        raise ValueError("only 2-D and 3-D images are supported")
    if int(header.get("ElementNumberOfChannels", "1")) != 1:
        raise ValueError("only scalar images are supported")
    msb = header.get("BinaryDataByteOrderMSB", header.get("ElementByteOrderMSB", "False"))
    dt = np.dtype(_TYPES[header["ElementType"]]).newbyteorder(">" if msb.lower() == "true" else "<")
    if header["ElementDataFile"] == "LOCAL":
        payload = raw[pos:]
    else:
        with open(os.path.join(os.path.dirname(path), header["ElementDataFile"]), "rb") as fh:
            payload = fh.read()
    if header.get("CompressedData", "False").lower() == "true":
        payload = zlib.decompress(payload)
    w, h = dims[0], dims[1]
    d = dims[2] if nd == 3 else 1
    vol = np.frombuffer(payload, dtype=dt, count=w * h * d).astype(np.float32).reshape(d, h, w)
    spacing = [float(x) for x in header.get("ElementSpacing", "1 1 1").split()]
    origin = [float(x) for x in header.get("Offset", header.get("Origin", "0 0 0")).split()]
    return vol, dict(spacing=spacing, origin=origin)


def write(path, vol, spacing=None, origin=None, compress=True, element_type="MET_FLOAT",
          msb=False):
    """Writes a (D, H, W) volume; element_type/msb allow writing test inputs."""
    vol = np.asarray(vol)
    d, h, w = vol.shape
    nd = 3 if d > 1 else 2
    dt = np.dtype(_TYPES[element_type]).newbyteorder(">" if msb else "<")
    payload = vol.astype(dt).tobytes()
    if compress:
        payload = zlib.compress(payload)
    spacing = spacing or [1.0] * nd
    origin = origin or [0.0] * nd
    lines = ["ObjectType = Image", f"NDims = {nd}", "BinaryData = True",
             f"BinaryDataByteOrderMSB = {'True' if msb else 'False'}",
             f"CompressedData = {'True' if compress else 'False'}"]
    if compress:
        lines.append(f"CompressedDataSize = {len(payload)}")
    lines += ["Offset = " + " ".join(repr(float(x)) for x in origin[:nd]),
              "ElementSpacing = " + " ".join(repr(float(x)) for x in spacing[:nd]),
              "DimSize = " + " ".join(str(x) for x in ([w, h, d] if nd == 3 else [w, h])),
              f"ElementType = {element_type}"]
    if path.lower().endswith(".mhd"):
        data_name = os.path.basename(path)[:-4] + (".zraw" if compress else ".raw")
        lines.append(f"ElementDataFile = {data_name}")
        with open(path, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        with open(os.path.join(os.path.dirname(path), data_name), "wb") as fh:
            fh.write(payload)
    else:
        lines.append("ElementDataFile = LOCAL")
        with open(path, "wb") as fh:
            fh.write(("\n".join(lines) + "\n").encode("latin-1") + payload)
