"""Builds the gfx950 HIP library in-tree: ambient_image_processor_amd/libambient_amd.so.

hipcc compiles every csrc/*.hip for --offload-arch=gfx950 (cross-compiles
without a GPU).  Flags:
  -ffp-contract=off   only explicit fma/fmaf calls fuse (the kernels spell out
                      the reference's multiply-add placement)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions in generated code
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD_DIR = os.path.join(ROOT, "_build")
LIB_NAME = "libambient_amd.so"
LIB_PATH = os.path.join(PKG, LIB_NAME)
SOURCES = ["tgv_sweeps.hip", "tgvk.hip", "dct.hip", "dct_mfma.hip", "energy.hip", "adjoint.hip",
           "resize.hip", "capi.hip"]
ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fvisibility=hidden", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")
    return obj


def build(verbose: bool = False) -> str:
    """Compile every HIP source for gfx950 and link libambient_amd.so (always rebuilds)."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB_PATH)
    _build_compat_demo()
    _build_cli()
    _build_div_rn_check()
    if verbose:
        print("built", LIB_PATH)
    return LIB_PATH


COMPAT_DEMO = os.path.join(BUILD_DIR, "processor_demo")


def _build_compat_demo() -> str:
    """A C++ caller that uses only the reference's launcher templates (through
    include/ambient_amd_reference_compat.hpp), linked against the library."""
    src = os.path.join(ROOT, "tests", "compat", "processor_demo.cpp")
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), src, "-o",
           COMPAT_DEMO, "-L" + PKG, "-lambient_amd", "-Wl,-rpath," + PKG,
           "-Wl,-rpath,$ORIGIN/../ambient_image_processor_amd", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compat demo build failed:\n" + r.stdout + r.stderr)
    return COMPAT_DEMO


DIV_RN_CHECK = os.path.join(BUILD_DIR, "div_rn_check")


def _build_div_rn_check() -> str:
    """The device check of the fused sweep's correctly rounded division
    (tests/compat/div_rn_check.hip, same flags as the library)."""
    src = os.path.join(ROOT, "tests", "compat", "div_rn_check.hip")
    cmd = [HIPCC] + COMMON_FLAGS + [src, "-o", DIV_RN_CHECK]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("division check build failed:\n" + r.stdout + r.stderr)
    return DIV_RN_CHECK


CLI = os.path.join(BUILD_DIR, "tgv2_deshade_application")
CLI_DOWNSAMPLED = os.path.join(BUILD_DIR, "tgvk_deshade_downsampled_application")
METAIMAGE_CONVERT = os.path.join(BUILD_DIR, "metaimage_convert")


def _build_cli() -> None:
    """The tgv2_deshade_application and tgvk_deshade_downsampled_application CLIs
    (reference argument contracts, MetaImage files) and the MetaImage converter,
    plain C++ (g++, zlib)."""
    cli = os.path.join(PKG, "cli")
    common = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + cli]
    jobs = [
        common + [os.path.join(cli, "tgv2_deshade_main.cpp"), "-o", CLI, "-L" + PKG,
                  "-lambient_amd", "-lz", "-Wl,-rpath," + PKG,
                  "-Wl,-rpath,$ORIGIN/../ambient_image_processor_amd", "-Wl,-rpath,/opt/rocm/lib"],
        common + [os.path.join(cli, "tgvk_deshade_downsampled_main.cpp"), "-o", CLI_DOWNSAMPLED,
                  "-L" + PKG, "-lambient_amd", "-lz", "-Wl,-rpath," + PKG,
                  "-Wl,-rpath,$ORIGIN/../ambient_image_processor_amd", "-Wl,-rpath,/opt/rocm/lib"],
        common + [os.path.join(cli, "metaimage_convert.cpp"), "-o", METAIMAGE_CONVERT, "-lz"],
    ]
    for cmd in jobs:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("CLI build failed:\n" + r.stdout + r.stderr)


if __name__ == "__main__":
    build(verbose=True)
