"""Build libvision_amd.so (host C++ + HIP kernels for gfx950) in-tree.

Host sources are compiled with g++, HIP sources with hipcc for gfx950 only.  Object
files go to vision_amd/_build/, the shared library to vision_amd/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import tempfile
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libvision_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-I", str(CSRC)]
# packed-fp32 instruction selection is disabled for the device code.  FP
# contraction is off so the kernels round like the scalar host arithmetic;
# the traversal's box tests use explicit fmaf where they want fused ops.
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics", "-ffp-contract=off"]


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", str(src), "-o", str(obj)]
    else:
        cmd = [CXX, *COMMON, "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


DATA = PKG / "data"
RGB2SPEC = DATA / "rgb2spec_srgb.f32"


def build_rgb2spec(verbose: bool = False) -> Path:
    """Generate the sRGB -> sigmoid-spectrum table (gen/rgb2spec.cpp) unless
    it is present and newer than its generator and the CIE data."""
    src = PKG / "gen" / "rgb2spec.cpp"
    cie = DATA / "cie_xyz_d65.f32"
    if RGB2SPEC.exists() and RGB2SPEC.stat().st_mtime > max(src.stat().st_mtime, cie.stat().st_mtime):
        return RGB2SPEC
    BUILD.mkdir(exist_ok=True)
    exe = BUILD / "rgb2spec"
    res = subprocess.run([CXX, "-O2", "-std=c++17", "-o", str(exe), str(src), "-lpthread"],
                         capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"rgb2spec compile failed\n{res.stderr}")
    tmp = RGB2SPEC.with_suffix(".tmp")
    res = subprocess.run([str(exe), str(cie), str(tmp), "64"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"rgb2spec generation failed\n{res.stderr}")
    os.replace(tmp, RGB2SPEC)
    if verbose:
        print(f"generated {RGB2SPEC}")
    return RGB2SPEC


LLVM_BIN = Path(os.environ.get("ROCM_PATH", "/opt/rocm")) / "llvm" / "bin"


def device_functions(obj: Path) -> list[tuple[str, int]]:
    """Non-kernel functions (name, bytes) in the gfx950 code object of a
    compiled HIP object.

    Every device function is meant to be inlined into its kernels.  A
    function the compiler keeps out of line returns through s[30:31]; when
    it is longer than the short-branch range (+-128 KiB) the branch
    relaxation of this toolchain builds its long jumps in s[30:31] too, the
    return address is lost and the kernel jumps astray (a GPU memory fault),
    so the build refuses any out-of-line device function."""
    with tempfile.TemporaryDirectory() as td:
        fb, co = Path(td) / "fatbin", Path(td) / "gfx950.co"
        r = subprocess.run([str(LLVM_BIN / "llvm-objcopy"), f"--dump-section=.hip_fatbin={fb}", str(obj),
                            str(Path(td) / "host.o")], capture_output=True, text=True)
        if r.returncode != 0 or not fb.exists():
            raise RuntimeError(f"no device code in {obj}: {r.stderr}")
        r = subprocess.run([str(LLVM_BIN / "clang-offload-bundler"), "--unbundle", "--type=o", f"--input={fb}",
                            "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"no gfx950 code object in {obj}: {r.stderr}")
        r = subprocess.run([str(LLVM_BIN / "llvm-readelf"), "-s", "-W", str(co)], capture_output=True, text=True,
                           check=True)
    funcs, kernels = {}, set()
    for line in r.stdout.splitlines():
        f = line.split()
        if len(f) >= 8 and f[3] == "FUNC":
            funcs[f[7]] = int(f[2])
        elif len(f) >= 8 and f[7].endswith(".kd"):
            kernels.add(f[7][:-3])
    return sorted((n, sz) for n, sz in funcs.items() if n not in kernels)


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    build_rgb2spec(verbose)
    sources = sorted(CSRC.glob("*.cpp")) + sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    for o in objs:
        if o.name.endswith(".hip.o"):
            calls = device_functions(o)
            if calls:
                raise RuntimeError(f"{o.name}: device functions kept out of line (mark them __forceinline__): "
                                   + ", ".join(f"{n} ({sz} B)" for n, sz in calls))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "--offload-arch=gfx950", "-fPIC", "-o", str(tmp), *map(str, objs),
           "-Wl,--no-undefined", "-lpthread", "-lz"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
