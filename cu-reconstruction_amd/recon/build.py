"""Builds the native libraries of the TSDF fusion path in-tree.

Outputs (all under cu-reconstruction_amd/lib/):
  librecon.so        HIP kernels for gfx950 + the C-ABI of include/recon/recon.h
  librecon_host.so   C++ drop-in mirror of the reference's TSDF API over the C-ABI
  librecon_scene.so  synthetic depth renderer (host C++)
  tsdf_wrapper_test  C++ program exercising librecon_host.so (run by a GPU test)
and oracle/lib/liboracle.so through oracle/Makefile (the CPU checker).

Every call compiles from scratch.  Device code is built with IEEE float32
semantics: -ffp-contract=off (no FMA contraction), default correctly-rounded
division and square root, no fast-math, denormals kept; packed-FP32
instructions are disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parents[1]
ROOT = PKG.parent
LIB = PKG / "lib"
OBJ = PKG / "_build"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{INCLUDE}", f"-I{PKG / 'csrc'}",
]
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", f"-I{INCLUDE}"]

# host-side pieces of the drop-in beside tsdf_volume.cpp (librecon_host.so)
HOST_SOURCES = ["config.cpp", "png_reader.cpp", "tsdf_surface.cpp", "driver_capi.cpp"]
KERNEL_SOURCES = ["recon_common.hip", "tsdf_volume.hip", "marching_cubes.hip", "depth_points.hip",
                  "tsdf_batch.hip", "tsdf_diff.hip", "diff_batch.hip", "render_points.hip", "mc_weld.hip", "color.hip", "sparse.hip"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(str(c) for c in cmd)}")
    if verbose and (r.stdout or r.stderr):
        # hipcc's host pass warns that it does not know packed-fp32-ops
        lines = [l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l]
        if lines:
            print("\n".join(lines), flush=True)


def _kernel_sources() -> list[Path]:
    return [PKG / "csrc" / s for s in KERNEL_SOURCES]


def build(verbose: bool = False, jobs: int = 8) -> dict[str, Path]:
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    LIB.mkdir(exist_ok=True)
    OBJ.mkdir(exist_ok=True)
    srcs = _kernel_sources()
    objs = [OBJ / (s.stem + ".o") for s in srcs]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, [HIPCC, *HIP_FLAGS, "-fvisibility=hidden", "-c", s, "-o", o],
                          verbose)
                for s, o in zip(srcs, objs)]
        # host-side libraries do not depend on the kernels' objects
        scene_so = LIB / "librecon_scene.so"
        futs.append(ex.submit(_run, [CXX, *CXX_FLAGS, "-fopenmp", "-shared",
                                     PKG / "scene" / "scene.cpp", "-o", scene_so], verbose))
        futs.append(ex.submit(_run, ["make", "-s", "-B", "-C", ROOT / "oracle"], verbose))
        for f in futs:
            f.result()
    recon_so = LIB / "librecon.so"
    _run([HIPCC, *HIP_FLAGS, "-shared", *objs, "-o", recon_so], verbose)
    out = {"recon": recon_so, "scene": scene_so, "oracle": ROOT / "oracle" / "lib" / "liboracle.so"}
    host_src = PKG / "host" / "tsdf_volume.cpp"
    if host_src.exists():
        host_so = LIB / "librecon_host.so"
        host_srcs = [host_src] + [PKG / "host" / f for f in HOST_SOURCES
                                  if (PKG / "host" / f).exists()]
        _run([HIPCC, *HIP_FLAGS, "-shared", *host_srcs, f"-L{LIB}", "-lrecon", "-lz",
              "-Wl,-rpath,$ORIGIN", "-o", host_so], verbose)
        out["host"] = host_so
        main_src = PKG / "host" / "tsdf_surface_main.cpp"
        if main_src.exists():
            exe = LIB / "tsdf_surface"
            _run([HIPCC, *HIP_FLAGS, main_src, f"-L{LIB}", "-lrecon_host", "-lrecon",
                  "-Wl,-rpath,$ORIGIN", "-o", exe], verbose)
            out["tsdf_surface"] = exe
        test_src = ROOT / "tests" / "cpp" / "tsdf_wrapper_test.cpp"
        if test_src.exists():
            exe = LIB / "tsdf_wrapper_test"
            oracle_lib = ROOT / "oracle" / "lib"
            _run([HIPCC, *HIP_FLAGS, test_src, f"-I{PKG / 'scene'}", f"-I{ROOT / 'oracle'}",
                  f"-L{LIB}", f"-L{oracle_lib}", "-lrecon_host", "-lrecon", "-lrecon_scene",
                  "-loracle", "-Wl,-rpath,$ORIGIN:$ORIGIN/../../oracle/lib", "-o", exe], verbose)
            out["wrapper_test"] = exe
    return out


if __name__ == "__main__":
    for k, v in build(verbose="-v" in sys.argv).items():
        print(f"{k}: {v}")
