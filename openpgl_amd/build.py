"""Builds openpgl_amd/libopenpgl_amd.so (HIP kernels for gfx950 + host C ABI)
in-tree with hipcc.  `python -m openpgl_amd.build` or build() from Python."""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libopenpgl_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PGL_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-Wno-unused-but-set-variable", "-I" + CSRC]
# device code is built without packed-FP32 instructions (v_pk_*_f32)
HIPFLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]
SOURCES = [
    ("kernels_spatial.hip", HIPFLAGS),
    ("kernels_region.hip", HIPFLAGS),
    # the region kernel built a second time for two workgroups per CU (k_region_update_wide)
    ("kernels_region.hip", HIPFLAGS + ["-DPGL_REGION_OCC=2"], "kernels_region_wide.o"),
    ("kernels_query.hip", HIPFLAGS),
    ("kernels_sort.hip", HIPFLAGS),
    ("kernels_path.hip", HIPFLAGS),
    ("field.cpp", ["-x", "hip", "--offload-arch=" + ARCH, "-ffp-contract=off", "-Xclang", "-target-feature",
                   "-Xclang", "-packed-fp32-ops"]),
    ("region_sched.cpp", ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops"]),
    ("region_batch.cpp", ["-x", "hip", "--offload-arch=" + ARCH, "-ffp-contract=off", "-Xclang", "-target-feature",
                          "-Xclang", "-packed-fp32-ops"]),
    ("field_io.cpp", ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
                      "-packed-fp32-ops"]),
    ("path_batch.cpp", ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops"]),
    ("api_paths.cpp", ["-x", "hip", "--offload-arch=" + ARCH, "-ffp-contract=off", "-Xclang", "-target-feature",
                       "-Xclang", "-packed-fp32-ops"]),
    ("server.cpp", ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]),
    ("api.cpp", ["-x", "hip", "--offload-arch=" + ARCH, "-ffp-contract=off", "-Xclang", "-target-feature",
                 "-Xclang", "-packed-fp32-ops"]),
]


def _compile(src, extra, objName=None):
    obj = os.path.join(BUILD, objName or os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda a: _compile(*a), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
