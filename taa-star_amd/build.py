"""Build libtaastar.so (gfx950 kernels + host runtime) in-tree.

Usage: python taa-star_amd/build.py [--jobs N]
Outputs: taa-star_amd/taastar/libtaastar.so (objects under taa-star_amd/_build/).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "taastar", "libtaastar.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
INC = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]

# Device code: gfx950 only; -ffp-contract=off keeps the FP32 path in taa.comp's operation order;
# packed-fp32 instructions are disabled (compiler-generated v_pk_*_f32 are not used here).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
# Host code: no contraction either, so the scene generator is bit-stable across hosts.
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fopenmp", "-Wall",
             "-Wno-unused-result", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include")]

SOURCES = [
    ("taa_kernels.hip", "hip"),
    ("taa_fast.hip", "hip"),
    ("taa_velocity.hip", "hip"),
    ("taa_edge_aa.hip", "hip"),
    ("taa_noalloc.hip", "hip"),
    ("taa_api.cpp", "cxx"),
    ("taa_host.cpp", "cxx"),
    ("taa_scene.cpp", "cxx"),
    ("taa_serve.cpp", "cxx"),
    ("taa_settings.cpp", "cxx"),
]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise SystemExit("build failed: " + os.path.basename(cmd[-3] if len(cmd) > 3 else cmd[0]))
    return r


def compile_one(src_kind):
    src, kind = src_kind
    path = os.path.join(CSRC, src)
    obj = os.path.join(OUT_DIR, src + ".o")
    if kind == "hip":
        cmd = [HIPCC, *HIP_FLAGS, *INC, "-c", path, "-o", obj]
    else:
        cmd = ["g++", *CXX_FLAGS, *INC, "-c", path, "-o", obj]
    run(cmd)
    return obj


def build(jobs=8):
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s[0]))]
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs, "-lgomp"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    jobs = 8
    if "--jobs" in sys.argv:
        jobs = int(sys.argv[sys.argv.index("--jobs") + 1])
    print(build(jobs))
