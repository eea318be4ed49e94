"""Builds libexageostat_amd.so in-tree with hipcc for gfx950.

Every HIP source is compiled with -Xclang -target-feature -Xclang -packed-fp32-ops
(packed FP32 instructions are switched off for the device; the host pass warns
that it does not know the feature, which is harmless).
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libexageostat_amd.so")
# C++ host API (ExaGeoStat<T>, Configurations, Kernel<T> plugins) over the C ABI
API_SRC = os.path.join(HERE, "api", "ExaGeoStat.cpp")
API_LIB = os.path.join(HERE, "libexageostat_cpp.so")
CXX = os.environ.get("CXX", "g++")

SOURCES = ["egs_covgen.hip", "egs_chol.hip", "egs_update.hip", "egs_mixed.hip", "egs_solve.hip", "egs_capi.cpp",
           "egs_bobyqa.cpp", "egs_serve.cpp", "egs_devmem.cpp"]
ARCH = os.environ.get("EGS_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH, "-Wall", "-Wno-unused-result", "-Wno-unused-value",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errors = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, errors))
    return obj, errors


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    for _, msg in results:
        if verbose and msg.strip():
            print(msg, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    build_api()
    build_tests()
    return LIB


def build_api():
    """libexageostat_cpp.so: host-only C++, linked against the C ABI library."""
    tmp = API_LIB + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"),
           API_SRC, "-o", tmp, "-L" + HERE, "-lexageostat_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("C++ API build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, API_LIB)
    return API_LIB


def build_tests():
    """C++ test programs under tests/cpp (binaries in tests/cpp/_bin)."""
    src_dir = os.path.join(ROOT, "tests", "cpp")
    out_dir = os.path.join(src_dir, "_bin")
    os.makedirs(out_dir, exist_ok=True)
    outs = []
    for name in sorted(os.listdir(src_dir)):
        if not name.endswith(".cpp"):
            continue
        out = os.path.join(out_dir, name[:-4])
        cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"), os.path.join(src_dir, name),
               "-o", out + ".tmp", "-L" + HERE, "-lexageostat_cpp", "-lexageostat_amd", "-Wl,-rpath,$ORIGIN/../../../exageostatcpp_amd"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("test build failed for %s:\n%s" % (name, r.stdout + r.stderr))
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


if __name__ == "__main__":
    print(build(verbose=True))
