"""Build recipe for the gfx950 library and the drop-in executables.

Everything is compiled in-tree with hipcc (no JIT caches), so the built
artefacts travel with the source tree:

  threepoint_amd/_lib/libthreepoint_amd.so      HIP kernels + C ABI + cubature
  threepoint_amd/_lib/calculateApertureStatistics.x
  threepoint_amd/_lib/calculateGamma.x
  threepoint_amd/_lib/calculatePowerspectrum.x
  threepoint_amd/_lib/calculateSecondOrderApertureStatistics.x
  threepoint_amd/_lib/calculateApertureStatisticsCovariance.x
  threepoint_amd/_lib/benchMap3.x                the batched <Map^3> engine from C++ (for profilers)
  threepoint_amd/_lib/threepoint_cosmosis.so    the cosmosis module
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libthreepoint_amd.so")

HIP_SOURCES = ["tp_tables.hip", "tp_map3.hip", "tp_runtime.hip", "tp_gamma.hip", "tp_mlp.hip", "tp_power.hip"]
CPP_SOURCES = ["tp_cubature.cpp", "tp_census.cpp", "tp_classes.cpp", "tp_io.cpp"]
APPS = ["calculateApertureStatistics", "calculateGamma", "calculatePowerspectrum",
        "calculateSecondOrderApertureStatistics", "calculateApertureStatisticsCovariance", "benchMap3"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-fopenmp",
    "-Wno-unused-result",
    # packed FP32 instructions stay off (v_pk_* FP32 forms are avoided on this pool)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        ln for ln in proc.stdout.splitlines()
        if "packed-fp32-ops" not in ln and "is not a recognized feature" not in ln
    )
    if proc.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    return out


def build_library(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES + CPP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared"] + COMMON + srcs + ["-o", tmp]
    out = _run(cmd)
    if verbose and out.strip():
        print(out)
    os.replace(tmp, LIB)
    return LIB


def build_apps(verbose=False):
    built = []
    for app in APPS:
        src = os.path.join(CSRC, "apps", app + ".cpp")
        if not os.path.exists(src):
            continue
        exe = os.path.join(LIBDIR, app + ".x")
        cmd = [HIPCC] + COMMON + [src, "-L" + LIBDIR, "-lthreepoint_amd", "-Wl,-rpath,$ORIGIN", "-o", exe]
        _run(cmd)
        built.append(exe)
    return built


COSMOSIS_MODULE = os.path.join(LIBDIR, "threepoint_cosmosis.so")


def build_cosmosis_module(verbose=False):
    """The cosmosis module (setup/execute/cleanup); cosmosis provides the datablock symbols at load."""
    src = os.path.join(CSRC, "cosmosis", "threepoint_module.cpp")
    cmd = [HIPCC, "-shared"] + COMMON + [src, "-L" + LIBDIR, "-lthreepoint_amd", "-Wl,-rpath,$ORIGIN",
                                          "-o", COSMOSIS_MODULE]
    _run(cmd)
    return COSMOSIS_MODULE


def build(verbose=False):
    lib = build_library(verbose)
    apps = build_apps(verbose)
    return [lib] + apps + [build_cosmosis_module(verbose)]


if __name__ == "__main__":
    for p in build(verbose=True):
        print("built", p)
