"""Build recipe for the gfx950 library and its Python binding (in-tree).

Produces, next to this file:
  libpgeof_hip.so                      HIP kernels + host runtime + C ABI (hipcc)
  pgeof_ext.<python-ext-suffix>         pybind11 module over the C ABI (g++)

Run ``python -m point_geometric_features_amd._build`` or call ``build()``.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("PGEOF_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["pgeof_search.hip", "pgeof_features.hip", "pgeof_noalloc.hip", "pgeof_capi.cpp", "pgeof_serve.cpp"]

# packed FP32 instructions are kept out of the device code on purpose
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unknown-warning-option",
]

LIB = os.path.join(HERE, "libpgeof_hip.so")


def ext_path() -> str:
    return os.path.join(HERE, "pgeof_ext" + sysconfig.get_config_var("EXT_SUFFIX"))


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose: bool = False) -> None:
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmds.append([HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(len(cmds), 8)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)

    import pybind11  # noqa: WPS433 (build-time dependency)

    inc = [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}"]
    tmp = ext_path() + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden", *inc,
          os.path.join(CSRC, "pgeof_ext.cpp"), "-o", tmp, f"-L{HERE}", "-lpgeof_hip",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, ext_path())


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB, ext_path())
