"""Builds libsnicit.so (HIP, gfx950), the oracle library and the drop-in CLI.

Everything is built in-tree so the shared objects travel with the repository snapshot.
"""
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "snicit_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libsnicit.so")
CLI = os.path.join(PKG, "snicit_sdgc")
CLI_BEYOND = os.path.join(PKG, "snicit_beyond")
ORACLE_DIR = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions stay off: all HIP code is compiled without them.
HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
SOURCES = ["kernels.hip", "gen.cpp", "net.cpp", "api.cpp", "server.cpp"]


def _run(cmd, cwd=ROOT):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "is not a recognized feature" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_lib():
    objs = []
    odir = os.path.join(PKG, "_obj")
    os.makedirs(odir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(odir, src.replace(".", "_") + ".o")
        lang = ["-x", "hip"]
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    for name, out in (("cli_sdgc.cpp", CLI), ("cli_beyond.cpp", CLI_BEYOND)):
        _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
              os.path.join(CSRC, name), "-o", out, "-L", PKG, "-lsnicit", "-Wl,-rpath,$ORIGIN"])


def build_oracle():
    _run(["make", "-C", ORACLE_DIR, "-j4"])


def build_all():
    build_lib()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"lib": build_lib, "oracle": build_oracle, "all": build_all}[what]()
