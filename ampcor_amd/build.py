"""Build recipe of the native code: libampcor_hip.so (hipcc, gfx950) and the
CPython module ampcor_hip (g++), both in-tree next to this file.

    python -m ampcor_amd.build

Flags: -ffp-contract=off keeps every rounding where the source puts it (the
FP32 coarse stage is bit-reproducible on the CPU); packed FP32 instructions
are disabled (-packed-fp32-ops).
"""

import os
import shutil
import subprocess
import sys
import sysconfig
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
ARCH = os.environ.get("AMPCOR_HIP_ARCH", "gfx950")

HIP_SOURCES = ["kernels.hip", "mixed.hip", "mixed_nk_a.hip", "mixed_nk_b.hip", "mixed_nk_c.hip",
               "mixed_nk_d.hip", "fp8.hip", "plan.hip", "server.hip", "noalloc.hip"]
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    proc = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in proc.stderr.splitlines() if noise not in l and "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{err}\n{proc.stdout}")
    return err


def library_path():
    return os.path.join(HERE, "libampcor_hip.so")


def module_path():
    return os.path.join(HERE, "ampcor_hip" + sysconfig.get_config_var("EXT_SUFFIX"))


def build(verbose=False):
    """Compiles the HIP sources in parallel (one object each), links the library, then
    builds the CPython module against it."""
    lib = library_path()
    hipcc = _hipcc()
    with tempfile.TemporaryDirectory(prefix="ampcor_hip_build_") as tmp:
        objs = [os.path.join(tmp, os.path.splitext(s)[0] + ".o") for s in HIP_SOURCES]
        cmds = [[hipcc] + HIPCC_FLAGS + [f"-I{INCLUDE}", "-c", os.path.join(CSRC, s), "-o", o]
                for s, o in zip(HIP_SOURCES, objs)]
        with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as pool:
            outs = list(pool.map(_run, cmds))
        outs.append(_run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC"] + objs
                         + ["-o", lib + ".tmp"]))
    os.replace(lib + ".tmp", lib)
    out = "\n".join(o for o in outs if o)
    if verbose and out:
        print(out, file=sys.stderr)
    mod = module_path()
    py_inc = sysconfig.get_paths()["include"]
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{py_inc}", f"-I{INCLUDE}",
           os.path.join(CSRC, "ext_ampcor_hip.cc"), f"-L{HERE}", "-lampcor_hip",
           "-Wl,-rpath,$ORIGIN", "-o", mod + ".tmp"]
    out = _run(cmd)
    os.replace(mod + ".tmp", mod)
    if verbose and out:
        print(out, file=sys.stderr)
    return lib, mod


if __name__ == "__main__":
    for path in build(verbose=True):
        print(path)
