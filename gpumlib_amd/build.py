"""Build libgpumlib_nmf.so for gfx950 in-tree (gpumlib_amd/lib/).

Every .hip kernel file is compiled by hipcc for --offload-arch=gfx950, the C-ABI
and class-handle layers by hipcc as host C++, and the input generators by g++
without floating-point contraction.  Objects go to gpumlib_amd/_build/.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libgpumlib_nmf.so")
EXAMPLES = os.path.join(HERE, "examples")
NATIVE_TESTS = os.path.join(os.path.dirname(HERE), "tests", "native")
BIN_DIR = os.path.join(HERE, "bin")

ARCH = os.environ.get("GPUMLIB_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# Packed FP32 instructions stay off in all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wno-unused-result"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def _jobs():
    jobs = []
    for f in sorted(os.listdir(CSRC)):
        src = os.path.join(CSRC, f)
        obj = os.path.join(OBJ_DIR, f + ".o")
        if f.endswith(".hip"):
            jobs.append([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
        elif f.endswith(".cpp"):
            # host-only C++ over the HIP runtime API; the generators must not contract a*b+c
            jobs.append(["g++", *HOST_FLAGS, "-ffp-contract=off", "-pthread", "-D__HIP_PLATFORM_AMD__",
                         "-I/opt/rocm/include", "-c", src, "-o", obj])
    return jobs


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    jobs = _jobs()
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    # example drivers over the C++ drop-in API, and the native checks of tests/native
    os.makedirs(BIN_DIR, exist_ok=True)
    sources = [os.path.join(EXAMPLES, f) for f in sorted(os.listdir(EXAMPLES)) if f.endswith(".cpp")]
    if os.path.isdir(NATIVE_TESTS):
        sources += [os.path.join(NATIVE_TESTS, f) for f in sorted(os.listdir(NATIVE_TESTS)) if f.endswith(".cpp")]
    for src in sources:
        exe = os.path.join(BIN_DIR, os.path.basename(src)[:-4])
        _run(["g++", *HOST_FLAGS, "-pthread", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
              src, "-o", exe + ".tmp", f"-L{OUT_DIR}", "-lgpumlib_nmf",
              f"-Wl,-rpath,$ORIGIN/../lib", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
        os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
