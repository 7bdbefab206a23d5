"""In-tree build of the HIP library (gfx950) and the C++ host.

    python -m lbm_amd.build            # everything
Outputs: lbm_amd/liblbm_hip.so, lbm_amd/bin/lbm (when the host sources exist).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# packed FP32 instructions are disabled for every HIP translation unit
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
            "-packed-fp32-ops", "-ffp-contract=off", "-Wno-unused-result", "-Wno-unused-value"]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        print(out)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


# translation units of liblbm_hip.so: the API and set-up kernels, the step
# kernels of each storage type, the warm-start network and the Poisson
# solver (compiled in parallel)
HIP_SOURCES = ["lbm_hip.hip", "lbm_step_f32.hip", "lbm_step_f64.hip", "lbm_step_f16.hip", "lbm_step_bf16.hip",
               "lbm_surrogate.hip", "lbm_poisson.hip"]


def build_hip(extra_flags=(), out=None, jobs=None):
    """Compile every HIP translation unit for gfx950 and link liblbm_hip.so."""
    from concurrent.futures import ThreadPoolExecutor

    out = out or os.path.join(HERE, "liblbm_hip.so")
    objdir = os.path.join(os.path.dirname(out), "_build")
    os.makedirs(objdir, exist_ok=True)
    flags = [*HIPFLAGS, *extra_flags]

    def compile_one(name):
        obj = os.path.join(objdir, name.replace(".hip", ".o"))
        run([HIPCC, *flags, "-c", "-o", obj, os.path.join(HERE, "csrc", name)])
        return obj

    jobs = jobs or min(len(HIP_SOURCES), os.cpu_count() or 1)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp", *objs])
    os.replace(out + ".tmp", out)
    return out


def build_host():
    mk = os.path.join(HERE, "host", "Makefile")
    if os.path.exists(mk):
        run(["make", "-C", os.path.join(HERE, "host"), "-j8"])


def build_all():
    build_hip()
    build_host()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
