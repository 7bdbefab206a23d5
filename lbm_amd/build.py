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
            "-packed-fp32-ops", "-Wno-unused-result", "-Wno-unused-value"]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        print(out)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build_hip():
    src = os.path.join(HERE, "csrc", "lbm_hip.hip")
    out = os.path.join(HERE, "liblbm_hip.so")
    run([HIPCC, *HIPFLAGS, "-shared", "-o", out + ".tmp", src])
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
