"""In-tree build of libggbtd.so for gfx950 (hipcc cross-compiles without a GPU)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "ggbtd.hip")
OUT = os.path.join(HERE, "libggbtd.so")

HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
               # keep the compiler from forming packed FP32 instructions
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-pass-failed",
               # the propagator tables' batched eigendecomposition (rocSOLVER on rocBLAS)
               "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas"]


def build(verbose=False):
    """Compile csrc/ggbtd.hip into ggbtddft_amd/libggbtd.so (always recompiles)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), SRC, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose and (res.stdout or res.stderr):
        print(res.stdout + res.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
