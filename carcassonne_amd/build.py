"""Compile the HIP sources in csrc/ into carcassonne_amd/libcarcassonne_hip.so (gfx950 only)."""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcarcassonne_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(HERE, "..", "include")]
# -amdgpu-mfma-vgpr-form: MFMA accumulators in ordinary VGPRs.  Without it the register allocator puts
# them in AGPRs and, in loops that carry several accumulators, copies them between the two register
# files around the MFMAs (64 copies per 32 MFMAs in the FP64 GEMM's K loop).  Not for the fused stage-3
# kernels: the FP64 one at P = 36 needs more than the 256 ordinary VGPRs for its nine tiles' accumulators,
# and the forced form made it 3.3x slower (0.28 -> 0.93 ms per matvec at D = 6, chi = 64).
VGPR_FORM = ["-mllvm", "-amdgpu-mfma-vgpr-form"]
VGPR_FORM_EXCLUDED = {"cc_stage3_fused.hip"}


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OUT_DIR, os.path.basename(src)[:-4] + ".o")
    extra = [] if os.path.basename(src) in VGPR_FORM_EXCLUDED else VGPR_FORM
    cmd = [HIPCC] + FLAGS + extra + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for {}:\n{}".format(src, r.stderr))
    return obj


def build(verbose=False, jobs=8):
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB, "from", len(srcs), "sources")
    return LIB


if __name__ == "__main__":
    build(verbose=True, jobs=int(sys.argv[1]) if len(sys.argv) > 1 else 8)
