"""Build the HIP shading library (liboslmi.so) for gfx950 with hipcc.

Usage: python -m openshadinglanguage_amd.build   (or build() from Python)
The library is written in-tree to openshadinglanguage_amd/_lib/.
"""
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "_lib")
LIB = os.path.join(OUT_DIR, "liboslmi.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: one IEEE op per source operator (values agree
# bit-for-bit with the reference's operation order); packed FP32 off.
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


# Per-source extra flags. The surrogate's MFMA kernels keep their
# accumulators in VGPRs (the SiLU reads them there) instead of the
# accumulation registers, which then hold the inference weights.
EXTRA_FLAGS = {"oslmi_surrogate.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def build(verbose=False, jobs=8):
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = []
    procs = []
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    for src in sources():
        obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [HIPCC] + HIPFLAGS + EXTRA_FLAGS.get(os.path.basename(src), []) + inc + [
            "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    bake_tables()
    return LIB


def bake_tables():
    """Bakes libbsdl's energy-compensation tables with the freshly built
    library (host code, no device) into _lib/bsdl_luts.bin, where the
    library loads them from; the build-time step libbsdl's genluts is."""
    import ctypes
    path = os.path.join(OUT_DIR, "bsdl_luts.bin")
    if os.path.exists(path):
        os.remove(path)
    code = ("import ctypes,sys; L=ctypes.CDLL(sys.argv[1]); b=(ctypes.c_float*8448)(); "
            "sys.exit(L.oslmi_bsdl_tables(b, ctypes.c_int64(8448)))")
    # a child process, so this process never maps the HIP runtime
    subprocess.check_call([sys.executable, "-c", code, LIB])
    if not os.path.exists(path):
        raise RuntimeError("table bake did not write %s" % path)
    del ctypes


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    if p.returncode != 0:
        sys.stderr.write(out.decode(errors="replace"))
        raise RuntimeError("hipcc failed on %s" % src)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
