"""Build the gfx950 HIP libraries in-tree (wavefd_amd/lib/*.so).

One shared library per reference library (src/CMakeLists.txt:11-18 names them), so the
reference's Julia bindings can point at these files unchanged:
    libprop3DAcoIsoDenQ_DEO2_FDTD.so  libprop3DAcoVTIDenQ_DEO2_FDTD.so
    libprop3DAcoTTIDenQ_DEO2_FDTD.so  libspacetime.so
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")

LIBS = {
    "libprop3DAcoIsoDenQ_DEO2_FDTD.so": ("prop3DAcoIsoDenQ_DEO2_FDTD.hip", ["-lhipfft"]),
    "libprop3DAcoVTIDenQ_DEO2_FDTD.so": ("prop3DAcoVTIDenQ_DEO2_FDTD.hip", ["-lhipfft"]),
    "libprop3DAcoTTIDenQ_DEO2_FDTD.so": ("prop3DAcoTTIDenQ_DEO2_FDTD.hip", ["-lhipfft"]),
    "libprop2DAcoIsoDenQ_DEO2_FDTD.so": ("prop2DAcoIsoDenQ_DEO2_FDTD.hip", ["-lhipfft"]),
    "libprop2DAcoVTIDenQ_DEO2_FDTD.so": ("prop2DAcoVTIDenQ_DEO2_FDTD.hip", ["-lhipfft"]),
    "libprop2DAcoTTIDenQ_DEO2_FDTD.so": ("prop2DAcoTTIDenQ_DEO2_FDTD.hip", ["-lhipfft"]),
    "libspacetime.so": ("spacetime.hip", []),
    # snapshot codec behind the Compressor's lossy mode (src/compressedio.jl); additive
    "libwavefd_codec.so": ("codec.hip", []),
    # ginsu windows, shared image and device helpers of the Born driver; additive
    "libwavefd_ginsu.so": ("ginsu.hip", []),
    # learned snapshot codec (quantiser + convolutional decoder on the matrix cores); additive
    "libwavefd_lcodec.so": ("lcodec.hip", []),
}

ARCH = os.environ.get("WAVEFD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags():
    return [
        "--offload-arch=" + ARCH,
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        # IEEE-exact expression evaluation: the kernels evaluate the reference's expressions
        # operation by operation (no fused multiply-add contraction)
        "-ffp-contract=off",
        "-fvisibility=hidden",
        # no packed FP32 (v_pk_*_f32) code generation
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-I" + CSRC,
        "-L/opt/rocm/lib",
        "-Wl,-rpath,/opt/rocm/lib",
    ]


def build(verbose=False, only=None):
    os.makedirs(LIBDIR, exist_ok=True)
    procs = []
    for lib, (src, extra) in LIBS.items():
        if only and lib not in only:
            continue
        out = os.path.join(LIBDIR, lib)
        tmp = out + ".tmp"
        cmd = [hipcc()] + flags() + [os.path.join(CSRC, src), "-o", tmp] + extra
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((lib, out, tmp, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for lib, out, tmp, pr in procs:
        log = pr.communicate()[0].decode(errors="replace")
        # hipcc's host pass warns that it does not know the packed-fp32 feature; harmless
        log = "\n".join(x for x in log.splitlines() if "packed-fp32-ops" not in x and x.strip() != "1 warning generated.")
        if pr.returncode != 0:
            failed.append((lib, log))
            continue
        os.replace(tmp, out)
        if verbose and log.strip():
            print(log)
    if failed:
        for lib, log in failed:
            sys.stderr.write("=== %s failed ===\n%s\n" % (lib, log))
        raise RuntimeError("HIP build failed: " + ", ".join(l for l, _ in failed))
    return [os.path.join(LIBDIR, lib) for lib in LIBS if not only or lib in only]


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
