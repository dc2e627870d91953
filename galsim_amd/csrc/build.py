"""Build galsim_amd/libgsmi.so in-tree: host numerics (g++, no FP contraction) + HIP kernels
(hipcc, gfx950).  Usage: python -m galsim_amd.csrc.build  (or galsim_amd.csrc.build.build())."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(PKG, "libgsmi.so")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GSMI_ARCH", "gfx950")

HOST_SRCS = ["host_numerics.cpp", "capi_host.cpp"]
HIP_SRCS = ["gsmi_kernels.hip", "gsmi_adjoint.hip", "gsmi_fields.hip", "gsmi_fft.hip",
            "gsmi_real.hip"]

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             # loaded by torch's bundled HIP runtime: keep the code object version it reads
             "-mcode-object-version=5",
             # packed FP32 instructions are kept out of the device code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: %s" % os.path.basename(cmd[-1]))
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    procs = []
    for src in HOST_SRCS:
        obj = os.path.join(BUILD, src + ".o")
        cmd = ["g++"] + HOST_FLAGS + ["-c", os.path.join(HERE, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
        objs.append(obj)
    for src in HIP_SRCS:
        obj = os.path.join(BUILD, src + ".o")
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(HERE, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
        objs.append(obj)
    failed = False
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(" ".join(cmd) + "\n" + out + err)
        elif verbose and (out or err):
            sys.stderr.write(out + err)
    if failed:
        raise RuntimeError("galsim_amd native build failed")
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
