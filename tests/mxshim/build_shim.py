"""Builds the CPU-side test libraries next to libefie_hip.so (g++): the mx
API shim and the MEX entry points against it (_build/libmxshim.so,
_build/mex_<name>.so for each of MEX_NAMES)
and a host build of the device sine/cosine (_build/libefie_math_host.so).
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PKG = os.path.join(ROOT, "matlab-cuda-efief-vs-cpu_amd")
OUT = os.path.join(PKG, "_build")
MEX_NAMES = ["patch_sweep_cuda", "gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST", "patch_sweep_fastquad_cuda",
             "ggassemble_ZV_EFIE_RWG_DIFF_GPU_FAST", "assemble_ZV_EFIE_RWG_DIFF_GPU_FAST"]


def build():
    os.makedirs(OUT, exist_ok=True)
    shim = os.path.join(OUT, "libmxshim.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-std=c99", "-Wall", "-I", HERE,
                    os.path.join(HERE, "mxshim.c"), "-o", shim], check=True)
    outs = [shim]
    for name in MEX_NAMES:
        so = os.path.join(OUT, f"mex_{name}.so")
        subprocess.run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-I", HERE,
                        "-I", os.path.join(ROOT, "include"), os.path.join(PKG, "mex", name + ".cpp"),
                        "-o", so, "-L", OUT, "-lefie_hip", "-lmxshim", "-Wl,-rpath,$ORIGIN"], check=True)
        outs.append(so)
    mathlib = os.path.join(OUT, "libefie_math_host.so")
    subprocess.run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
                    os.path.join(HERE, "efie_math_host.cpp"), "-o", mathlib], check=True)
    outs.append(mathlib)
    return outs


if __name__ == "__main__":
    for p in build():
        print(p)
