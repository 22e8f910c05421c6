"""Builds the mx API shim and the two MEX entry points against it (g++), for
the MEX-layer tests:  _build/libmxshim.so, _build/mex_patch_sweep_cuda.so,
_build/mex_gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST.so (next to libefie_hip.so).
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PKG = os.path.join(ROOT, "matlab-cuda-efief-vs-cpu_amd")
OUT = os.path.join(PKG, "_build")
MEX_NAMES = ["patch_sweep_cuda", "gassemble_ZV_EFIE_RWG_DIFF_GPU_FAST"]


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
    return outs


if __name__ == "__main__":
    for p in build():
        print(p)
