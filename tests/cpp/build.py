"""Build recipe for the C++ front-end test program (tests/cpp/test_frontend).

Host-only C++ compiled with hipcc (for the HIP runtime headers), linked
against the in-tree libremora_hip.so and the oracle with $ORIGIN-relative
rpaths so the binary runs wherever the tree is copied.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "test_frontend")


def build(verbose: bool = False) -> str:
    cmd = [os.environ.get("HIPCC", "/opt/rocm/bin/hipcc"), "--offload-arch=gfx950", "-std=c++17", "-O2", "-Wall",
           os.path.join(HERE, "test_frontend.cpp"), "-I", os.path.join(ROOT, "include"),
           "-L", os.path.join(ROOT, "remora_amd", "lib"), "-lremora_hip",
           "-L", os.path.join(ROOT, "oracle"), "-loracle",
           "-Wl,-rpath,$ORIGIN/../../remora_amd/lib:$ORIGIN/../../oracle", "-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
