"""Build recipe for the CPU oracle (test infrastructure only).

Compiles ``oracle/remora_cpu.c`` -- the plain-C restatement of Remora's default
CPU kernels -- into ``oracle/liboracle.so`` with the flags the reference's CPU
configuration names: ``-O3 -DREMORA_USE_SIMD -DNDEBUG`` and an AVX target, so
Remora's vector width is 32 bytes (include/remora/kernels/default/simd.hpp:49-53).
``-march=x86-64-v3`` (AVX2 + FMA) is used instead of ``-march=native`` so the
same binary runs on any host the library is tested on; with ``__AVX__``
defined the kernel's block constants are the ones ``-march=native`` gives on
an AVX/AVX-512 machine.

Remora's own CPU path is not built: every header on it includes Boost
(simd.hpp:34, boost_align/*, cpu/dense.hpp:37-40) and Boost is absent from
this image, so there is no ``oracle/_ref`` build.

Mutant builds (``mutant=...``) plant one deliberate mistake, used only by the
tests that show the oracle's pins catch such mistakes.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "remora_cpu.c")
LIB = os.path.join(HERE, "liboracle.so")

CFLAGS = ["-O3", "-march=x86-64-v3", "-DREMORA_USE_SIMD", "-DNDEBUG", "-std=gnu11",
          "-fPIC", "-shared", "-Wall", "-Wno-unused-function"]

MUTANTS = ("UGEMM_DROP_LAST_K", "TRSM_TRANSPOSED_A", "POTRF_SIGN", "SYRK_WRONG_TRIANGLE", "UPDATE_UNSCALED_TEMP", "PSTRF_LAST_MAX",
           "TRSV_SKIP_FIRST", "GEMV_NO_ALPHA")


def build(out: str = LIB, mutant: str | None = None, verbose: bool = False) -> str:
    cmd = ["gcc", *CFLAGS, SRC, "-o", out, "-lm"]
    if mutant is not None:
        if mutant not in MUTANTS:
            raise ValueError(f"unknown mutant {mutant!r}")
        cmd.insert(1, f"-DORC_MUTANT_{mutant}")
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
