"""Build recipe of the CPU oracle (test infrastructure only).

Compiles oracle/cpu_ref.c into oracle/_build/libmossq_oracle.so with gcc and
OpenMP.  The target ISA is x86-64-v3 (AVX2/FMA) rather than -march=native so
the library built here also runs on the GPU machine's host CPU.

Complex products are compiled inline (-fcx-limited-range): (ac - bd, ad + bc),
as Julia's ComplexF64 multiply computes them, instead of a call to the C99
runtime helper that re-checks infinities; for finite operands the results are
the same bits.

The reference (MOSSQuantum) is Julia-only and has no C/C++ sources, so there is
no oracle/_ref build of the reference itself.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_build" / "libmossq_oracle.so"
CFLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-fcx-limited-range"]


def build() -> Path:
    OUT.parent.mkdir(exist_ok=True)
    tmp = OUT.with_suffix(".so.tmp")
    cmd = ["gcc", *CFLAGS, str(HERE / "cpu_ref.c"), "-o", str(tmp), "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stderr}")
    tmp.replace(OUT)
    return OUT


if __name__ == "__main__":
    print(build())
