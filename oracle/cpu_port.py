"""Oracle (test infrastructure only): the C++/OpenMP CPU restatement of the SVI
erosion step (oracle/cpu/svicpu.cpp), used as the timed CPU baseline and
checked against the NumPy oracle by tests/test_oracle_flow.py.

    python -m oracle.cpu_port      # build oracle/_cpu/libsvicpu.so with g++ -fopenmp
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SRC = HERE / "cpu" / "svicpu.cpp"
LIB = HERE / "_cpu" / "libsvicpu.so"
# x86-64-v3 (AVX2 + FMA): built here, run on any current x86 host
FLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-std=c++17", "-shared", "-fPIC", "-Wall"]


def build(verbose: bool = False) -> Path:
    LIB.parent.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    r = subprocess.run(["g++", *FLAGS, "-o", str(tmp), str(SRC)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on {SRC.name}:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        P, I, D = C.c_void_p, C.c_int, C.c_double
        lib.svicpu_run.restype = I
        lib.svicpu_run.argtypes = [P, I, I, I, D, D, D, I, I, D, D, D, P, I, D, D, D, I, D, D, I, P]
        _lib = lib
    return _lib


def run_steps(phi: np.ndarray, h: float, law, steps: int, width: float = 6.0, cfl: float = 0.5,
              feedback_cfl: float = 1.0, reinit_distance: float = 1.5, reinit_iterations: int = 20,
              cutoff=(4.0, 5.5)):
    """``steps`` accepted TVD-RK3 steps of an SVI law (oracle.flow.SVILaw) on the
    float32 field ``phi`` (3-D, C order).  Returns (phi after the steps,
    [(dt, volume, reinitialised)] per step, OpenMP threads used)."""
    from .flow import ball_cell_weights
    f = np.ascontiguousarray(phi, dtype=np.float32).copy()
    w = np.ascontiguousarray(ball_cell_weights(law.radius, h), dtype=np.float64)
    m = (w.shape[0] - 1) // 2
    rec = np.zeros(4 * max(steps, 1), dtype=np.float64)
    nx, ny, nz = f.shape
    threads = _load().svicpu_run(f.ctypes.data, nx, ny, nz, h, width, law.radius, law.k, law.s,
                                 law.t if law.t is not None else 0.0, law.eps, law.n_cap, w.ctypes.data, m, cfl,
                                 feedback_cfl, reinit_distance, reinit_iterations, cutoff[0], cutoff[1], steps,
                                 rec.ctypes.data)
    if threads < 0:
        raise ValueError("svicpu_run rejected its arguments")
    rows = [(rec[4 * i], rec[4 * i + 1], bool(rec[4 * i + 3])) for i in range(steps)]
    return f, rows, int(threads)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
