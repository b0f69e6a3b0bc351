"""CPU oracle of the compact local-inertial step (test infrastructure only).

``nhf_oracle.c`` restates the reference's passes in plain C; ``build()`` compiles it into
``oracle/_build/liboracle_f32.so`` (F32, the reference's arithmetic) and ``liboracle_f64.so``
(the FP64 comparison run). :class:`HostDevice` lets the product's host driver run on it, so
parity tests drive the same host protocol through both libraries. Only ``tests/``,
``__graft_entry__.smoke`` and the benchmark's CPU-baseline leg import this package.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
BUILD_DIR = ORACLE_DIR / '_build'
LIBRARIES = {'f32': BUILD_DIR / 'liboracle_f32.so', 'f64': BUILD_DIR / 'liboracle_f64.so'}


def build() -> None:
    """Compile the C restatement (gcc, OpenMP, no FMA contraction)."""
    subprocess.run(['make', '-C', str(ORACLE_DIR), '-s', 'all'], check=True)


_loaded: dict = {}


def load(precision: str = 'f32') -> ctypes.CDLL:
    from nh_flood_2d_amd import abi
    if precision not in _loaded:
        path = LIBRARIES[precision]
        if not path.exists():
            build()
        _loaded[precision] = abi.bind(ctypes.CDLL(str(path)))
    return _loaded[precision]


class HostDevice:
    """Host arrays standing in for device buffers, so the oracle runs behind ``Engine``."""

    def __init__(self, precision: str = 'f32'):
        self.lib = load(precision)
        self.keep = []

    def array(self, values: np.ndarray):
        buf = np.ascontiguousarray(values).copy()
        self.keep.append(buf)
        return buf

    def empty(self, nbytes: int):
        buf = np.zeros(max(nbytes, 1), dtype=np.uint8)
        self.keep.append(buf)
        return buf

    @staticmethod
    def read(buf, dtype) -> np.ndarray:
        return buf.view(dtype).copy()

    @staticmethod
    def write(buf, values: np.ndarray) -> None:
        buf[...] = np.ascontiguousarray(values).view(buf.dtype)

    def upload(self, buf, values: np.ndarray) -> None:
        self.write(buf, values)

    def download(self, buf, dtype) -> np.ndarray:
        return self.read(buf, dtype)

    @staticmethod
    def zero(buf) -> None:
        buf[...] = 0

    @staticmethod
    def ptr(buf) -> int:
        return buf.ctypes.data

    @staticmethod
    def stream_handle():
        return None

    def synchronize(self) -> None:
        pass


def engine_factory(precision: str = 'f32'):
    """An Engine constructor bound to the oracle library (for Simulation(engine_factory=...))."""
    from nh_flood_2d_amd.engine import Engine

    def make(layout, params, state, pipe_q=None):
        return Engine(layout, params, state, device=HostDevice(precision), pipe_q=pipe_q)
    return make
