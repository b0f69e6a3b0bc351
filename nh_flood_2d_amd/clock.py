"""Host half of the 2^-40 s tick clock (core/clock.py:9-30).

The device half (StepClock.choose / commit) runs inside the HIP kernels; the host only converts
seconds to ticks, mirrors the last completed tick from the trace, and decodes trace rows.
"""
from __future__ import annotations

import math

import numpy as np

TICKS_PER_SECOND = 1 << 40
MAX_TICKS = (1 << 64) - 1


def seconds_to_ticks(seconds: float) -> int:
    """Exact F64 seconds -> ticks, rounding half to even (clock.py:14-24)."""
    seconds = float(seconds)
    if not math.isfinite(seconds) or seconds < 0:
        raise ValueError('Clock seconds must be finite and nonnegative')
    num, den = seconds.as_integer_ratio()
    q, r = divmod(num * TICKS_PER_SECOND, den)
    twice = 2 * r
    ticks = q + int(twice > den or (twice == den and q & 1))
    if ticks > MAX_TICKS:
        raise ValueError('Clock exceeds the two-U32 range')
    return ticks


def ticks_to_seconds(ticks: int) -> float:
    return int(ticks) / TICKS_PER_SECOND


def row_end_ticks(row) -> int:
    return (int(row[1]) << 32) | int(row[2])


def f32_from_bits(bits) -> float:
    return float(np.array([bits], dtype=np.uint32).view(np.float32)[0])
