"""Exact water-budget accumulation, restated from core/device_budget.py:10-84 (test infrastructure).

Each finite F32 value is an integer multiple of 2**-149. Values are added per (block, term) into
sign-separated planes of base-2**16 digits, so no subtraction or borrow happens while
accumulating; :meth:`flush` merges blocks and signs exactly, converts once to F64 and clears the
interval digits. The product path keeps its budget in FP64 with compensation; this exact form is
the reference the tests compare it with.
"""
from __future__ import annotations

import math

import numpy as np


class ExactBudget:
    def __init__(self, block_count: int, term_count: int, max_steps: int):
        self.word_count = (277 + max_steps.bit_length() + 15) // 16
        self.max_count = max_steps
        self.count = 0
        self.transfer_count = 0
        self.total_units = [0] * term_count
        self.step = np.zeros((block_count, term_count), dtype=np.float32)
        self.digits = np.zeros((2, block_count, term_count, self.word_count), dtype=np.int64)
        self.invalid = 0

    def _accumulate(self, block: int, term: int) -> None:
        bits = int(self.step[block, term:term + 1].view(np.uint32)[0])
        if bits & 0x7fffffff == 0:
            return
        exponent = (bits >> 23) & 255
        if exponent == 255:
            self.invalid = 1
            return
        sign = bits >> 31
        mantissa = bits & 0x7fffff
        if exponent:
            mantissa |= 1 << 23
        offset = max(exponent - 1, 0)          # value = mantissa * 2**(offset - 149)
        word, shift = divmod(offset, 16)
        shifted = mantissa << shift             # at most 39 bits: spans three digits
        carry = 0
        j = word
        while j < self.word_count:
            digit = (shifted >> (16 * (j - word))) & 0xffff if j - word < 3 else 0
            total = int(self.digits[sign, block, term, j]) + digit + carry
            self.digits[sign, block, term, j] = total & 0xffff
            carry = total >> 16
            if j >= word + 2 and carry == 0:
                break
            j += 1

    def collect(self) -> None:
        for block in range(self.step.shape[0]):
            for term in range(self.step.shape[1]):
                self._accumulate(block, term)

    def step_finished(self) -> None:
        self.count += 1

    def flush(self) -> np.ndarray:
        if self.invalid:
            raise FloatingPointError('Nonfinite F32 device budget diagnostic')
        sums = self.digits.sum(axis=1)
        for term in range(len(self.total_units)):
            self.total_units[term] += sum((int(sums[0, term, j]) - int(sums[1, term, j])) << (16 * j)
                                          for j in range(self.word_count))
        totals = np.array([math.ldexp(float(u), -149) for u in self.total_units], dtype=np.float64)
        self.digits[...] = 0
        self.count = 0
        self.transfer_count += 1
        return totals
