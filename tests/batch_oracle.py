"""A scenario batch made of independent oracle contexts (CPU), with BatchEngine's interface.

BatchRunner drives it exactly as it drives the HIP batch: views are plain oracle Engines that
remember whether their slot began a batch, and run_batch steps only those slots. It lets the
runner's admission and interleaving be tested without a GPU.
"""
from __future__ import annotations

import numpy as np

import oracle
from nh_flood_2d_amd import abi
from nh_flood_2d_amd.engine import Engine


class OracleSlot(Engine):
    def __init__(self, batch, slot, layout, params, state):
        super().__init__(layout, params, state, device=oracle.HostDevice(batch.precision))
        self.batch = batch
        self.slot = slot
        self.pending = False

    def begin(self, target_ticks, origin_ticks):
        super().begin(target_ticks, origin_ticks)
        self.pending = True

    def run_batch(self, steps):
        return self.batch.run_batch(steps)[self.slot].copy()

    def close(self):
        super().close()
        self.batch.owner[self.slot] = None


class OracleBatch:
    def __init__(self, slots: int, precision: str = 'f32'):
        self.slots = slots
        self.precision = precision
        self.owner = [None] * slots
        self.ctx = None
        self.batches = 0

    def free_slots(self):
        return [s for s in range(self.slots) if self.owner[s] is None]

    def view(self, slot=None):
        def factory(layout, params, state, pipe_q=None):
            s = slot if slot is not None else self.free_slots()[0]
            v = OracleSlot(self, s, layout, params, state)
            self.owner[s] = v
            self.ctx = True
            return v
        return factory

    def run_batch(self, steps: int) -> np.ndarray:
        trace = np.zeros((self.slots, steps, abi.TRACE_COLS), dtype=np.uint32)
        for s, v in enumerate(self.owner):
            if v is not None and v.pending:
                trace[s] = Engine.run_batch(v, steps)
                v.pending = False
        self.batches += 1
        return trace
