"""MI355X-native compact local-inertial flood step (nh-flood-2d ``solver_compact`` path).

Hand-written gfx950 HIP kernels (``csrc/nhflood.hip`` -> ``libnhflood.so``) behind the C ABI of
``include/nhflood.h``; Python host code mirrors the reference's entry points, its JSON
configuration, NE/NS ingestion and output layout.

The package exports the same four entry points as the reference's ``nh_flood_2d.core``
(``src/nh_flood_2d/core/__init__.py:1-4``): ``run_simulation``, ``solver``, ``warmup_solver``
and ``solver_coupled``.
"""
__version__ = '0.2.0'

from .solver_compact import run_simulation, solver, solver_coupled, warmup_solver  # noqa: E402

__all__ = ['run_simulation', 'solver', 'warmup_solver', 'solver_coupled', '__version__']
