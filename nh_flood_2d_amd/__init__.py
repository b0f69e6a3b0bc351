"""MI355X-native compact local-inertial flood step (nh-flood-2d ``solver_compact`` path).

Hand-written gfx950 HIP kernels (``csrc/nhflood.hip`` -> ``libnhflood.so``) behind the C ABI of
``include/nhflood.h``; Python host code mirrors the reference's ``run_simulation`` / ``solver``
/ ``warmup_solver`` entry points, its JSON configuration, NE/NS ingestion and output layout.
"""
__version__ = '0.1.0'
