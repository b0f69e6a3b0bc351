"""The package-level drop-in entry points on the HIP engine, end to end from files.

A user of the reference calls ``nh_flood_2d.core.run_simulation(domain_cfg, force_cfg)``
(``src/nh_flood_2d/core/__init__.py:1-4``, ``core/solver_compact.py:152-167``). Here the same call
is ``nh_flood_2d_amd.run_simulation``: the domain and forcing JSON and the NE/NS text of the
CPU-runnable configuration (256 x 256 cells of 10 m, 50 mm/h for 1800 s on a dry start, closed
boundaries, 3600 s) are written to disk, read back as a user would, and run once on the HIP
library and once with the F32 oracle as the device step. The returned mass-balance rows agree
and every depth frame is equal bit for bit. ``solver_coupled`` without a pipe configuration is
``solver`` (``core/coupled/driver.py:12-20``) and gives the same frames again.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import nh_flood_2d_amd
from nh_flood_2d_amd import solver_compact, synth
from nh_flood_2d_amd.config import DomainConfig, ForceConfig
from nh_flood_2d_amd.output import list_depth_frames, read_depth_frame

import cases
import oracle

pytestmark = pytest.mark.gpu


def _write_inputs(root: Path, name: str):
    """The 256 x 256-cell, 10 m CPU-runnable configuration's mesh, rain raster and JSON
    configuration under root/name; returns the JSON paths."""
    sc = synth.plain_uniform(0, n=256, dx=10.)
    path = root / name
    cases.write_mesh(path, sc.mesh)
    extent = 256 * 10.
    rain = cases.write_rain(path, np.array([[[25.]], [[0.]]], np.float32), 0., extent, extent,
                            [(0, 1800), (1800, 3600)])
    domain = DomainConfig(ne=str(path / 'ne.txt'), ns=str(path / 'ns.txt'), epsg_code=cases.EPSG,
                          domain_dir=str(path / 'out'), start=cases.START, duration=3600, warmup_duration=0,
                          land_parameters={int(k): v for k, v in sc.land_parameters.items()},
                          yield_step=900, sita=1.0, afa=0.5, min_h=0.02)
    force = ForceConfig(boundaries=[], rainfall=rain)
    (path / 'domain.json').write_text(domain.model_dump_json())
    (path / 'force.json').write_text(force.model_dump_json())
    return path / 'domain.json', path / 'force.json'


def _load(domain_json: Path, force_json: Path):
    return (DomainConfig.model_validate_json(domain_json.read_text()),
            ForceConfig.model_validate_json(force_json.read_text()))


def _frames(domain: DomainConfig):
    return [(seconds, read_depth_frame(path)) for seconds, path in list_depth_frames(domain.output_dir)]


def test_package_run_simulation_matches_oracle_from_files(gpu_device, tmp_path, monkeypatch):
    print('[entry points] run_simulation, 256 x 256 cells, 3600 s, HIP then oracle', flush=True)
    assert callable(nh_flood_2d_amd.run_simulation) and nh_flood_2d_amd.solver_coupled is not None
    rows, frames = {}, {}
    for name, factory in (('hip', None), ('oracle', oracle.engine_factory('f32'))):
        monkeypatch.setattr(solver_compact, 'ENGINE_FACTORY', factory)
        domain, force = _load(*_write_inputs(tmp_path, name))
        rows[name] = nh_flood_2d_amd.run_simulation(domain, force)
        frames[name] = _frames(domain)
        manifest = json.loads((Path(domain.domain_dir) / 'manifest.json').read_text())
        assert manifest['status'] == 'completed'
        if name == 'hip':
            assert manifest['device_name'] and manifest['device_name'] != 'injected engine'
    hip, ref = rows['hip'], rows['oracle']
    assert hip.keys() == ref.keys()
    scale = max(abs(ref['storage_m3']), 1.)
    for key, value in ref.items():
        if isinstance(value, float):
            # storage and inputs come from the same depth bits; the device budget sums are FP64
            # per wavefront on the GPU and exact per cell in the oracle
            assert abs(hip[key] - value) <= 1e-9 * scale, (key, hip[key], value)
        else:
            assert hip[key] == value, key
    assert hip['storage_m3'] == ref['storage_m3']
    assert [t for t, _ in frames['hip']] == [t for t, _ in frames['oracle']] == [0, 900, 1800, 2700, 3600]
    for (t, a), (_, b) in zip(frames['hip'], frames['oracle']):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=f'frame {t}')
    assert frames['hip'][-1][1].max() > 0.1

    # solver_coupled on the preprocessed domain, no pipe: the same run as solver
    monkeypatch.setattr(solver_compact, 'ENGINE_FACTORY', None)
    domain, force = _load(*_write_inputs(tmp_path, 'coupled'))
    solver_compact.preprocess(domain, force)
    row = nh_flood_2d_amd.solver_coupled(domain, force)
    assert row['storage_m3'] == hip['storage_m3']
    for (t, a), (_, b) in zip(_frames(domain), frames['hip']):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=f'coupled frame {t}')
    print(f'[entry points] storage {hip["storage_m3"]:.6f} m3, residual {hip["residual_m3"]:.3g} m3', flush=True)
