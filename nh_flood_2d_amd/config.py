"""Domain and forcing JSON models.

The JSON schema is the reference's (``src/nh_flood_2d/input/domain.py:11-135`` and
``src/nh_flood_2d/input/force.py:1-76``): the same field names, defaults, bounds and
``extra='forbid'`` / no-NaN validation, so an existing ``domain.json`` / ``force.json`` loads
unchanged. Path helpers (``ne_fdb``, ``output_dir`` ...) keep the reference's directory layout.
"""
from __future__ import annotations

from datetime import datetime
from pathlib import Path
from typing import Annotated, Literal

from pydantic import AwareDatetime, BaseModel, ConfigDict, Field, field_validator, model_validator


def absolute_time(value: str) -> datetime:
    """ISO 8601 timestamp that must carry its own UTC offset (input/force.py:6-11)."""
    parsed = datetime.fromisoformat(value.replace('Z', '+00:00'))
    if parsed.tzinfo is None or parsed.utcoffset() is None:
        raise ValueError(f'Timestamp must include a timezone: {value}')
    return parsed


class InputModel(BaseModel):
    model_config = ConfigDict(extra='forbid', allow_inf_nan=False)


class HortonInfiltration(InputModel):
    f0_mm_h: float = Field(ge=0)
    fc_mm_h: float = Field(ge=0)
    k_per_h: float = Field(ge=0)

    @model_validator(mode='after')
    def _capacity_order(self):
        if self.f0_mm_h < self.fc_mm_h:
            raise ValueError('f0_mm_h must be at least fc_mm_h')
        return self


class LandParameters(InputModel):
    manning_n: float = Field(ge=0)
    infiltration: Literal['none'] | HortonInfiltration


class DomainConfig(InputModel):
    ne: str
    ns: str
    epsg_code: int
    domain_dir: str
    afa: float = Field(default=0.5, gt=0, le=1)
    sita: float = Field(default=1.0, ge=0, le=1)
    min_h: float = Field(default=0.02, gt=0)
    start: AwareDatetime = Field(strict=True)
    duration: int = Field(gt=0)
    warmup_duration: int = Field(ge=0)
    land_parameters: dict[Annotated[int, Field(ge=0, le=255)], LandParameters] = Field(min_length=1)
    yield_step: int = Field(default=300, gt=0)
    restart_state: str = ''
    stations: dict[str, tuple[float, float]] = {}
    observation_dir: str = ''

    @model_validator(mode='after')
    def _clock_range(self):
        if self.duration + self.warmup_duration >= 2 ** 24:
            raise ValueError('Duration plus warmup must fit the two-U32 clock (less than 2**24 seconds)')
        return self

    @field_validator('ne', 'ns')
    @classmethod
    def _mesh_exists(cls, value):
        if not Path(value).exists():
            raise ValueError(f'Path does not exist: {value}')
        return value

    @field_validator('domain_dir')
    @classmethod
    def _make_domain_dir(cls, value):
        Path(value).mkdir(parents=True, exist_ok=True)
        return value

    def _ensure(self, path: Path) -> str:
        path.mkdir(parents=True, exist_ok=True)
        return str(path)

    @property
    def preprocessed_dir(self) -> str:
        return self._ensure(Path(self.domain_dir) / 'preprocessed')

    @property
    def ne_fdb(self) -> str:
        return str(Path(self.preprocessed_dir) / 'ne.fdb')

    @property
    def ns_fdb(self) -> str:
        return str(Path(self.preprocessed_dir) / 'ns.fdb')

    @property
    def boundary_fdb(self) -> str:
        return str(Path(self.preprocessed_dir) / 'boundary.fdb')

    @property
    def tmp_ne(self) -> str:
        return str(Path(self.domain_dir) / 'temp_ne.txt')

    @property
    def tmp_ns(self) -> str:
        return str(Path(self.domain_dir) / 'temp_ns.txt')

    @property
    def output_dir(self) -> str:
        return self._ensure(Path(self.domain_dir) / 'output')

    @property
    def frame_dir(self) -> str:
        return self._ensure(Path(self.output_dir) / 'frames')

    @property
    def flood_map_dir(self) -> str:
        return self._ensure(Path(self.domain_dir) / 'flood_maps')

    @property
    def max_inundation_dir(self) -> str:
        return self._ensure(Path(self.domain_dir) / 'max_inundation')


class BoundaryConfig(InputModel):
    name: str = Field(min_length=1)
    kind: Literal['level', 'discharge']
    line: str
    series: str


class RainfallConfig(InputModel):
    raster: str
    intervals: str


class GatePointConfig(InputModel):
    """One gate (input/force.py:32-35): a point inside its gate cells and its control point."""
    name: str = Field(min_length=1)
    inside_point: tuple[float, float]
    control_point: tuple[float, float]


class GateConfig(InputModel):
    """Gate control (input/force.py:38-54), same fields and checks."""
    land_type: int = Field(ge=0, le=255)
    pond_land_type: int = Field(ge=0, le=255)
    enabled: bool
    trigger_level: float
    opening_seconds: float = Field(gt=0)
    open_elevation: float
    gates: list[GatePointConfig] = Field(min_length=1)

    @model_validator(mode='after')
    def _distinct(self):
        if self.pond_land_type == self.land_type:
            raise ValueError('pond_land_type must differ from gate land_type')
        names = [g.name for g in self.gates]
        if len(names) != len(set(names)):
            raise ValueError('Gate names must be unique')
        return self


class ForceConfig(InputModel):
    boundaries: list[BoundaryConfig]
    rainfall: RainfallConfig
    gate: GateConfig | None = None

    @model_validator(mode='after')
    def _unique_names(self):
        names = [b.name for b in self.boundaries]
        if len(set(names)) != len(names):
            raise ValueError('Boundary names must be unique')
        return self

    def input_paths(self) -> list[str]:
        return [p for b in self.boundaries for p in (b.line, b.series)] + [
            self.rainfall.raster, self.rainfall.intervals]


def load_domain_config(path: str) -> DomainConfig:
    return DomainConfig.model_validate_json(Path(path).read_text())


def load_force_config(path: str) -> ForceConfig:
    return ForceConfig.model_validate_json(Path(path).read_text())
