"""Hydrographs at stations: simulated samples against observations and against each other (the
reference's output/hydrograph.py).

* :func:`station_series` reads a station's samples from ``output/stations.csv`` (the solver's
  station output), at and after the run start, by time (hydrograph.py:13-21).
* :func:`observation_series` reads ``<observation_dir>/<station>.csv``: the one column whose name
  holds "water" and the quantity, timed by its Date (day first) and Time columns (:24-32).
* :func:`comparison_data` lines up several runs' samples (shifted by `translation_second`, their
  first `forward_ignore_second` dropped) with the observations, clamped to the span all of them
  cover; every run must use the same observation file (:79-142).
* :func:`compare_hydrograph` returns RMSE and Nash-Sutcliffe efficiency of each run against the
  observations, or against one of the runs, after interpolating each run linearly onto the
  reference series' times (:145-237); plots as the reference does when asked.
* :func:`draw_hydrograph` and :func:`compare_hydrograph_panels` are the plots (:35-76, :240-388).

Plots use matplotlib; ``show=False`` with ``output_path`` renders off screen.
"""
from __future__ import annotations

from pathlib import Path
from typing import Literal

import numpy as np
import pandas as pd

LABEL = {'depth': 'Water Depth', 'level': 'Water Level'}


def station_series(cfg, station: str, quantity: Literal['depth', 'level']) -> pd.DataFrame:
    path = Path(cfg.domain_dir) / 'output' / 'stations.csv'
    samples = pd.read_csv(path)
    picked = samples[(samples['station'] == station) & (samples['elapsed_seconds'] >= 0)]
    if picked.empty:
        raise ValueError(f'No simulated samples for station {station} in {path}')
    out = pd.DataFrame({'datetime': pd.to_datetime(picked['time']).dt.tz_localize(None),
                        quantity: picked[quantity].to_numpy()})
    return out.sort_values('datetime')


def observation_series(path, quantity: Literal['depth', 'level']) -> pd.DataFrame:
    table = pd.read_csv(path)
    columns = [c for c in table.columns if 'water' in c.lower() and quantity in c.lower()]
    if len(columns) != 1:
        raise ValueError(f'Need exactly one water {quantity} column in {path}')
    out = pd.DataFrame({'datetime': pd.to_datetime(table['Date'] + ' ' + table['Time'], dayfirst=True),
                        quantity: pd.to_numeric(table[columns[0]])})
    return out.dropna(subset=[quantity]).sort_values('datetime')


def _within(frame: pd.DataFrame, start, end) -> pd.DataFrame:
    return frame[(frame['datetime'] >= start) & (frame['datetime'] <= end)]


def comparison_data(cfgs: list, station: str, quantity: Literal['depth', 'level'], clamped: bool = True,
                    translation_second: int = 0, forward_ignore_second: int = 0, show_obs: bool = True):
    """(observations or None, [(run label = domain directory name, samples)])."""
    if not cfgs:
        raise ValueError('At least one domain config is required for hydrograph comparison')
    files = list(dict.fromkeys(str(Path(c.observation_dir) / f'{station}.csv') for c in cfgs
                               if (Path(c.observation_dir) / f'{station}.csv').exists()))
    if len(files) > 1:
        raise ValueError(f'Configs reference different observation files for station {station}: {files}. '
                         'All configs must use the same observation data.')
    observed = observation_series(files[0], quantity) if show_obs and files else None
    runs = []
    for cfg in cfgs:
        series = station_series(cfg, station, quantity)
        series['datetime'] += pd.to_timedelta(translation_second, unit='s')
        series = series.sort_values('datetime')
        if forward_ignore_second > 0:
            series = series[series['datetime'] >= series['datetime'].min()
                            + pd.to_timedelta(forward_ignore_second, unit='s')]
        runs.append((Path(cfg.domain_dir).name, series))
    if clamped:
        spans = [(s['datetime'].min(), s['datetime'].max()) for _, s in runs if not s.empty]
        if observed is not None and not observed.empty:
            spans.append((observed['datetime'].min(), observed['datetime'].max()))
        if spans:
            start, end = max(a for a, _ in spans), min(b for _, b in spans)
            if observed is not None:
                observed = _within(observed, start, end)
            runs = [(label, _within(s, start, end)) for label, s in runs]
    return observed, runs


def nse(observed: np.ndarray, simulated: np.ndarray) -> float:
    """Nash-Sutcliffe efficiency; -inf for a constant observed series."""
    spread = np.sum((observed - observed.mean()) ** 2)
    return float(1.0 - np.sum((observed - simulated) ** 2) / spread) if spread != 0 else float('-inf')


def _seconds(frame: pd.DataFrame) -> np.ndarray:
    return frame['datetime'].astype(np.int64).to_numpy() / 1e9


def scores(reference: pd.DataFrame, runs: list, quantity: str, skip: str | None = None):
    """RMSE and NSE of each run (but `skip`) interpolated onto the reference series' times."""
    t_ref, v_ref = _seconds(reference), reference[quantity].to_numpy()
    rmse, efficiency = [], []
    for label, series in runs:
        if label == skip:
            continue
        values = np.interp(t_ref, _seconds(series), series[quantity].to_numpy())
        rmse.append(float(np.sqrt(np.mean((v_ref - values) ** 2))))
        efficiency.append(nse(v_ref, values))
    return rmse, efficiency


def _pyplot(show: bool):
    import matplotlib
    if not show:
        matplotlib.use('Agg')
    import matplotlib.pyplot as plt
    return plt


def compare_hydrograph(cfgs: list, station: str, quantity: Literal['depth', 'level'], clamped: bool = True,
                       translation_second: int = 0, forward_ignore_second: int = 0, show_obs: bool = True,
                       baseline=None, show: bool = True, output_path=None) -> tuple[list[float], list[float]]:
    """RMSE and NSE of every run against the observations, or with `baseline` (one of `cfgs`)
    against that run; ([], []) without observations or baseline. Plots the series when `show`
    or `output_path`."""
    observed, runs = comparison_data(cfgs, station, quantity, clamped, translation_second, forward_ignore_second,
                                     show_obs)
    if show or output_path:
        plt = _pyplot(show)
        fig = plt.figure(figsize=(12, 6))
        if observed is not None:
            plt.plot(observed['datetime'], observed[quantity], label=f'Observed {LABEL[quantity]} (m)', linewidth=2)
        for label, series in runs:
            plt.plot(series['datetime'], series[quantity], label=f'Simulated {LABEL[quantity]} (m) [{label}]',
                     linewidth=2)
        plt.title(f'{LABEL[quantity]} Hydrograph Comparison at {station}', fontsize=16, fontweight='bold')
        plt.xlabel('Time', fontsize=14)
        plt.ylabel(f'{LABEL[quantity]} (m)', fontsize=14)
        plt.legend()
        plt.tight_layout()
        if output_path:
            fig.savefig(output_path, dpi=150, bbox_inches='tight')
        if show:
            plt.show()
        plt.close(fig)
    if baseline is not None:
        name = Path(baseline.domain_dir).name
        base = next((series for label, series in runs if label == name), None)
        if base is None:
            raise ValueError(f'Baseline config "{name}" was not found in cfgs. The baseline must be one of the '
                             'provided configs.')
        return scores(base, runs, quantity, skip=name)
    if observed is None:
        return [], []
    return scores(observed, runs, quantity)


def draw_hydrograph(cfg, station: str, quantity: Literal['depth', 'level'], clamped: bool = True,
                    translation_second: int = 0, show: bool = True, output_path=None):
    """Observed against simulated at one station (the observation file must exist)."""
    path = Path(cfg.observation_dir) / f'{station}.csv'
    if not path.exists():
        raise FileNotFoundError(f'Observation file for station {station} not found at {path}')
    observed, runs = comparison_data([cfg], station, quantity, clamped, translation_second, 0, True)
    plt = _pyplot(show)
    fig = plt.figure(figsize=(12, 6))
    plt.plot(observed['datetime'], observed[quantity], label=f'Observed {LABEL[quantity]} (m)', linewidth=2)
    plt.plot(runs[0][1]['datetime'], runs[0][1][quantity], label=f'Simulated {LABEL[quantity]} (m)', linewidth=2)
    plt.title(f'Hydrograph at Station {station}', fontsize=16)
    plt.xlabel('Time', fontsize=14)
    plt.ylabel(f'{LABEL[quantity]} (m)', fontsize=14)
    plt.legend()
    plt.tight_layout()
    if output_path:
        fig.savefig(output_path, dpi=150, bbox_inches='tight')
    if show:
        plt.show()
    plt.close(fig)
    return observed, runs[0][1]


def compare_hydrograph_panels(cfgs: list, stations: list[str], quantity: Literal['depth', 'level'],
                              clamped: bool = True, translation_second: int = 0, forward_ignore_second: int = 0,
                              show_obs: bool = False, show: bool = True, output_path=None, title: str | None = None,
                              figsize=None, dpi: int = 300):
    """One panel per station, shared axes over every series' span and value range (5 % padding,
    at least 0.05 m), one legend for the figure. Returns the figure."""
    if not stations:
        raise ValueError('At least one station name is required for panel plotting')
    panels, values, starts, ends = [], [], [], []
    for station in stations:
        observed, runs = comparison_data(cfgs, station, quantity, clamped, translation_second,
                                         forward_ignore_second, show_obs)
        panels.append((station, observed, runs))
        for series in ([observed] if observed is not None else []) + [s for _, s in runs]:
            if not series.empty:
                values.extend(series[quantity].tolist())
                starts.append(series['datetime'].min())
                ends.append(series['datetime'].max())
    if not values:
        raise ValueError('No hydrograph data available for panel plotting')
    plt = _pyplot(show)
    import matplotlib.dates as mdates
    low, high = min(values), max(values)
    pad = max((high - low) * 0.05, 0.05)
    fig, axes = plt.subplots(1, len(stations), figsize=figsize or (10 * len(stations), 5.4), sharex=True,
                             sharey=True)
    axes = np.atleast_1d(axes)
    for axis, (station, observed, runs) in zip(axes, panels):
        if observed is not None:
            axis.plot(observed['datetime'], observed[quantity], label=f'Observed {LABEL[quantity]} (m)',
                      linewidth=2)
        for label, series in runs:
            axis.plot(series['datetime'], series[quantity], label=f'Simulated {LABEL[quantity]} (m) [{label}]',
                      linewidth=2)
        axis.set_xlim(min(starts), max(ends))
        axis.set_ylim(low - pad, high + pad)
        axis.set_xlabel('Time', fontsize=13)
        axis.grid(axis='y', linestyle='--', linewidth=0.8, alpha=0.28)
        axis.xaxis.set_major_locator(mdates.HourLocator(byhour=[0, 12]))
        axis.xaxis.set_major_formatter(mdates.DateFormatter('%m-%d %H'))
        axis.text(0.98, 0.02, station, transform=axis.transAxes, ha='right', va='bottom', fontsize=17,
                  fontweight='bold')
    axes[0].set_ylabel(f'{LABEL[quantity]} (m)', fontsize=13)
    handles, labels = [], []
    for axis in axes:
        for handle, label in zip(*axis.get_legend_handles_labels()):
            if label not in labels:
                handles.append(handle)
                labels.append(label)
    if handles:
        fig.legend(handles, labels, loc='upper center', bbox_to_anchor=(0.5, 0.91), ncol=max(1, min(3, len(labels))))
    fig.suptitle(title or f'{LABEL[quantity]} Hydrograph Comparison', fontsize=20, y=0.965)
    fig.tight_layout(rect=(0.0, 0.0, 1.0, 0.90))
    if output_path is not None:
        out = Path(output_path)
        out.parent.mkdir(parents=True, exist_ok=True)
        fig.savefig(out, dpi=dpi, bbox_inches='tight', facecolor='white')
    if show:
        plt.show()
    return fig
