"""Synthetic accessor chunks of the benchmark workloads (numpy, seeded).

The 36-antenna array follows the fake accessor of askap/gridding/TestCFGenPerformance.cc:70-133
(36 antennas, stub data); uvw use the textbook earth-rotation relation (as the simulator's,
askap/simulation/Simulator.cc:932-963).

* ``sphfunc_disc``   -- plain SphFunc gridding: single-pol visibilities with uv uniform in a
                        disc, w = 0 (the tConvolve / tGridding style case).
* ``wproject_array`` -- W-projection: 36 antennas uniform in a disc, earth rotation over the
                        given hour-angle span at a declination, linear channels, 4 pols.
* ``batched_beams``  -- many W-projection grids at once: 36 beams with uneven channel counts
                        (8..64), visibility counts (log-uniform 1e4..2e6 per polarisation) and
                        grid sizes (1024^2, 2048^2, 4096^2), rows of the 36-antenna array.
* ``learned_wdisc``  -- the ``sphfunc_disc`` rows with w drawn uniformly in [-wmax, wmax]
                        (continuous w) for W-projection with learned CFs.
* ``awproject_feeds`` -- A-projection: the ``sphfunc_disc`` uv generator at 32 channels around
                        1.4 GHz, 9 feeds pointed on a 3x3 grid of 1 degree spacing, a 12 m
                        uniform disk illumination.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

C_LIGHT = 299792458.0
ARCSEC = math.pi / 180.0 / 3600.0


@dataclass
class Workload:
    name: str
    uvw: np.ndarray          # (nrow, 3) float64 metres
    freq: np.ndarray         # (nchan,) Hz
    vis: np.ndarray          # (npol, nchan, nrow) complex64 (casacore cube order)
    noise: np.ndarray        # same shape, complex64
    nx: int
    ny: int
    cell: float              # radians
    gridder: str
    params: dict = field(default_factory=dict)
    feed1: np.ndarray | None = None       # (nrow,) int32 feed of each row
    pointing1: np.ndarray | None = None   # (nrow, 2) pointing (lon, lat) rad of each row
    centre: tuple = (0.0, 0.0)            # image centre (lon, lat) rad
    pa: np.ndarray | None = None          # (nrow,) parallactic angle rad (no effect on a symmetric disk)

    @property
    def nrow(self):
        return self.uvw.shape[0]

    @property
    def nchan(self):
        return self.freq.shape[0]

    @property
    def npol(self):
        return self.vis.shape[0]


def _cn(rng, shape):
    """Complex normal CN(0,1): real and imaginary parts N(0, 1/2)."""
    s = math.sqrt(0.5)
    return (rng.normal(0.0, s, shape) + 1j * rng.normal(0.0, s, shape)).astype(np.complex64)


def sphfunc_disc(nvis: int = 160_000, radius_m: float = 2000.0, freq_hz: float = 1.4e9, n: int = 2048,
                 cell_arcsec: float = 5.0, support: int = 3, oversample: int = 128, seed: int = 20261203) -> Workload:
    rng = np.random.default_rng(seed)
    r = radius_m * np.sqrt(rng.random(nvis))
    th = 2.0 * math.pi * rng.random(nvis)
    uvw = np.stack([r * np.cos(th), r * np.sin(th), np.zeros(nvis)], axis=1)
    vis = _cn(rng, (1, 1, nvis))
    noise = np.ones_like(vis)
    return Workload("sphfunc_disc", uvw, np.array([freq_hz]), vis, noise, n, n, cell_arcsec * ARCSEC, "GPUSphFunc",
                    {"support": support, "oversample": oversample, "alpha": 1.0})


def array_layout(nant: int = 36, diameter_m: float = 6000.0, seed: int = 20261203) -> np.ndarray:
    """East/north antenna positions uniform in a disc of the given diameter."""
    rng = np.random.default_rng(seed)
    r = 0.5 * diameter_m * np.sqrt(rng.random(nant))
    th = 2.0 * math.pi * rng.random(nant)
    return np.stack([r * np.cos(th), r * np.sin(th)], axis=1)


def earth_rotation_uvw(enu: np.ndarray, nint: int, hours: float, dec_deg: float, lat_deg: float = -26.7) -> np.ndarray:
    """uvw (metres) of every baseline (i<j) for nint integrations over the hour-angle span."""
    lat = math.radians(lat_deg)
    dec = math.radians(dec_deg)
    e, nn = enu[:, 0], enu[:, 1]
    # local (east, north, up=0) -> equatorial (X towards H=0, Y east, Z pole)
    X = -math.sin(lat) * nn
    Y = e
    Z = math.cos(lat) * nn
    i, j = np.triu_indices(enu.shape[0], k=1)
    bx, by, bz = X[j] - X[i], Y[j] - Y[i], Z[j] - Z[i]
    H = np.linspace(-hours / 2, hours / 2, nint) * (2 * math.pi / 24.0)
    sH, cH = np.sin(H)[:, None], np.cos(H)[:, None]
    u = sH * bx + cH * by
    v = -math.sin(dec) * cH * bx + math.sin(dec) * sH * by + math.cos(dec) * bz
    w = math.cos(dec) * cH * bx - math.cos(dec) * sH * by + math.sin(dec) * bz
    return np.stack([u.ravel(), v.ravel(), w.ravel()], axis=1)


def wproject_array(nint: int = 300, nchan: int = 64, npol: int = 4, fmin: float = 700e6, fmax: float = 1000e6,
                   hours: float = 8.0, dec_deg: float = -45.0, n: int = 4096, cell_arcsec: float = 4.5,
                   nwplanes: int = 129, oversample: int = 4, limitsupport: int = 64, maxsupport: int = 2048,
                   cutoff: float = 1e-3, seed: int = 20261203, with_vis: bool = True) -> Workload:
    enu = array_layout(seed=seed)
    uvw = earth_rotation_uvw(enu, nint, hours, dec_deg)
    freq = np.linspace(fmin, fmax, nchan) if nchan > 1 else np.array([fmax])
    rng = np.random.default_rng(seed + 1)
    nrow = uvw.shape[0]
    vis = _cn(rng, (npol, nchan, nrow)) if with_vis else np.zeros((npol, nchan, nrow), np.complex64)
    noise = np.ones((npol, nchan, nrow), dtype=np.complex64)
    wmax = float(np.max(np.abs(uvw[:, 2]))) * fmax / C_LIGHT * 1.0001
    params = {"wmax": wmax, "nwplanes": nwplanes, "oversample": oversample, "maxsupport": maxsupport,
              "limitsupport": limitsupport, "cutoff": cutoff}
    return Workload("wproject_array", uvw, freq, vis, noise, n, n, cell_arcsec * ARCSEC, "GPUWProject", params)


def awproject_feeds(nvis: int = 160_000, nchan: int = 32, fcentre: float = 1.4e9, bandwidth: float = 32e6,
                    radius_m: float = 2000.0, n: int = 2048, cell_arcsec: float = 5.0, nfeed_side: int = 3,
                    spacing_deg: float = 1.0, nint: int = 160, pa_step_deg: float = 5.0, oversample: int = 8,
                    maxsupport: int = 128, cutoff: float = 1e-3, diameter: str = "12m",
                    centre: tuple = (0.5, -0.7), seed: int = 20261203) -> Workload:
    """A-projection workload: rows of the sphfunc_disc generator (uv uniform in a disc, w = 0) at
    nchan linear channels over fcentre +- bandwidth/2; row r belongs to feed r % 9 of a 3x3 grid
    of pointings spacing_deg apart around the image centre, and to integration r // (nvis/nint)
    whose parallactic angle steps by pa_step_deg (a uniform disk is rotationally symmetric, so the
    angle changes no convolution function: AWProjectVisGridder.cc:312,373)."""
    rng = np.random.default_rng(seed)
    r = radius_m * np.sqrt(rng.random(nvis))
    th = 2.0 * math.pi * rng.random(nvis)
    uvw = np.stack([r * np.cos(th), r * np.sin(th), np.zeros(nvis)], axis=1)
    freq = fcentre + np.linspace(-bandwidth / 2, bandwidth / 2, nchan)
    vis = _cn(rng, (1, nchan, nvis))
    noise = np.ones_like(vis)
    nfeed = nfeed_side * nfeed_side
    feed1 = (np.arange(nvis) % nfeed).astype(np.int32)
    sp = math.radians(spacing_deg)
    fx = (feed1 % nfeed_side - nfeed_side // 2) * sp
    fy = (feed1 // nfeed_side - nfeed_side // 2) * sp
    pointing = np.stack([centre[0] + fx / math.cos(centre[1]), centre[1] + fy], axis=1)
    pa = np.radians(pa_step_deg) * (np.arange(nvis) // max(1, nvis // nint))
    params = {"wmax": 1.0, "nwplanes": 1, "oversample": oversample, "maxsupport": maxsupport, "maxfeeds": nfeed,
              "cutoff": cutoff, "diameter": diameter, "blockage": "0m", "frequencydependent": True}
    return Workload("awproject_feeds", uvw, freq, vis, noise, n, n, cell_arcsec * ARCSEC, "GPUAWProject", params,
                    feed1=feed1, pointing1=np.ascontiguousarray(pointing), centre=centre, pa=pa)


def batched_beams(nbeam: int = 36, chan_range=(8, 64), nvis_range=(1e4, 2e6), sizes=(1024, 2048, 4096),
                  npol: int = 4, fmin: float = 700e6, fmax: float = 1000e6, nwplanes: int = 129, oversample: int = 4,
                  limitsupport: int = 64, maxsupport: int = 2048, cutoff: float = 1e-3, seed: int = 20261204):
    """One W-projection Workload per beam.  Beam b has nchan_b channels drawn uniformly from
    chan_range over fmin..fmax, nvis_b (row x channel samples per polarisation) log-uniform in
    nvis_range, and a grid of sizes[k] pixels with the cell scaled so that the array's longest
    baseline keeps the same fraction of the grid (4.5 arcsec at 4096).  Rows are integrations of
    the 36-antenna array (earth rotation over 8 h at dec -45 deg) drawn at random; w-planes span
    the |w| of each beam's rows."""
    rng = np.random.default_rng(seed)
    enu = array_layout(seed=20261203)
    base = earth_rotation_uvw(enu, 300, 8.0, -45.0)
    out = []
    for b in range(nbeam):
        nchan = int(rng.integers(chan_range[0], chan_range[1] + 1))
        nvis = float(np.exp(rng.uniform(math.log(nvis_range[0]), math.log(nvis_range[1]))))
        nrow = max(1, int(round(nvis / nchan)))
        n = int(sizes[int(rng.integers(0, len(sizes)))])
        rows = rng.integers(0, base.shape[0], nrow)
        uvw = np.ascontiguousarray(base[rows])
        freq = np.linspace(fmin, fmax, nchan)
        vis = _cn(rng, (npol, nchan, nrow))
        noise = np.ones((npol, nchan, nrow), dtype=np.complex64)
        wmax = float(np.max(np.abs(uvw[:, 2]))) * fmax / C_LIGHT * 1.0001
        params = {"wmax": wmax, "nwplanes": nwplanes, "oversample": oversample, "maxsupport": maxsupport,
                  "limitsupport": limitsupport, "cutoff": cutoff}
        out.append(Workload(f"beam{b}", uvw, freq, vis, noise, n, n, 4.5 * ARCSEC * 4096 / n, "GPUWProject", params))
    return out


def learned_wdisc(nvis: int = 160_000, radius_m: float = 2000.0, w_m: float = 300.0, freq_hz: float = 1.4e9,
                  n: int = 2048, cell_arcsec: float = 5.0, nwplanes: int = 129, oversample: int = 4,
                  cutoff: float = 1e-2, maxsupport: int = 256, seed: int = 20261203) -> Workload:
    """The sphfunc_disc rows (uv uniform in a disc, one channel) with w uniform in [-w_m, w_m]
    metres instead of 0: every sample has its own w, between the planes of any w sampling."""
    wl = sphfunc_disc(nvis=nvis, radius_m=radius_m, freq_hz=freq_hz, n=n, cell_arcsec=cell_arcsec, seed=seed)
    rng = np.random.default_rng(seed + 7)
    uvw = wl.uvw.copy()
    uvw[:, 2] = rng.uniform(-w_m, w_m, wl.nrow)
    wmax = w_m * freq_hz / C_LIGHT * 1.0001
    params = {"wmax": wmax, "nwplanes": nwplanes, "oversample": oversample, "maxsupport": maxsupport,
              "cutoff": cutoff}
    return Workload("learned_wdisc", uvw, wl.freq, wl.vis, wl.noise, n, n, wl.cell, "GPUWProject", params)
