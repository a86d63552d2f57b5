"""Analytic model configurations (the inputs initial.nc, forcing.nc and
vmix.nc would hold), generated from fixed seeds.

The reference reads these from netCDF (main.f90:263-345, mod_mympi.f90:
493-560); its input files are not available, so the workloads are defined
analytically:

* ``box``    - closed flat-bottom basin, 60x40 horizontal x 10 levels of
               400 m, 1 degree spacing over 20-60N, zonal wind
               0.1*cos(pi*y/Ly) N m-2, T = 2 + 20 exp(-z/800 m), S = 35,
               at rest.  The reference is always zonally periodic
               (mod_param.f90:325-329); the box is closed by land in the
               first and last interior columns.
* ``global`` - quasi-global 1 degree 360x180x30, zonally periodic, seeded
               Gaussian-smoothed bathymetry scaled to 0-5500 m, land where
               shallower than 50 m and poleward of 70 degrees (barotropic
               CFL of dtbt = 60 s), 30 levels stretched from 20 m,
               tanh latitudinal T profile plus N(0, 0.05 K) noise, smooth
               S in 34-36 psu, analytic annual wind (0.05 N m-2, tapered
               on the equator) and restoring SST.

Arrays are Fortran-ordered (i fastest) float64, without halo, as
pcom_inputs in include/pcom_gpu.h expects them.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi


@dataclass
class Config:
    name: str
    glo_ni: int
    glo_nj: int
    nk: int
    lon: np.ndarray
    lat: np.ndarray
    z: np.ndarray
    itn: np.ndarray          # (glo_ni, glo_nj) int32
    tmn: np.ndarray
    smn: np.ndarray
    pt: np.ndarray           # (glo_ni, glo_nj, nk)
    sa: np.ndarray
    taux: np.ndarray         # (glo_ni, glo_nj, 12)
    tauy: np.ndarray
    bct: np.ndarray
    bcs: np.ndarray
    pa: np.ndarray
    fw: np.ndarray
    vmix: np.ndarray = None  # (glo_ni, glo_nj, nk) or None
    dtbt: int = 60
    dtbc: int = 3600
    start: tuple = (1, 1, 1, 0)
    notes: dict = field(default_factory=dict)

    def c_inputs(self):
        """Build the pcom_inputs struct; keeps the arrays alive on self."""
        keep = {}

        def f64(a):
            a = np.asfortranarray(a, dtype=np.float64)
            keep[id(a)] = a
            return abi.dptr(a)

        itn = np.asfortranarray(self.itn, dtype=np.int32)
        keep["itn"] = itn
        s = abi.PcomInputs()
        s.glo_ni, s.glo_nj, s.nk = self.glo_ni, self.glo_nj, self.nk
        s.lon, s.lat, s.z = f64(self.lon), f64(self.lat), f64(self.z)
        s.itn = abi.iptr(itn)
        s.tmn, s.smn = f64(self.tmn), f64(self.smn)
        s.pt, s.sa = f64(self.pt), f64(self.sa)
        s.vmix = f64(self.vmix) if self.vmix is not None else abi.dptr(None)
        s.taux, s.tauy = f64(self.taux), f64(self.tauy)
        s.bct, s.bcs = f64(self.bct), f64(self.bcs)
        s.pa, s.fw = f64(self.pa), f64(self.fw)
        self._keep = keep
        return s

    def write_raw(self, path):
        """Write the inputs as the raw stream the Fortran driver reads
        (pcom_amd/fortran/pcom_main.f90): int32 glo_ni, glo_nj, nk,
        has_vmix; float64 lon, lat, z; int32 itn; float64 tmn, smn, pt, sa,
        [vmix,] taux, tauy, bct, bcs, pa, fw; Fortran order, no halo."""
        with open(path, "wb") as f:
            hv = 0 if self.vmix is None else 1
            f.write(np.array([self.glo_ni, self.glo_nj, self.nk, hv], dtype="<i4").tobytes())
            for a in (self.lon, self.lat, self.z):
                f.write(np.asarray(a, dtype="<f8").tobytes())
            f.write(np.asfortranarray(self.itn, dtype="<i4").tobytes(order="F"))
            arrs = [self.tmn, self.smn, self.pt, self.sa] + ([self.vmix] if hv else [])
            arrs += [self.taux, self.tauy, self.bct, self.bcs, self.pa, self.fw]
            for a in arrs:
                f.write(np.asfortranarray(a, dtype="<f8").tobytes(order="F"))

    def c_opts(self, precision=abi.PCOM_FP64, device=0, restart_run=0, trop_tile=(0, 0, 0),
               trop_mode=abi.PCOM_TROP_AUTO):
        o = abi.PcomOpts()
        o.trop_tile_i, o.trop_tile_j, o.trop_substeps = trop_tile
        o.trop_mode = trop_mode
        o.dtbt, o.dtbc = self.dtbt, self.dtbc
        o.year, o.month, o.day, o.hour = self.start
        o.precision = precision
        o.restart_run = restart_run
        o.device = device
        return o


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


def box(glo_ni=60, glo_nj=40, nk=10, dz=400.0):
    """Closed flat-bottom box basin at 20-60N (1 degree spacing)."""
    lon = 0.5 + np.arange(glo_ni, dtype=np.float64)
    lat = 20.5 + np.arange(glo_nj, dtype=np.float64) * (40.0 / glo_nj)
    z = dz * (np.arange(nk, dtype=np.float64) + 0.5)
    itn = np.full((glo_ni, glo_nj), nk, dtype=np.int32)
    itn[0, :] = 0
    itn[-1, :] = 0
    tprof = 2.0 + 20.0 * np.exp(-z / 800.0)
    pt = np.broadcast_to(tprof[None, None, :], (glo_ni, glo_nj, nk)).copy()
    sa = np.full((glo_ni, glo_nj, nk), 35.0)
    ly = 40.0
    tx = 0.1 * np.cos(np.pi * (lat - 20.0) / ly)
    taux = np.broadcast_to(tx[None, :, None], (glo_ni, glo_nj, 12)).copy()
    zeros = np.zeros((glo_ni, glo_nj, 12))
    bct = np.broadcast_to(pt[:, :, :1], (glo_ni, glo_nj, 12)).copy()
    bcs = np.full((glo_ni, glo_nj, 12), 35.0)
    return Config(
        name="box", glo_ni=glo_ni, glo_nj=glo_nj, nk=nk,
        lon=_f(lon), lat=_f(lat), z=_f(z), itn=np.asfortranarray(itn),
        tmn=_f(tprof), smn=_f(np.full(nk, 35.0)), pt=_f(pt), sa=_f(sa),
        taux=_f(taux), tauy=_f(zeros), bct=_f(bct), bcs=_f(bcs),
        pa=_f(zeros), fw=_f(zeros.copy()), vmix=None,
        notes={"km_guard": "km(:,:,12) not written when nk+1 < 12"},
    )


def box_pa(seed=11):
    """The box basin with a seeded atmospheric pressure: a smooth monthly
    field of a few hundred Pa (bnd%pa drives int_readyc's pressure force,
    mod_int.f90:100-103, 140-145), for tests of the pressure gradient, the
    forcing hook and restarts; everything else as ``box``."""
    cfg = box()
    rng = np.random.default_rng(seed)
    ni, nj = cfg.glo_ni, cfg.glo_nj
    x = np.arange(ni)[:, None, None] / ni
    y = np.arange(nj)[None, :, None] / nj
    m = np.arange(12)[None, None, :]
    ph = rng.uniform(0.0, 2.0 * np.pi, size=12)[None, None, :]
    pa = 400.0 * np.sin(2.0 * np.pi * x + ph) * np.cos(np.pi * y) + 150.0 * np.cos(2.0 * np.pi * m / 12.0) * y
    cfg.pa = _f(pa)
    cfg.name = "box_pa"
    return cfg


def _smooth(field2d, sigma):
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(field2d, sigma=sigma, mode=("wrap", "reflect"))


def stretched_levels(nk, depth, top=20.0, power=1.5):
    """Layer thicknesses growing from `top` m, summing to `depth` m."""
    s = (np.arange(nk, dtype=np.float64) / max(nk - 1, 1)) ** power
    c = (depth - top * nk) / s.sum()
    dz = top + c * s
    z = np.cumsum(dz) - 0.5 * dz
    return z, dz


def global_1deg(glo_ni=360, glo_nj=180, nk=30, depth=5500.0, polar_cut=70.0,
                seed_bathy=1, seed_t=2, seed_s=3):
    """Quasi-global zonally periodic configuration at 360/glo_ni degrees."""
    dlon = 360.0 / glo_ni
    dlat = 180.0 / glo_nj
    lon = dlon * (np.arange(glo_ni, dtype=np.float64) + 0.5)
    lat = -90.0 + dlat * (np.arange(glo_nj, dtype=np.float64) + 0.5)
    z, dz = stretched_levels(nk, depth)
    sig = 3.0 * (glo_ni / 360.0) if glo_ni != 360 else 3.0
    rng = np.random.default_rng(seed_bathy)
    h = _smooth(rng.standard_normal((glo_ni, glo_nj)), sig)
    h = (h - h.min()) / (h.max() - h.min()) * depth
    itn = np.sum(z[None, None, :] <= h[:, :, None], axis=2).astype(np.int32)
    land = (h < 50.0) | (np.abs(lat)[None, :] > polar_cut)
    itn[land] = 0
    tsurf = 15.0 + 12.5 * np.tanh((35.0 - np.abs(lat)) / 10.0)
    rng_t = np.random.default_rng(seed_t)
    decay = np.exp(-z / 1000.0)
    pt = 2.0 + (tsurf[None, :, None] - 2.0) * decay[None, None, :]
    pt = pt + 0.05 * rng_t.standard_normal((glo_ni, glo_nj, nk))
    pt = np.broadcast_to(pt, (glo_ni, glo_nj, nk)).copy()
    rng_s = np.random.default_rng(seed_s)
    s2 = _smooth(rng_s.uniform(size=(glo_ni, glo_nj)), sig)
    s2 = 34.0 + 2.0 * (s2 - s2.min()) / (s2.max() - s2.min())
    sa = np.broadcast_to(s2[:, :, None], (glo_ni, glo_nj, nk)).copy()
    wet = itn > 0
    tmn = np.array([pt[:, :, k][wet].mean() for k in range(nk)])
    smn = np.array([sa[:, :, k][wet].mean() for k in range(nk)])
    months = np.arange(12, dtype=np.float64)
    season = np.cos(2.0 * np.pi * months / 12.0)
    hemi = np.sign(lat)
    # trades and westerlies of 0.05 N m-2, tapered to zero on the equator:
    # int_ts steps T/S forward with centred advection (mod_int.f90:484,
    # mod_op.f90:280-325), stable only while ah >= u^2 dtbc / 2, so the
    # configuration keeps currents well below 1 m/s
    tx = -0.05 * np.cos(3.0 * np.pi * lat / 140.0) * (1.0 - np.exp(-(lat / 8.0) ** 2))
    taux = tx[None, :, None] * (1.0 + 0.2 * hemi[None, :, None] * season[None, None, :])
    taux = np.broadcast_to(taux, (glo_ni, glo_nj, 12)).copy()
    bct = tsurf[None, :, None] - 2.0 * hemi[None, :, None] * season[None, None, :]
    bct = np.broadcast_to(bct, (glo_ni, glo_nj, 12)).copy()
    bcs = np.broadcast_to(s2[:, :, None], (glo_ni, glo_nj, 12)).copy()
    zeros = np.zeros((glo_ni, glo_nj, 12))
    return Config(
        name="global", glo_ni=glo_ni, glo_nj=glo_nj, nk=nk,
        lon=_f(lon), lat=_f(lat), z=_f(z), itn=np.asfortranarray(itn),
        tmn=_f(tmn), smn=_f(smn), pt=_f(pt), sa=_f(sa),
        taux=_f(taux), tauy=_f(zeros), bct=_f(bct), bcs=_f(bcs),
        pa=_f(zeros), fw=_f(zeros.copy()), vmix=None,
        notes={"polar_cut_deg": polar_cut, "min_depth_m": 50.0,
               "seeds": (seed_bathy, seed_t, seed_s)},
    )


def small(glo_ni=8, glo_nj=6, nk=4, seed=7):
    """Tiny periodic-plus-walls grid with irregular bathymetry, for
    operator-level parity (every stagger, halo and wall case is hit)."""
    cfg = global_1deg(glo_ni=glo_ni, glo_nj=glo_nj, nk=nk, depth=2000.0,
                      polar_cut=90.0, seed_bathy=seed, seed_t=seed + 1,
                      seed_s=seed + 2)
    rng = np.random.default_rng(seed)
    itn = rng.integers(0, nk + 1, size=(glo_ni, glo_nj)).astype(np.int32)
    itn[rng.uniform(size=itn.shape) < 0.6] = nk
    cfg.itn = np.asfortranarray(itn)
    cfg.lat = _f(np.linspace(-50.0, 50.0, glo_nj))
    cfg.name = "small"
    return cfg


def eddy_025deg():
    """Eddy-permitting quarter-degree configuration, 1440 x 720 x 50: the
    quasi-global geometry, stratification and forcing of global_1deg at 0.25
    degrees (bathymetry smoothed over the same physical distance), with the
    steps the CFL limits there allow: dtbt = 15 s (gravity waves of about
    230 m/s over 9.5 km cells at the 70 degree cut) and dtbc = 900 s, 120
    barotropic substeps per step as at 1 degree."""
    cfg = global_1deg(glo_ni=1440, glo_nj=720, nk=50)
    cfg.dtbt, cfg.dtbc = 15, 900
    cfg.name = "eddy"
    return cfg


CONFIGS = {"box": box, "box_pa": box_pa, "global": global_1deg, "small": small, "eddy": eddy_025deg}


if __name__ == "__main__":
    import sys
    if len(sys.argv) != 3 or sys.argv[1] not in CONFIGS:
        print(f"usage: python -m pcom_amd.configs {{{'|'.join(CONFIGS)}}} OUT.bin  "
              "(raw inputs for the Fortran driver)", file=sys.stderr)
        sys.exit(2)
    CONFIGS[sys.argv[1]]().write_raw(sys.argv[2])
