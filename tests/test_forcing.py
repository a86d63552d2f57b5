"""The forcing path at the boundary (main.f90:158-202, mod_int.f90:90-103):

* caller forcing records (the reference's daily and 6-hourly forcing files)
  applied through pcom_set_forcing in the reference's order: int_bnd, the
  records, den_prho;
* int_readyc's atmospheric pressure gradient follows every change of bnd
  pa, whether from the daily refresh, a caller's record, pcom_set_field or
  a restart file read mid-day.

CPU tests check the oracle's side of the hook; GPU tests compare the HIP
library with the oracle bit for bit, driven the same way."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from pcom_amd import configs, restart
from test_gpu_parity import STEP_FIELDS, _compare


def _records(cfg, seed):
    rng = np.random.default_rng(seed)
    shp = (cfg.glo_ni, cfg.glo_nj)
    x = np.linspace(0.0, 2.0 * np.pi, cfg.glo_ni)[:, None]
    y = np.linspace(-1.0, 1.0, cfg.glo_nj)[None, :]
    return {
        "taux": 0.08 * np.cos(np.pi * y / 2.0) * (1.0 + 0.2 * rng.uniform(size=shp)),
        "tauy": 0.01 * np.sin(x) * rng.uniform(size=shp),
        "pa": 300.0 * np.sin(x + rng.uniform(0.0, 6.0)) * np.cos(np.pi * y / 2.0),
        "bct": 15.0 + 5.0 * y + rng.uniform(-0.5, 0.5, size=shp),
    }


def _drive(x, cfg, nsteps, six_hourly=True, keys=("taux", "tauy", "pa", "bct")):
    """The Fortran driver's loop with forcing files: before each step, the
    daily records when the day changes, then the 6-hourly ones."""
    for i in range(nsteps):
        daily, six = x.forcing_due()
        if daily:
            r = _records(cfg, 1000 + i)
            x.set_forcing(True, **{k: r[k] for k in keys})
        if six_hourly and six:
            r = _records(cfg, 2000 + i)
            x.set_forcing(False, pa=r["pa"], bct=r["bct"])
        x.step(1)


def test_oracle_forcing_records_replace_interior_and_swap():
    """A record lands on the interior points, the E/W halo columns are the
    periodic copies, and the wall rows keep int_bnd's values; a 6-hourly
    record before a due daily refresh is refused."""
    cfg = configs.box_pa()
    o = Oracle(cfg)
    r = _records(cfg, 5)
    with pytest.raises(ValueError):
        o.set_forcing(False, pa=r["pa"])
    o.set_forcing(True)
    before = o.get("bnd_pa")
    o.set_forcing(False, pa=r["pa"])
    pa = o.get("bnd_pa")
    np.testing.assert_array_equal(pa[1:-1, 1:-1], r["pa"])
    np.testing.assert_array_equal(pa[0], pa[-2])
    np.testing.assert_array_equal(pa[-1], pa[1])
    np.testing.assert_array_equal(pa[1:-1, [0, -1]], before[1:-1, [0, -1]])
    # den_prho ran again after the record (T, S unchanged: same values)
    assert np.any(o.get("prho_t") != 0.0)


def test_oracle_daily_refresh_taken_by_caller_is_not_repeated():
    """After pcom_set_forcing(interpolate=1) the step does not run its own
    int_bnd (which would overwrite the records)."""
    cfg = configs.box_pa()
    o = Oracle(cfg)
    r = _records(cfg, 6)
    o.set_forcing(True, taux=r["taux"])
    o.step(1)
    np.testing.assert_array_equal(o.get("bnd_taux")[1:-1, 1:-1], r["taux"])


@pytest.mark.gpu
def test_gpu_forcing_hook_bitwise(gpu_lib):
    """30 hourly steps from midnight (two daily refreshes, six 6-hourly
    ones) driven by caller records on both: every field bit for bit."""
    from pcom_amd.lib import Model
    cfg = configs.box_pa()
    m, o = Model(cfg), Oracle(cfg)
    print("forcing hook: box_pa, 30 steps", flush=True)
    _drive(m, cfg, 30)
    _drive(o, cfg, 30)
    bad = _compare(m, o, STEP_FIELDS)
    assert not bad, "\n".join(bad)
    assert np.abs(m.get("pax")).max() > 0.0


@pytest.mark.gpu
def test_gpu_set_bnd_pa_refreshes_gradient(gpu_lib):
    """A caller's set of bnd pa mid-day changes int_readyc's pressure force
    from the next step on, as the reference's every-step gradient does."""
    from pcom_amd.lib import Model
    cfg = configs.box_pa()
    m, o = Model(cfg), Oracle(cfg)
    print("set bnd_pa: box_pa", flush=True)
    m.step(3)
    o.step(3)
    r = _records(cfg, 9)
    pa = m.get("bnd_pa")
    pa[1:-1, 1:-1] = r["pa"]
    pa[0], pa[-1] = pa[-2], pa[1]
    m.set("bnd_pa", pa)
    o.set("bnd_pa", pa)
    m.step(2)
    o.step(2)
    bad = _compare(m, o, STEP_FIELDS)
    assert not bad, "\n".join(bad)


@pytest.mark.gpu
def test_gpu_mid_day_restart_past_midnight(gpu_lib, tmp_path):
    """A restart file written at 18:00 (its bnd fields include a nonzero
    pa) continues on the GPU past midnight exactly as the oracle's restart
    run does, and as the uninterrupted run does."""
    from pcom_amd.lib import Model
    cfg = configs.box_pa()
    a = Model(cfg)
    print("mid-day restart: box_pa, 18 + 12 steps", flush=True)
    a.step(18)
    c = a.counters()
    cd = (c.year, c.month, c.day, c.hour, c.minute, c.second)
    pd = (c.year, c.month, c.day, c.hour - 1, c.minute, c.second)
    assert c.hour == 18
    path = str(tmp_path / "rst.bin")
    restart.write(path, a, cd, pd, c.step)
    b = Model(cfg, restart_run=1)
    restart.load(path, b)
    o = Oracle(cfg, restart_run=1)
    restart.load(path, o)
    for x in (a, b, o):
        x.step(12)
    assert np.abs(o.get("bnd_pa")).max() > 0.0
    bad = _compare(b, o, STEP_FIELDS)
    assert not bad, "\n".join(bad)
    for n in ("tc", "sc", "uc", "vc", "chc", "buc", "pax"):
        np.testing.assert_array_equal(b.get(n)[1:-1, 1:-1], a.get(n)[1:-1, 1:-1], err_msg=f"restart vs run {n}")
