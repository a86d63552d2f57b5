"""The Fortran driver (pcom_amd/fortran) through its ISO_C_BINDING layer.

CPU: the interoperable types have the C structs' sizes, and the driver
reaches the library (which reports that no GPU is visible).  GPU: one day of
the box configuration run by the Fortran driver gives the same time means,
bit for bit, as the same day run through the Python binding, and a restart
run of the driver continues exactly like an uninterrupted run
(the reference's tools/debug_restart.sh procedure); with forcing files and
6-hourly output the driver's records reach the library as the same records
handed over through the Python binding do.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from pcom_amd import abi, configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "pcom_amd", "fortran", "pcom_main")

NAMELIST = """&time_ctrl
bdate = '{bdate}',
edate = '{edate}',
dtbt = 60,
dtbc = 3600,
/
&io_ctrl
fname_ini = '{ini}',
fname_rst = '{rst}',
out_dir = '{out}/',
out_per = '{per}',
restart_per = 1,
restart_run = {restart},
{ff}/
"""


def _driver():
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (make -C pcom_amd/fortran, or __graft_entry__.build())")
    return DRIVER


def _namelist(tmp, bdate, edate, restart=0, rst="", per="day", ff="", cfg=None):
    cfg = cfg or configs.box()
    ini = os.path.join(tmp, f"{cfg.name}.bin")
    if not os.path.exists(ini):
        cfg.write_raw(ini)
    path = os.path.join(tmp, f"namelist_{restart}")
    with open(path, "w") as f:
        f.write(NAMELIST.format(bdate=bdate, edate=edate, ini=ini, rst=rst, out=tmp, restart=restart, per=per,
                                ff=ff))
    return path


def _read_means(path, g_ni, g_nj, nk):
    with open(path, "rb") as f:
        head = np.frombuffer(f.read(4 * 9), dtype="<i4")
        assert tuple(head[:3]) == (g_ni, g_nj, nk)
        n3, n2 = g_ni * g_nj * nk, g_ni * g_nj
        m3 = np.frombuffer(f.read(8 * 6 * n3), dtype="<f8").reshape((g_ni, g_nj, nk, 6), order="F")
        m2 = np.frombuffer(f.read(8 * 2 * n2), dtype="<f8").reshape((g_ni, g_nj, 2), order="F")
    out = {n: m3[..., q] for q, n in enumerate(["pt", "sa", "rho", "u", "v", "w"])}
    out["ch"], out["ssh"] = m2[..., 0], m2[..., 1]
    return tuple(int(x) for x in head[3:]), out


def test_driver_abi_sizes():
    out = subprocess.run([_driver(), "--abi"], capture_output=True, text=True, timeout=60)
    sizes = [int(x) for x in out.stdout.split()[1:6]]
    assert sizes == [C.sizeof(abi.PcomInputs), C.sizeof(abi.PcomOpts), C.sizeof(abi.PcomCounters),
                     C.sizeof(abi.PcomMeans), C.sizeof(abi.PcomForcing)]


def test_driver_reaches_library_without_gpu(tmp_path):
    """Without a GPU the driver's pcom_create call returns PCOM_ERR_NOGPU
    and the driver stops with the library's message."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    nml = _namelist(str(tmp_path), "0001-01-01 00:00:00", "0001-01-02 00:00:00")
    out = subprocess.run([_driver(), nml], capture_output=True, text=True, timeout=120)
    assert out.returncode != 0
    assert "pcom_create failed (-5): no HIP device visible" in out.stdout


@pytest.mark.gpu
def test_driver_day_and_restart_match_python(gpu_lib, tmp_path):
    from pcom_amd.lib import Model
    tmp = str(tmp_path)
    # day 1 from the driver, writing means and a restart file at its end
    nml = _namelist(tmp, "0001-01-01 00:00:00", "0001-01-02 00:00:00")
    r = subprocess.run([_driver(), nml], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "integrated to 0001-01-02 00:00:00" in r.stdout
    cfg = configs.box()
    m = Model(cfg)
    m.step(24)
    ref1 = m.output_means()
    date, got1 = _read_means(os.path.join(tmp, "means_0001-01-02_00-00-00.bin"), 60, 40, 10)
    assert date == (1, 1, 2, 0, 0, 0)
    for n in ref1:
        np.testing.assert_array_equal(got1[n], ref1[n], err_msg=f"day 1 {n}")
    # day 2 as a restart run of the driver, against the uninterrupted run
    rst = os.path.join(tmp, "restart_0001-01-02_00-00-00.bin")
    assert os.path.exists(rst)
    nml2 = _namelist(tmp, "0001-01-01 00:00:00", "0001-01-03 00:00:00", restart=1, rst=rst)
    r = subprocess.run([_driver(), nml2], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "restart-run" in r.stdout
    m.step(24)
    ref2 = m.output_means()
    date, got2 = _read_means(os.path.join(tmp, "means_0001-01-03_00-00-00.bin"), 60, 40, 10)
    assert date == (1, 1, 3, 0, 0, 0)
    for n in ref2:
        np.testing.assert_array_equal(got2[n], ref2[n], err_msg=f"day 2 (restart) {n}")


def _forcing_files(tmp, cfg, nrec):
    """nrec records of each forcing file (raw float64, Fortran order):
    ff_ws holds taux then tauy per record."""
    rng = np.random.default_rng(17)
    shp = (cfg.glo_ni, cfg.glo_nj)
    y = np.linspace(-1.0, 1.0, cfg.glo_nj)[None, :]
    rec = {"taux": [0.1 * np.cos(np.pi * y / 2.0) * (1.0 + 0.3 * rng.uniform(size=shp)) for _ in range(nrec)],
           "tauy": [0.02 * rng.uniform(-1.0, 1.0, size=shp) for _ in range(nrec)],
           "pa": [400.0 * rng.uniform(-1.0, 1.0) * np.cos(np.pi * y / 2.0) * np.ones(shp) for _ in range(nrec)],
           "bct": [12.0 + 6.0 * y + rng.uniform(-1.0, 1.0, size=shp) for _ in range(nrec)]}
    paths = {k: os.path.join(tmp, f"ff_{k}.bin") for k in ("ws", "pa", "bct")}
    with open(paths["ws"], "wb") as f:
        for r in range(nrec):
            f.write(np.asfortranarray(rec["taux"][r]).tobytes(order="F"))
            f.write(np.asfortranarray(rec["tauy"][r]).tobytes(order="F"))
    for k in ("pa", "bct"):
        with open(paths[k], "wb") as f:
            for r in range(nrec):
                f.write(np.asfortranarray(rec[k][r]).tobytes(order="F"))
    ff = "".join(f"ff_{k} = '{p}',\n" for k, p in paths.items())
    return rec, ff


@pytest.mark.gpu
def test_driver_forcing_files_6hour_match_python(gpu_lib, tmp_path):
    """30 hours of box_pa by the driver with ff_ws, ff_pa and ff_bct and
    out_per = '6hour' (main.f90:164-202, 228-236): its five 6-hourly means
    equal, bit for bit, those of the Python binding handed record nday on
    each daily refresh and record (nday-1)*4 + hour/6 + 1 every 6 hours."""
    from pcom_amd.lib import Model
    tmp = str(tmp_path)
    cfg = configs.box_pa()
    rec, ff = _forcing_files(tmp, cfg, 8)
    nml = _namelist(tmp, "0001-01-01 00:00:00", "0001-01-02 06:00:00", per="6hour", ff=ff, cfg=cfg)
    r = subprocess.run([_driver(), nml], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    m = Model(cfg)
    nday = 0
    ref = {}
    for i in range(30):
        daily, six = m.forcing_due()
        if daily:
            nday += 1
            m.set_forcing(True, **{k: rec[k][nday - 1] for k in rec})
        if six:
            q = (nday - 1) * 4 + m.counters().hour // 6
            m.set_forcing(False, **{k: rec[k][q] for k in rec})
        m.step(1)
        c = m.counters()
        if c.hour % 6 == 0:
            ref[(c.day, c.hour)] = m.output_means()
    assert nday == 2 and len(ref) == 5
    for (day, hour), want in ref.items():
        date, got = _read_means(os.path.join(tmp, f"means_0001-01-{day:02d}_{hour:02d}-00-00.bin"), 60, 40, 10)
        assert date == (1, 1, day, hour, 0, 0)
        for n in want:
            np.testing.assert_array_equal(got[n], want[n], err_msg=f"{day} {hour}h {n}")
