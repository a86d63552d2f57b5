"""GPU parity: the HIP library (through its C ABI) against the FP64 oracle.

The FP64 form must reproduce the oracle bit for bit (compared as values:
+0.0 == -0.0) on every field, since both evaluate the reference's
statements with the same IEEE operations in the same order.
"""
import numpy as np
import pytest

from pcom_amd import abi, configs
from pcom_amd.lib import Model, PcomError, den_rho as gpu_den_rho

from oracle.oracle import Oracle, den_rho as orc_den_rho
from test_oracle_pins import unstable_box, worst_instability

pytestmark = pytest.mark.gpu

# fields that a step leaves defined (alpha / wm as int_ssh / int_ts leave them)
STEP_FIELDS = [
    "tc", "sc", "uc", "vc", "auc", "avc", "aup", "avp", "aupp", "avpp", "fx", "fy", "am",
    "alpha", "adp", "wm", "prho_t", "prho_s", "act", "acs", "acr", "acu", "acv", "acw",
    "chc", "chp", "mch", "mch2", "acch", "acssh", "buc", "bvc", "bup", "bvp", "but", "bvt",
    "pax", "pay", "bphi_xn", "bphi_xs", "bphi_ye", "bphi_yw", "bgra_xn", "bgra_xs", "bgra_ye",
    "bgra_yw", "bnd_taux", "bnd_tauy", "bnd_t", "bnd_s", "bnd_pa", "bnd_fw", "ssh",
]
STATIC_FIELDS = ["graphihx", "graphihy", "cor", "phih_t", "prh_t", "phih_u", "prh_u", "km", "kh"]


def _diff(name, a, b):
    """Describe a mismatch (value equality, NaN-aware)."""
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    if same.all():
        return None
    idx = np.argwhere(~same)
    first = tuple(int(x) + 1 for x in idx[0])  # 1-based like the reference
    d = np.abs(a - b)
    d = d[np.isfinite(d)]
    return (f"{name}: {int((~same).sum())} of {a.size} differ, first at {first} "
            f"gpu={a[tuple(idx[0])]!r} oracle={b[tuple(idx[0])]!r}, max|d|={d.max() if d.size else 'nan'}")


def _compare(gpu, orc, names):
    bad = []
    for n in names:
        r = _diff(n, gpu.get(n), orc.get(n))
        if r:
            bad.append(r)
    return bad


@pytest.fixture(scope="module")
def small_cfg():
    return configs.small()


def test_den_rho_bitwise(gpu_lib):
    rng = np.random.default_rng(0)
    t = rng.uniform(-2, 32, 4096)
    s = rng.uniform(0, 40, 4096)
    p = rng.uniform(0, 6e7, 4096)
    t[:3], s[:3], p[:3] = [3.0, 0.0, 25.0], [35.5, 0.0, 35.0], [300e5, 0.0, 0.0]
    np.testing.assert_array_equal(gpu_den_rho(t, s, p), orc_den_rho(t, s, p))


def test_static_fields(gpu_lib, small_cfg):
    m = Model(small_cfg)
    o = Oracle(small_cfg)
    bad = _compare(m, o, STATIC_FIELDS)
    for g in range(1, 5):
        for w in range(4):
            r = _diff(f"metric g{g} #{w}", m.metric(g, w), o.metric(g, w))
            if r:
                bad.append(r)
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("ga,gb", [(a, b) for a in range(1, 5) for b in range(1, 5) if a != b])
def test_op_ter(gpu_lib, small_cfg, ga, gb):
    m = Model(small_cfg)
    o = Oracle(small_cfg)
    rng = np.random.default_rng(10 * ga + gb)
    v = rng.standard_normal((m.ni, m.nj, 3))
    # inputs of the reference's operators are halo-swapped (periodic)
    v[0] = v[-2]
    v[-1] = v[1]
    for dft in (0.0, 1.0):
        np.testing.assert_array_equal(m.op_ter(ga, gb, v, dft), o.op_ter(ga, gb, v, dft))


@pytest.mark.parametrize("ga,gb", [(a, b) for a in range(1, 5) for b in range(1, 5)])
def test_op_px(gpu_lib, small_cfg, ga, gb):
    m = Model(small_cfg)
    o = Oracle(small_cfg)
    rng = np.random.default_rng(100 + 10 * ga + gb)
    v = rng.standard_normal((m.ni, m.nj, 2))
    v[0] = v[-2]
    v[-1] = v[1]
    np.testing.assert_array_equal(m.op_px1(ga, gb, v), o.op_px1(ga, gb, v))
    np.testing.assert_array_equal(m.op_px2(ga, gb, v), o.op_px2(ga, gb, v))


@pytest.mark.parametrize("ga,gb,gc", [(3, 3, 1), (4, 2, 3), (2, 4, 1)])
def test_op_div_lap(gpu_lib, small_cfg, ga, gb, gc):
    m = Model(small_cfg)
    o = Oracle(small_cfg)
    rng = np.random.default_rng(7)
    va = rng.standard_normal((m.ni, m.nj, 2))
    vb = rng.standard_normal((m.ni, m.nj, 2))
    for v in (va, vb):
        v[0] = v[-2]
        v[-1] = v[1]
    np.testing.assert_array_equal(m.op_div(ga, gb, gc, va, vb), o.op_div(ga, gb, gc, va, vb))
    np.testing.assert_array_equal(m.op_lap(3, 3, va[:, :, 0]), o.op_lap(3, 3, va[:, :, 0]))


@pytest.mark.parametrize("cfgname,nsteps", [("small", 1), ("small", 3), ("box", 2)])
def test_step_bitwise(gpu_lib, cfgname, nsteps):
    cfg = configs.CONFIGS[cfgname]()
    m = Model(cfg)
    o = Oracle(cfg)
    for s in range(nsteps):
        m.step(1)
        o.step(1)
        bad = _compare(m, o, STEP_FIELDS)
        assert not bad, f"after step {s + 1}:\n" + "\n".join(bad)
    cm, co = m.counters(), o.counters()
    for f in ("step", "n_ts", "n_uv", "n_w", "n_ch", "n_ssh", "year", "month", "day", "hour"):
        assert getattr(cm, f) == getattr(co, f), f


@pytest.mark.parametrize("tile", [(4, 4, 1), (5, 3, 3), (8, 8, 4), (24, 12, 2), (7, 7, 5), (8, 6, 5), (20, 14, 3),
                                  (7, 8, 2), (7, 8, 3), (16, 17, 3), (20, 13, 3)])
def test_subcycle_tiling_is_exact(gpu_lib, tile):
    """The LDS-tiled barotropic subcycle gives the reference's bits for any
    tile shape and any number of substeps between exchanges, also where the
    last column of tiles is narrower than a halo (7 x 8 x 3 on the box: 4
    columns against a halo of 6, which then reaches two tiles east)."""
    for cfgname in ("small", "box"):
        cfg = configs.CONFIGS[cfgname]()
        m = Model(cfg, trop_tile=tile)
        o = Oracle(cfg)
        m.step(2)
        o.step(2)
        bad = _compare(m, o, ["chc", "chp", "mch", "mch2", "acch", "buc", "bvc", "bup", "bvp", "tc", "uc"])
        assert not bad, f"{cfgname} tile {tile}:\n" + "\n".join(bad)


@pytest.mark.parametrize("tile", [(4, 4, 1), (5, 3, 3), (24, 12, 2), (7, 7, 5)])
def test_launch_mode_is_exact(gpu_lib, tile):
    """The subcycle in launch mode (one launch per S substeps, the kernel
    boundary being the exchange) gives the reference's bits."""
    for cfgname in ("small", "box"):
        cfg = configs.CONFIGS[cfgname]()
        m = Model(cfg, trop_tile=tile, trop_mode=abi.PCOM_TROP_LAUNCHES)
        assert m.subcycle_mode() == abi.PCOM_TROP_LAUNCHES
        o = Oracle(cfg)
        m.step(2)
        o.step(2)
        bad = _compare(m, o, ["chc", "chp", "mch", "mch2", "acch", "buc", "bvc", "bup", "bvp", "tc", "uc"])
        assert not bad, f"{cfgname} tile {tile}:\n" + "\n".join(bad)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_launch_mode_matches_persistent_at_1deg(gpu_lib, form):
    """Both modes run the same arithmetic: the exact one in FP64, the mixed
    forms' reciprocal divisions and hardware roots in mixed."""
    cfg = configs.global_1deg()
    a = Model(cfg, precision=abi.FORMS[form])
    assert a.subcycle_mode() == abi.PCOM_TROP_PERSISTENT  # the 1-degree grid fits
    b = Model(cfg, precision=abi.FORMS[form], trop_mode=abi.PCOM_TROP_LAUNCHES)
    a.step(3)
    b.step(3)
    for n in STEP_FIELDS:
        assert np.array_equal(a.get(n), b.get(n)), n


def test_large_grid_runs_in_launch_mode(gpu_lib):
    """A 0.25 degree grid (1440 x 720) has more tiles than the device holds
    co-resident: the automatic mode is launch mode, the persistent mode is
    refused, and the bits do not depend on the launch tiling."""
    cfg = configs.global_1deg(glo_ni=1440, glo_nj=720, nk=5)
    cfg.dtbt, cfg.dtbc = 15, 900
    with pytest.raises(PcomError):
        Model(cfg, trop_mode=abi.PCOM_TROP_PERSISTENT)
    a = Model(cfg)
    assert a.subcycle_mode() == abi.PCOM_TROP_LAUNCHES
    b = Model(cfg, trop_tile=(12, 10, 1))
    a.step(2)
    b.step(2)
    for n in ["chc", "chp", "mch", "mch2", "acch", "buc", "bvc", "tc", "uc"]:
        assert np.array_equal(a.get(n), b.get(n)), n
    assert a.check_finite()


def test_subcycle_wait_bound_reports_and_ends(gpu_lib):
    """A neighbour wait of the persistent subcycle past its bound (forced from
    the host: one unsuccessful poll ends a wait) stops its tile; the launch
    ends, the next synchronising call reports PCOM_ERR_STATE, and so does
    every later one.  A new context on the device then steps as the oracle
    does, bit for bit."""
    cfg = configs.box()
    m = Model(cfg)
    assert m.subcycle_mode() == abi.PCOM_TROP_PERSISTENT
    m.wait_limit(1)
    m.step(3)
    with pytest.raises(PcomError) as ei:
        m.sync()
    assert ei.value.code == abi.PCOM_ERR_STATE and "wait" in str(ei.value)
    with pytest.raises(PcomError) as ei:
        m.get("chc")
    assert ei.value.code == abi.PCOM_ERR_STATE
    m.close()
    fresh = Model(cfg)
    o = Oracle(cfg)
    fresh.step(2)
    o.step(2)
    bad = _compare(fresh, o, ["chc", "mch2", "buc", "bvc", "tc", "uc"])
    assert not bad, "\n".join(bad)


def test_box_many_steps_bitwise(gpu_lib):
    """The box basin (60x40x10): 30 steps crossing a day boundary."""
    cfg = configs.box()
    m = Model(cfg)
    o = Oracle(cfg)
    m.step(30)
    o.step(30)
    bad = _compare(m, o, STEP_FIELDS)
    assert not bad, "\n".join(bad)
    assert m.check_finite()


def _runs_of_four(t):
    """Columns holding 4 or more equal consecutive levels (mixed layers)."""
    t = t[1:-1, 1:-1, :]
    run4 = (t[:, :, :-3] == t[:, :, 1:-2]) & (t[:, :, 1:-2] == t[:, :, 2:-1]) & (t[:, :, 2:-1] == t[:, :, 3:])
    return int(run4.any(axis=2).sum())


def test_deep_convection_mixed_form(gpu_lib):
    """The mixed forms' convective adjustment (a mixed range's sums updated
    as it grows instead of summed again, values written once per range) on
    the unstable box: after one step and after two more, every column is
    statically stable under convect's own test (which the reference's
    search guarantees); each column's heat and salt content sum(T dpr),
    sum(S dpr), which mixing conserves, matches the oracle's to 1e-4 K and
    1e-4 psu (the FP32 tendencies differ by far less); the first step mixes
    the same deep layers, and almost every point matches the oracle to
    1e-3."""
    cfg = unstable_box()
    m = Model(cfg, precision=abi.PCOM_MIXED)
    o = Oracle(cfg)
    lev = o.igrid(2)
    dpr = o.vgrid(2, 3)
    sea = lev[:, :, None] >= np.arange(1, cfg.nk + 1)[None, None, :]
    sea[[0, -1], :, :] = False
    sea[:, [0, -1], :] = False
    col = sea.any(axis=2)
    for n in (1, 2):
        m.step(n)
        o.step(n)
        assert m.check_finite()
        worst, scale = worst_instability(m, lev)
        assert worst <= 1e-12 * scale, (n, worst, scale)
        w = np.where(sea, dpr[None, None, :], 0.0)
        for name, tol in (("tc", 1e-4), ("sc", 1e-4)):
            a, b = m.get(name), o.get(name)
            hm = np.sum(a * w, axis=2)[col] / np.sum(w, axis=2)[col]
            ho = np.sum(b * w, axis=2)[col] / np.sum(w, axis=2)[col]
            assert np.max(np.abs(hm - ho)) <= tol, (n, name, np.max(np.abs(hm - ho)))
            far = np.abs(a - b)[sea] > 1e-3
            assert far.mean() < 0.01, (n, name, far.mean())
        if n == 1:
            # (later steps re-mix layers that the tendencies left nearly
            # neutral, where the sign of an FP32 rounding decides)
            rm, ro = _runs_of_four(m.get("tc")), _runs_of_four(o.get("tc"))
            assert ro > 1000 and abs(rm - ro) <= 0.02 * ro, (rm, ro)
    # the oracle's own columns pass the same stability test
    worst, scale = worst_instability(o, lev)
    assert worst <= 1e-12 * scale, (worst, scale)


def test_deep_convection_bitwise(gpu_lib):
    """Box columns made statically unstable to seeded random depths (water
    warmed by up to 8 K towards the bottom in 60% of the columns, salinity
    noise of +-0.5 psu): convect then mixes layers that deepen downwards and
    upwards many times over, and the result stays bit-identical to the
    oracle after one step and after two more in one call (whose later steps
    start from what the earlier ones formed)."""
    cfg = unstable_box()
    m = Model(cfg)
    o = Oracle(cfg)
    m.step(1)
    o.step(1)
    bad = _compare(m, o, STEP_FIELDS)
    assert not bad, "\n".join(bad)
    # the adjustment did mix deep layers: runs of 4 or more equal levels
    assert _runs_of_four(m.get("tc")) > 100
    m.step(2)
    o.step(2)
    bad = _compare(m, o, STEP_FIELDS)
    assert not bad, "\n".join(bad)
    assert m.check_finite()


def test_single_step_calls_and_set_field_bitwise(gpu_lib):
    """One step per pcom_step call (the Fortran driver's loop), with field
    reads between the calls (each joins the work a step leaves queued),
    then pcom_set_field of T and uc (which must drop what the last step
    prepared for the next) and more single steps: bit-identical to the
    oracle throughout."""
    cfg = configs.box()
    m = Model(cfg)
    o = Oracle(cfg)
    names = ["tc", "sc", "uc", "vc", "wm", "alpha", "adp", "act", "acw", "acr", "chc", "ssh"]
    for _ in range(4):
        m.step(1)
        o.step(1)
        bad = _compare(m, o, names)
        assert not bad, "\n".join(bad)
    rng = np.random.default_rng(11)
    t = o.get("tc")
    t = t + rng.uniform(-0.5, 0.5, t.shape) * (t != 0)
    u = o.get("uc") * 0.9
    # the oracle keeps tc/tp and uc/up apart (the library holds one buffer
    # for each pair, which the steps leave equal): set both names
    for x in (m, o):
        for name, v in (("tc", t), ("tp", t), ("uc", u), ("up", u)):
            x.set(name, v)
    for _ in range(3):
        m.step(1)
        o.step(1)
    bad = _compare(m, o, STEP_FIELDS)
    assert not bad, "\n".join(bad)


def test_global_1deg_bitwise_across_a_day(gpu_lib):
    """The full 1-degree workload (360x180x30): 25 steps, crossing the day
    boundary (daily forcing and den_prho), bit-identical to the oracle."""
    cfg = configs.global_1deg()
    m = Model(cfg)
    o = Oracle(cfg)
    m.step(25)
    o.step(25)
    bad = _compare(m, o, ["tc", "sc", "uc", "vc", "chc", "mch", "buc", "bvc", "am", "ssh", "acch", "act",
                          "acu", "acw", "acr", "wm", "fx", "auc", "bphi_xn", "bgra_yw"])
    assert not bad, "\n".join(bad)


def test_two_contexts_share_the_device(gpu_lib):
    """Two 1-degree contexts stepping on their own streams at once (each
    subcycle launch needs most of the device's CUs; launches are chained
    across contexts) give the same bits as one context alone."""
    cfg = configs.global_1deg()
    a, b = Model(cfg), Model(cfg, precision=abi.PCOM_MIXED)
    for _ in range(12):
        a.step(1)
        b.step(1)
    solo = Model(cfg)
    solo.step(12)
    bad = _compare(a, solo, ["tc", "uc", "chc", "buc", "mch"])
    assert not bad, "\n".join(bad)
    assert b.check_finite()
