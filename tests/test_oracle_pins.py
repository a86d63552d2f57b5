"""Pins the FP64 oracle's remaining routines against what the mathematics or
the reference's own definitions fix, independently of how the oracle
evaluates them:

* smag (mod_int.f90:403-440): closed forms of the Smagorinsky viscosity for
  a zonal jet U cos(lat) plus a meridional flow linear in longitude, its
  cap (dx^2 + dy^2) / (32 dtbc), the rows it writes and the halo swap;
* den_prho (mod_den.f90:120-153): the centred differences are the partial
  derivatives of den_rho (pinned by the JM95 / UNESCO check values) at the
  layer-centre pressure chc pr(k), to their O(delta^2) truncation;
* den_alpha (mod_den.f90:33-55): alpha chc dpr(k) is the hydrostatic
  volume of layer k, the integral of 1/rho over the layer's actual
  pressure range [chc pr1(k), chc pr1(k+1)];
* int_bnd (mod_int.f90:649-691): monthly fields anchored on the 15th,
  linear in the day between anchors over the length of the earlier month,
  through the December-January wrap and the Gregorian leap rule, with a
  weight that is a float32 number;
* int_readyc (mod_int.f90:90-151): at rest without wind, the column
  tendency is the atmospheric pressure force -a0 sqrt(chc) grad(pa) of a
  pressure linear in longitude, and the Adams-Bashforth history shifts
  with its bootstrap on steps 1 and 2 unless the run is a restart;
* int_ts's accumulators (mod_int.f90:456-461, 500-504): acr adds the
  density of the new T, S at the layer-centre pressure, acw the pressure
  velocity converted to height units with the step-mean mch2;
* convect at the end of int_ts (mod_int.f90:509-542): a step from columns
  made statically unstable leaves every column stable under convect's own
  test with the step's prho.

The GPU kernels match the oracle bit for bit (tests/test_gpu_parity.py), so
these pins carry over to them."""
import numpy as np
import pytest

from oracle.oracle import Oracle, constants, den_rho
from pcom_amd import configs

C = constants()
A_EARTH = C["a"]


def _periodic(v):
    v[0] = v[-2]
    v[-1] = v[1]
    return v


def _channel(glo_ni, glo_nj, nk=3, lat0=-60.0, lat1=60.0):
    """All-sea zonally periodic channel between walls at lat0 and lat1."""
    cfg = configs.box(glo_ni=glo_ni, glo_nj=glo_nj, nk=nk)
    cfg.lat = np.asfortranarray(lat0 + (np.arange(glo_nj) + 0.5) * ((lat1 - lat0) / glo_nj))
    cfg.lon = np.asfortranarray((np.arange(glo_ni) + 0.5) * (360.0 / glo_ni))
    cfg.itn = np.asfortranarray(np.full((glo_ni, glo_nj), nk, dtype=np.int32))
    return cfg


def _unwrapped_lon(o, g):
    """Longitudes (radians) of stagger g along i, continued through the
    halo columns without the periodic jump."""
    lon = o.metric(g, 4)[:, 2].copy()
    d = lon[2] - lon[1]
    return lon[1] + d * (np.arange(o.ni) - 1)


# --------------------------------------------------------------------------
# smag (mod_int.f90:403-440)

def _smag_expected(o, U, V):
    """The reference's Smagorinsky formula evaluated from the coordinates
    alone for u = U cos(lat), v = V lon (g3): centred differences over two
    cells of a uniform grid are exact for v and are the finite differences
    of cos(lat) for u; the tension and shear carry the reference's factor 4
    and its metric terms +v tan(lat)/a and -u tan(lat)/a."""
    lam = _unwrapped_lon(o, 3)
    phi = o.metric(3, 5)[2, :]
    ni, nj = o.ni, o.nj
    am = np.full((ni, nj), np.nan)
    for j in range(1, nj - 1):
        dx = A_EARTH * np.cos(phi[j]) * (lam[2:] - lam[:-2])        # i = 1..ni-2
        dy = A_EARTH * (phi[j + 1] - phi[j - 1])
        v = V * lam[1:-1]
        u = U * np.cos(phi[j])
        dt = v * np.tan(phi[j]) / A_EARTH
        ds = 4 * V * (lam[2:] - lam[:-2]) / dx + 4 * U * (np.cos(phi[j + 1]) - np.cos(phi[j - 1])) / dy \
            - u * np.tan(phi[j]) / A_EARTH
        c = (0.5 * C["smag_c"] / C["pi"] * (0.5 * dx + 0.5 * dy)) ** 2
        cap = (dx * dx + dy * dy) / (32.0 * o.cfg.dtbc)
        am[1:-1, j] = np.minimum(c * np.sqrt(dt * dt + ds * ds), cap)
    return am


def _smag_run(U, V, nk=2):
    cfg = _channel(24, 16, nk=nk)
    o = Oracle(cfg)
    lam = _unwrapped_lon(o, 3)
    phi = o.metric(3, 5)[2, :]
    ni, nj = o.ni, o.nj
    u = np.broadcast_to((U * np.cos(phi))[None, :, None], (ni, nj, nk)).copy()
    v = np.broadcast_to((V * lam)[:, None, None], (ni, nj, nk)).copy()
    # halo columns are the periodic copies the reference's swap leaves
    u, v = _periodic(u), _periodic(v)
    sentinel = np.full((ni, nj, nk), 7.25)
    o.set("am", sentinel)
    return o, o.smag(u, v)


def test_smag_closed_form_jet_and_shear():
    """u = U cos(lat) and v = V lon give, at every point whose i-1, i+1
    neighbours do not straddle the periodic seam, the closed form of the
    reference's tension/shear formula (not capped at these speeds)."""
    o, am = _smag_run(U=0.3, V=0.02)
    want = _smag_expected(o, 0.3, 0.02)
    ni, nj = o.ni, o.nj
    cols = slice(2, ni - 2)         # i-1 and i+1 both interior and unwrapped
    rows = slice(1, nj - 1)
    for k in range(am.shape[2]):
        np.testing.assert_allclose(am[cols, rows, k], want[cols, rows], rtol=1e-12)
    cap = (A_EARTH * 0.13) ** 2 / (32 * o.cfg.dtbc)
    assert np.all(want[cols, rows] < cap)   # the formula, not the cap, is tested


def test_smag_cap_rows_and_swap():
    """At 1e6 m/s the viscosity is (dx^2 + dy^2) / (4*8*dtbc) everywhere it
    is computed (rows 2..nj-1); wall rows 1 and nj keep what they held; the
    halo columns are the periodic copies of columns ni-1 and 2."""
    o, am = _smag_run(U=1.0e6, V=0.0)
    lam = _unwrapped_lon(o, 3)
    phi = o.metric(3, 5)[2, :]
    ni, nj = o.ni, o.nj
    for j in range(1, nj - 1):
        dx = A_EARTH * np.cos(phi[j]) * (lam[2:] - lam[:-2])
        dy = A_EARTH * (phi[j + 1] - phi[j - 1])
        np.testing.assert_allclose(am[1:-1, j, 0], (dx * dx + dy * dy) / (32.0 * o.cfg.dtbc), rtol=1e-13)
    assert np.all(am[:, 0, :] == 7.25) and np.all(am[:, nj - 1, :] == 7.25)
    np.testing.assert_array_equal(am[0], am[ni - 2])
    np.testing.assert_array_equal(am[ni - 1], am[1])


# --------------------------------------------------------------------------
# den_prho (mod_den.f90:120-153) and den_alpha (mod_den.f90:33-55)

def _random_state(o, rng, chc_spread=0.08):
    ni, nj, nk = o.ni, o.nj, o.nk
    t = rng.uniform(-1.0, 28.0, size=(ni, nj, nk))
    s = rng.uniform(33.0, 37.0, size=(ni, nj, nk))
    chc = 1.0 + chc_spread * rng.uniform(-0.5, 1.0, size=(ni, nj))
    return t, s, chc


def _derivative(f, x, h):
    """Richardson-extrapolated central difference: O(h^4) truncation."""
    d1 = (f(x + h) - f(x - h)) / (2 * h)
    d2 = (f(x + h / 2) - f(x - h / 2)) / h
    return (4 * d2 - d1) / 3


def test_den_prho_is_the_partial_derivative_at_layer_pressure():
    """prho = (d rho / dT, d rho / dS) at p = chc pr(k) on sea points: the
    reference's centred differences with deltat = 0.01 K and deltas = 0.001
    agree with the true derivatives of the pinned den_rho to their O(delta^2)
    truncation (< 1e-7 kg m-3 per unit).  Land keeps what it held."""
    cfg = configs.small()
    o = Oracle(cfg)
    rng = np.random.default_rng(41)
    t, s, chc = _random_state(o, rng)
    o.set("tc", t)
    o.set("sc", s)
    o.set("chc", chc)
    o.set("prho_t", np.full(t.shape, 123.0))
    o.set("prho_s", np.full(t.shape, -45.0))
    o.run_stage("den_prho")
    pt, ps = o.get("prho_t"), o.get("prho_s")
    wet = o.igrid(0) > 0
    pr2 = o.vgrid(2, 1)
    p = chc[:, :, None] * pr2[None, None, :]
    dt = _derivative(lambda x: den_rho(x, s, p), t, 0.02)
    ds = _derivative(lambda x: den_rho(t, x, p), s, 0.004)
    assert np.abs(pt - dt)[wet].max() < 1e-7
    assert np.abs(ps - ds)[wet].max() < 1e-7
    assert (pt[wet] < 0.0).mean() > 0.9 and np.all(ps[wet] > 0.7)   # expansion, haline contraction
    assert np.all(pt[~wet] == 123.0) and np.all(ps[~wet] == -45.0)


def test_den_alpha_is_hydrostatic_layer_volume():
    """alpha(k) chc dpr(k) equals the integral of 1/rho(T, S, p) dp over the
    layer's actual pressure range [chc pr1(k), chc pr1(k+1)] (5-point
    Gauss-Legendre) to the midpoint rule's O(dp^2 alpha'') error, well under
    the 1e-3 relative error of evaluating at an interface pressure or
    without the chc scaling.  Land keeps what it held."""
    cfg = configs.global_1deg(glo_ni=24, glo_nj=12, nk=30)
    o = Oracle(cfg)
    rng = np.random.default_rng(43)
    t, s, chc = _random_state(o, rng)
    o.set("tc", t)
    o.set("sc", s)
    o.set("chc", chc)
    o.set("alpha", np.full(t.shape, 5.0))
    o.run_stage("den_alpha_main")
    alpha = o.get("alpha")
    wet = o.igrid(0) > 0
    pr1 = o.vgrid(1, 1)
    dpr = o.vgrid(2, 3)
    x, w = np.polynomial.legendre.leggauss(5)
    want = np.zeros_like(alpha)
    for k in range(o.nk):
        lo, hi = chc * pr1[k], chc * pr1[k + 1]
        for xq, wq in zip(x, w):
            pq = 0.5 * (lo + hi) + 0.5 * (hi - lo) * xq
            want[:, :, k] += 0.5 * wq * (hi - lo) / den_rho(t[:, :, k], s[:, :, k], pq)
    got = alpha * chc[:, :, None] * dpr[None, None, :]
    rel = np.abs(got - want) / want
    # layers up to 3.6e6 Pa thick: midpoint error ~ 1e-6, an interface
    # pressure or a missing chc ~ 1e-3
    assert rel[wet].max() < 2e-5
    assert np.all(alpha[~wet] == 5.0)
    assert dpr.max() > 3e6 and (chc.max() - chc.min()) > 0.05


# --------------------------------------------------------------------------
# int_bnd (mod_int.f90:649-691)

def _days_in_month(y, m):
    leap = (y % 4 == 0 and y % 100 != 0) or (y % 400 == 0)
    return [31, 29 if leap else 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31][m - 1]


def _bnd_oracle():
    """small config whose monthly fields alternate 0 and 1 (even and odd
    months) for bct, scaled patterns for the rest: the interpolation weight
    f is then read back exactly as bnd_t (or 1 - bnd_t)."""
    cfg = configs.small()
    shp = (cfg.glo_ni, cfg.glo_nj, 12)
    rng = np.random.default_rng(5)
    pat = rng.uniform(1.0, 2.0, size=shp[:2])
    months = np.arange(1, 13, dtype=np.float64)
    cfg.bct = np.asfortranarray(np.broadcast_to((months % 2 == 0).astype(float)[None, None, :], shp).copy())
    cfg.taux = np.asfortranarray(pat[:, :, None] * (months ** 2)[None, None, :])
    cfg.tauy = np.asfortranarray(pat[:, :, None] * np.sin(months)[None, None, :])
    cfg.bcs = np.asfortranarray(34.0 + pat[:, :, None] * np.cos(months)[None, None, :])
    cfg.pa = np.asfortranarray(pat[:, :, None] * 100.0 * months[None, None, :])
    cfg.fw = np.asfortranarray(pat[:, :, None] * 1e-6 * np.sqrt(months)[None, None, :])
    return cfg, Oracle(cfg)


def _bnd_at(o, y, m, d):
    o.set_clock(0, (y, m, d, 0, 0, 0), (y, m, d, 0, 0, 0))
    o.run_stage("int_bnd")
    return {n: o.get("bnd_" + n)[1:-1, 1:-1] for n in ("taux", "tauy", "t", "s", "pa", "fw")}


def _weight(y, m, d):
    """(earlier month, later month, days since the earlier 15th, its month
    length): forcing is a mid-month value, interpolated linearly."""
    if d >= 15:
        m1, m2 = m, m % 12 + 1
        return m1, m2, d - 15, _days_in_month(y, m1)
    m1 = 12 if m == 1 else m - 1
    return m1, m, _days_in_month(y, m1) - 15 + d, _days_in_month(y, m1)


@pytest.mark.parametrize("ymd", [(1, 1, 1), (1, 1, 14), (1, 1, 15), (1, 2, 28), (1, 3, 10), (4, 3, 10),
                                 (100, 3, 10), (400, 3, 10), (1, 7, 31), (1, 12, 15), (1, 12, 20),
                                 (1, 12, 31), (3, 1, 10)])
def test_int_bnd_interpolates_between_mid_month_values(ymd):
    """bnd = F(m1) + f (F(m2) - F(m1)) with f = (days since the 15th of m1)
    / (length of m1): exactly F(m) on the 15th; f is a float32 number (the
    reference's default-REAL literals) within one float32 ulp of the exact
    ratio; December wraps to January; February has 29 days in Gregorian
    leap years (4, 400) and 28 otherwise (1, 100)."""
    cfg, o = _bnd_oracle()
    y, m, d = ymd
    b = _bnd_at(o, y, m, d)
    m1, m2, n, L = _weight(y, m, d)
    f_exact = n / L
    # bct alternates 0 / 1: the weight itself, or its complement
    t1 = float(m1 % 2 == 0)
    t2 = float(m2 % 2 == 0)
    f = (b["t"] - t1) / (t2 - t1)
    assert np.all(f == f.flat[0])
    f = f.flat[0]
    assert f == float(np.float32(f)), "the weight is evaluated in float32"
    assert abs(f - f_exact) <= np.spacing(np.float32(max(f_exact, 1e-30)))
    if d == 15:
        assert f == 0.0
    F = {"taux": cfg.taux, "tauy": cfg.tauy, "s": cfg.bcs, "pa": cfg.pa, "fw": cfg.fw}
    for name, arr in F.items():
        a1, a2 = arr[:, :, m1 - 1], arr[:, :, m2 - 1]
        np.testing.assert_allclose(b[name], a1 + f * (a2 - a1), rtol=1e-14, atol=1e-14 * np.abs(a1).max())
        if d == 15:
            np.testing.assert_array_equal(b[name], a1)


def test_int_bnd_is_linear_between_anchors():
    """Day by day from the 15th of one month to the 15th of the next, bnd
    advances in equal steps of (F(m2) - F(m1)) / (length of m1), across the
    month boundary, in a leap-year February too."""
    for y, m in ((1, 1), (4, 2), (1, 2), (1, 12)):
        cfg, o = _bnd_oracle()
        L = _days_in_month(y, m)
        m2 = m % 12 + 1
        y2 = y + 1 if m == 12 else y
        days = [(y, m, d) for d in range(15, L + 1)] + [(y2, m2, d) for d in range(1, 16)]
        vals = np.array([_bnd_at(o, *ymd)["taux"] for ymd in days])
        step = (cfg.taux[:, :, m2 - 1] - cfg.taux[:, :, m - 1]) / L
        np.testing.assert_allclose(np.diff(vals, axis=0), np.broadcast_to(step, vals[1:].shape),
                                   rtol=2e-6, atol=1e-6 * np.abs(cfg.taux).max())
        np.testing.assert_array_equal(vals[-1], cfg.taux[:, :, m2 - 1])
        assert len(days) == L + 1


# --------------------------------------------------------------------------
# int_readyc (mod_int.f90:90-151)

def test_int_readyc_pressure_force_at_rest():
    """At rest, without wind and with pa = P lon on T points, advection and
    friction vanish and the column tendency is the atmospheric pressure
    force: ut = -a0 sqrt(chc) pax sum(dpr msk)/prh, with pax the mean of the
    exact east-west gradient P / (a cos(lat)) over the two T-point rows
    beside the U row, and vt = 0.  A uniform chc = c makes sqrt(ter(chc)) =
    sqrt(c)."""
    cfg = _channel(24, 16, nk=3)
    z12 = np.zeros((cfg.glo_ni, cfg.glo_nj, 12))
    cfg.taux = np.asfortranarray(z12)
    cfg.tauy = np.asfortranarray(z12.copy())
    P = 250.0
    o0 = Oracle(cfg)
    lam1 = _unwrapped_lon(o0, 1)
    cfg.pa = np.asfortranarray(np.broadcast_to((P * lam1[1:-1])[:, None, None], z12.shape).copy())
    o = Oracle(cfg)
    o.set_clock(0, (1, 1, 1, 0, 0, 0), (1, 1, 1, 0, 0, 0))
    o.run_stage("int_bnd")
    ni, nj, nk = o.ni, o.nj, o.nk
    for n in ("uc", "vc", "up", "vp"):
        o.set(n, np.zeros((ni, nj, nk)))
    c0 = 1.021
    o.set("chc", np.full((ni, nj), c0))
    o.set("chp", np.full((ni, nj), c0))
    o.set_step_index(3)
    o.run_stage("int_readyc")
    ut, vt = o.get("but"), o.get("bvt")
    phi1 = o.metric(1, 5)[2, :]
    phi3 = o.metric(3, 5)[2, :]
    assert np.allclose(phi3[1:-1], 0.5 * (phi1[1:-1] + phi1[2:]))   # U row j between T rows j, j+1
    pax = 0.5 * (P / (A_EARTH * np.cos(phi1[1:-1])) + P / (A_EARTH * np.cos(phi1[2:])))
    lev_u = o.igrid(3)
    dpr = o.vgrid(2, 3)
    prh = o.get("prh_u")
    col = np.array([[dpr[:lev_u[i, j]].sum() for j in range(nj)] for i in range(ni)]) / prh
    want = -C["a0"] * np.sqrt(c0) * pax[None, :] * col[:, 1:-1]
    cols = slice(2, ni - 2)         # away from the periodic seam of pa
    wet = lev_u[cols, 1:-1] > 0
    assert wet.sum() > 100
    np.testing.assert_allclose(ut[cols, 1:-1][wet], want[cols][wet], rtol=1e-12)
    assert np.abs(vt[cols, 1:-1][wet]).max() < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("rst,step,expect", [
    (0, 1, ("N", "N", "N")),      # bootstrap: the whole history is the new advection
    (0, 2, ("N", "A", "A")),      # second step: the oldest level repeats the first
    (0, 3, ("N", "A", "B")),      # steady state: shift by one
    (1, 1, ("N", "A", "B")),      # restart run: no bootstrap
])
def test_int_readyc_adams_bashforth_history(rst, step, expect):
    """int_readyc shifts the advection history (auc -> aup -> aupp) before
    computing the new auc, then on steps 1 and 2 of a run that is not a
    restart fills the missing levels with the newest ones."""
    cfg = configs.small()
    o = Oracle(cfg, restart_run=rst)
    rng = np.random.default_rng(17)
    ni, nj, nk = o.ni, o.nj, o.nk
    mu = o.igrid(1).astype(float)
    o.set("uc", _periodic(rng.standard_normal((ni, nj, nk))) * mu * 0.1)
    o.set("vc", _periodic(rng.standard_normal((ni, nj, nk))) * mu * 0.1)
    A, B, Cc = (rng.standard_normal((ni, nj, nk)) for _ in range(3))
    for comp in ("u", "v"):
        o.set(f"a{comp}c", A)
        o.set(f"a{comp}p", B)
        o.set(f"a{comp}pp", Cc)
    o.set_step_index(step)
    o.run_stage("int_readyc")
    for comp in ("u", "v"):
        new = o.get(f"a{comp}c")
        assert np.abs(new - A).max() > 1e-12
        vals = {"N": new, "A": A, "B": B}
        for name, key in zip((f"a{comp}c", f"a{comp}p", f"a{comp}pp"), expect):
            np.testing.assert_array_equal(o.get(name), vals[key], err_msg=f"{name} should be {key}")


# --------------------------------------------------------------------------
# int_ts accumulators (mod_int.f90:456-461, 500-504)

def _ts_state(rng, o):
    ni, nj, nk = o.ni, o.nj, o.nk
    mu = o.igrid(1).astype(float)
    o.set("uc", _periodic(rng.standard_normal((ni, nj, nk))) * mu * 0.05)
    o.set("vc", _periodic(rng.standard_normal((ni, nj, nk))) * mu * 0.05)
    t = np.broadcast_to(np.linspace(20.0, 2.0, nk)[None, None, :], (ni, nj, nk)).copy()
    s = np.full((ni, nj, nk), 35.0)
    for n, v in (("tc", t), ("tp", t), ("sc", s), ("sp", s)):
        o.set(n, v)
    o.set("bnd_t", np.full((ni, nj), 25.0))
    chc = 1.0 + 0.03 * _periodic(rng.uniform(size=(ni, nj)))
    o.set("chc", chc)
    o.set("chp", chc)
    o.set("mch", chc * 1.01)
    o.set("mch2", chc * 0.98)
    return t, s, chc


def test_int_ts_density_accumulator_is_new_density():
    """acr gains rho = 1/den_alpha of the NEW T, S (after the step and
    convection) at the layer-centre pressure chc pr(k) (not mch): equal, bit
    for bit, to 1/alpha of the pinned den_alpha applied to the result; land
    is left as it was."""
    cfg = configs.small()
    o = Oracle(cfg)
    rng = np.random.default_rng(29)
    t0, s0, chc = _ts_state(rng, o)
    acr0 = np.full((o.ni, o.nj, o.nk), 3.5)
    o.set("acr", acr0)
    o.run_stage("int_ts")
    inc = o.get("acr") - acr0
    tn, sn = o.get("tc"), o.get("sc")
    wet = o.igrid(0) > 0
    assert np.abs(tn - t0)[wet].max() > 1e-3          # the step changed T
    o.set("tc", tn)
    o.set("sc", sn)
    o.set("alpha", np.ones_like(tn))
    o.run_stage("den_alpha_main")
    rho = 1.0 / o.get("alpha")
    np.testing.assert_allclose(inc[wet], rho[wet], rtol=1e-14)
    assert np.all(inc[~wet] == 0.0)


def test_int_ts_vertical_velocity_accumulator_in_height_units():
    """acw gains w = wm alpha / (g mch2): the pressure velocity of each
    layer top converted to height units with the layer's specific volume
    (dz = alpha dp / g) and the step-mean bottom-pressure ratio mch2; with
    alpha = 1/rho0 and mch2 = 1 this is wm / (rho0 g)."""
    cfg = configs.small()
    o = Oracle(cfg)
    rng = np.random.default_rng(31)
    _, _, chc = _ts_state(rng, o)
    ni, nj, nk = o.ni, o.nj, o.nk
    alpha = (1.0 / C["rho0"]) * (1.0 + 0.01 * np.arange(nk))[None, None, :] * np.ones((ni, nj, nk))
    o.set("alpha", alpha)
    o.set("acw", np.zeros((ni, nj, nk)))
    o.run_stage("int_ts")
    wm = o.get("wm")
    acw = o.get("acw")
    assert np.abs(wm).max() > 1e-6
    mch2 = chc * 0.98
    want = wm[:, :, :nk] * alpha / (C["g"] * mch2[:, :, None])
    np.testing.assert_allclose(acw, want, rtol=1e-14, atol=1e-30)
    k0 = wm[:, :, :nk] / (C["rho0"] * C["g"])
    np.testing.assert_allclose(acw[:, :, 0] * mch2, k0[:, :, 0], rtol=1e-14, atol=1e-30)


def unstable_box():
    """The box with its columns made statically unstable to seeded random
    depths: water warmed by up to 8 K towards the bottom in 60% of the
    columns, salinity noise of +-0.5 psu."""
    import dataclasses
    cfg = configs.box()
    rng = np.random.default_rng(7)
    shape = (cfg.glo_ni, cfg.glo_nj, cfg.nk)
    ramp = np.linspace(0.0, 1.0, cfg.nk)[None, None, :]
    warm = rng.uniform(0.0, 8.0, shape) * (rng.random(shape[:2] + (1,)) < 0.6) * ramp
    pt = np.asfortranarray(np.asarray(cfg.pt) + warm)
    sa = np.asfortranarray(np.asarray(cfg.sa) + rng.uniform(-0.5, 0.5, shape))
    return dataclasses.replace(cfg, pt=pt, sa=sa)


def worst_instability(x, lev):
    """Largest rhob(k) - rhoa(k+1) over the levels k < nlev of the inner
    columns of x (anything with get()), with convect's linearised density
    (mod_int.f90:509-542): rhob(k) = prho_t(k+1) T(k) + prho_s(k+1) S(k),
    rhoa(k+1) = prho_t(k+1) T(k+1) + prho_s(k+1) S(k+1); and the largest
    |rhoa| compared.  The adjustment mixes wherever the first is positive,
    and its search leaves no such level, so after it the first is <= 0 up
    to rounding."""
    t, s, p1, p2 = (x.get(n) for n in ("tc", "sc", "prho_t", "prho_s"))
    rb = p1[:, :, 1:] * t[:, :, :-1] + p2[:, :, 1:] * s[:, :, :-1]
    ra = p1[:, :, 1:] * t[:, :, 1:] + p2[:, :, 1:] * s[:, :, 1:]
    active = lev[:, :, None] > np.arange(1, t.shape[2])[None, None, :]
    active[[0, -1], :, :] = False
    active[:, [0, -1], :] = False
    return float(np.max(np.where(active, rb - ra, -np.inf))), float(np.max(np.abs(ra[active])))


def test_convect_leaves_unstable_columns_stable():
    """Steps from the statically unstable box: before the first step most
    columns are unstable somewhere; after it and after two more every
    column passes convect's test with the step's prho, and deep layers
    (4 or more levels) were mixed to equal values."""
    cfg = unstable_box()
    o = Oracle(cfg)
    lev = o.igrid(2)
    o.run_stage("den_prho")
    worst0, scale = worst_instability(o, lev)
    assert worst0 > 1e-3 * scale
    for n in (1, 2):
        o.step(n)
        worst, scale = worst_instability(o, lev)
        assert worst <= 1e-12 * scale, (n, worst, scale)
    t = o.get("tc")[1:-1, 1:-1, :]
    run4 = (t[:, :, :-3] == t[:, :, 1:-2]) & (t[:, :, 1:-2] == t[:, :, 2:-1]) & (t[:, :, 2:-1] == t[:, :, 3:])
    assert run4.any(axis=2).sum() > 100
