"""Pins the FP64 oracle's composite operators (op_adv, op_adv_ts, op_dif,
fri_r3d, upwelling, int_pgra, int_trop, int_clin, int_ssh, int_ts) against properties the
mathematics of each fixes, independently of how the oracle evaluates them:

* skew-symmetry of the momentum advection (it neither makes nor destroys
  kinetic energy);
* exact gradients of coordinate functions and summation by parts for the
  tracer advection;
* self-adjointness, conservation and dissipation of the tracer diffusion;
* rigid rotations of the sphere feel no lateral viscosity, and the column
  momentum budget is surface stress minus bottom drag;
* the continuity equation's vertical velocity for barotropic and one-level
  flows;
* the barotropic subcycle: the ocean at rest over topography stays at rest,
  a discrete geostrophic balance is steady, and a uniform flow performs the
  inertial oscillation of its Crank-Nicolson Coriolis step;
* closed forms of the pressure-gradient face integrals for a constant and a
  linear specific volume;
* the baroclinic step keeps the depth mean equal to the barotropic velocity,
  rotates by the Crank-Nicolson angle and integrates a quadratic tendency
  history exactly (Adams-Bashforth 3);
* the sea surface of the reference stratification at rest is flat (int_ssh
  with den_alpha and the vertical grid);
* surface restoring changes only the top layer, by exactly its flux
  (int_ts).

The GPU kernels match the oracle bit for bit (tests/test_gpu_parity.py),
so these pins carry over to them."""
import numpy as np
import pytest

from pcom_amd import configs
from oracle.oracle import Oracle, constants

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


def _small_global():
    return configs.global_1deg(glo_ni=36, glo_nj=18, nk=6)


# --------------------------------------------------------------------------
# op_adv (mod_op.f90:224-278)

@pytest.mark.parametrize("name", ["box", "global"])
def test_op_adv_is_skew_symmetric(name):
    """op_adv is div(var V) - var div(V) / 2 with centred face averages: for
    ANY var, flow (U, V, w) and surface pressure, sum over the U cells of
    volume * var * op_adv(var) telescopes to zero (kinetic energy is neither
    made nor destroyed by advection).  Volume = rh dx1 dx2 dpr of the U
    stagger; w vanishes at the surface and below the bottom.  A wrong face
    offset, a dropped or mis-scaled -var div(V)/2, or a flux paired with the
    wrong transport breaks the cancellation."""
    cfg = configs.box() if name == "box" else _small_global()
    o = Oracle(cfg)
    rng = np.random.default_rng(11)
    ni, nj, nk = o.ni, o.nj, o.nk
    mu = o.igrid(1).astype(float)
    mt = o.igrid(0).astype(float)
    lev_t = o.igrid(2)
    uc = _periodic(rng.standard_normal((ni, nj, nk))) * mu
    vc = _periodic(rng.standard_normal((ni, nj, nk))) * mu
    wm = _periodic(rng.standard_normal((ni, nj, nk + 1)))
    wm[:, :, 0] = 0.0
    wm[:, :, 1:nk] *= mt[:, :, 1:nk]
    wm[:, :, nk] = 0.0
    if name == "global":
        # w at the bottom face of every T column, and the flux-form pairing
        # of U points whose T corners differ in depth, vanish only for a
        # w that is zero wherever a neighbour is land at that face
        for k in range(1, nk):
            deep = lev_t > k
            nb = deep.copy()
            for di in (-1, 0, 1):
                for dj in (-1, 0, 1):
                    nb &= np.roll(np.roll(deep, di, 0), dj, 1)
            wm[:, :, k] *= nb
    ch = 1.0 + 0.1 * _periodic(rng.standard_normal((ni, nj)))
    o.set("uc", uc)
    o.set("vc", vc)
    o.set("wm", wm)
    o.set("chc", ch)
    sch = np.sqrt(o.op_ter(1, 3, ch, dft=1.0))
    sch[o.igrid(3) == 0] = 1.0
    for var in (uc, vc):
        ans = o.op_adv(var, sch)
        vol = (o.metric(3, 0) * o.metric(3, 2) * o.metric(3, 3))[:, :, None] * o.vgrid(2, 3)[None, None, :]
        t = (vol * var * ans)[1:-1, 1:-1, :]
        assert abs(t.sum()) < 1e-13 * np.abs(t).sum()
        assert np.abs(t).sum() > 0


def test_op_adv_of_constant_in_nondivergent_flow():
    """A flow with zero horizontal divergence on the sphere advects a
    constant to nothing: u/sch uniform (d u / d lon = 0), or v/sch =
    V0 / cos(lat) (d(v cos lat)/d lat = 0), with w = 0.  Pins the metric
    factor of the meridional transport."""
    o = Oracle(_channel(48, 32, nk=2))
    rng = np.random.default_rng(21)
    ni, nj, nk = o.ni, o.nj, o.nk
    mu = o.igrid(1).astype(float)
    ch = 1.0 + 0.1 * _periodic(rng.standard_normal((ni, nj)))
    o.set("chc", ch)
    o.set("wm", np.zeros((ni, nj, nk + 1)))
    sch = np.sqrt(o.op_ter(1, 3, ch, dft=1.0))
    const = np.full((ni, nj, nk), 2.5) * mu
    rows = slice(3, nj - 3)                   # away from the walls
    for u0, v0 in ((0.7, 0.0), (0.0, 0.4)):
        uc = np.repeat((u0 * sch)[:, :, None], nk, 2) * mu
        vc = np.repeat((v0 * sch / o.metric(3, 0))[:, :, None], nk, 2) * mu
        o.set("uc", uc)
        o.set("vc", vc)
        ans = o.op_adv(const, sch)
        flux = np.abs(o.op_adv(_periodic(rng.standard_normal((ni, nj, nk))) * mu, sch))[1:-1, rows]
        assert flux.max() > 0
        assert np.abs(ans[1:-1, rows]).max() < 1e-12 * flux.max(), (u0, v0)


# --------------------------------------------------------------------------
# op_adv_ts (mod_op.f90:280-325)

def _corner_mean(x):
    """Mean of a U-point field over the four U points around each T point
    (U(i-1:i, j-1:j) for T(i, j))."""
    out = np.full(x.shape, np.nan)
    out[1:, 1:] = 0.25 * (x[:-1, :-1] + x[1:, :-1] + x[:-1, 1:] + x[1:, 1:])
    return out


def test_op_adv_ts_of_coordinates():
    """The advective tracer operator of the coordinate functions is the
    velocity component along them: op_adv_ts(lon) = <u sch> / (a rh) and
    op_adv_ts(lat) = <v sch rh_u> / (a rh), <.> the mean over the four
    velocity points around the tracer point, with no vertical part.  Pins
    which transport pairs with which difference and the metric factor."""
    cfg = configs.global_1deg(glo_ni=24, glo_nj=16, nk=4)
    o = Oracle(cfg)
    rng = np.random.default_rng(12)
    ni, nj, nk = o.ni, o.nj, o.nk
    mu = o.igrid(1).astype(float)
    uc = _periodic(rng.standard_normal((ni, nj, nk))) * mu
    vc = _periodic(rng.standard_normal((ni, nj, nk))) * mu
    ch = 1.0 + 0.1 * _periodic(rng.standard_normal((ni, nj)))
    o.set("uc", uc)
    o.set("vc", vc)
    o.set("chc", ch)
    sch = np.sqrt(o.op_ter(1, 3, ch, dft=1.0))
    rh1 = o.metric(1, 0)
    wm = np.zeros((ni, nj, nk + 1))
    lon = np.repeat(o.metric(1, 4)[:, :, None], nk, 2)
    lat = np.repeat(o.metric(1, 5)[:, :, None], nk, 2)
    # longitude jumps by 2 pi across the periodic seam: use columns whose
    # differences do not cross it
    ans = o.op_adv_ts(lon, wm)
    want = np.stack([_corner_mean((uc * sch[:, :, None])[:, :, k]) for k in range(nk)], 2) / (A_EARTH * rh1[:, :, None])
    sl = (slice(2, ni - 2), slice(1, nj - 1))
    np.testing.assert_allclose(ans[sl], want[sl], rtol=0, atol=1e-14 * np.abs(want[sl]).max())
    ans = o.op_adv_ts(lat, wm)
    um2 = vc * (sch * o.metric(3, 0))[:, :, None]
    want = np.stack([_corner_mean(um2[:, :, k]) for k in range(nk)], 2) / (A_EARTH * rh1[:, :, None])
    sl = (slice(1, ni - 1), slice(2, nj - 2))
    np.testing.assert_allclose(ans[sl], want[sl], rtol=0, atol=1e-14 * np.abs(want[sl]).max())


def test_op_adv_ts_vertical_summation_by_parts():
    """For a tracer that varies only with depth and any w vanishing at the
    surface and bottom, the dpr-weighted column sum of the vertical
    advection w d(var)/dp equals -sum var dw (summation by parts):
    sum_k dpr_k A_k = sum_k var_k (w_{k+1} - w_k)."""
    o = Oracle(configs.global_1deg(glo_ni=24, glo_nj=16, nk=5))
    rng = np.random.default_rng(13)
    ni, nj, nk = o.ni, o.nj, o.nk
    o.set("uc", _periodic(rng.standard_normal((ni, nj, nk))))
    o.set("vc", _periodic(rng.standard_normal((ni, nj, nk))))
    phi = rng.standard_normal(nk)
    var = np.broadcast_to(phi[None, None, :], (ni, nj, nk)).copy()
    wm = _periodic(rng.standard_normal((ni, nj, nk + 1)))
    wm[:, :, 0] = 0.0
    wm[:, :, nk] = 0.0
    ans = o.op_adv_ts(var, wm)
    lhs = (ans * o.vgrid(2, 3)[None, None, :]).sum(2)
    rhs = (phi[None, None, :] * (wm[:, :, 1:] - wm[:, :, :-1])).sum(2)
    np.testing.assert_allclose(lhs, rhs, rtol=0, atol=1e-14 * np.abs(rhs).max())


# --------------------------------------------------------------------------
# op_dif (mod_op.f90:186-222)

@pytest.mark.parametrize("name", ["box", "global"])
def test_op_dif_self_adjoint_conservative_dissipative(name):
    """Tracer diffusion div(ch ah grad) + d/dp(g rho0 kh d/dz) with no flux
    through land, surface or bottom is, under the T-cell volume inner
    product <a, b> = sum rh dx1 dx2 dpr a b:
    self-adjoint, <x, D y> = <y, D x>; conservative, <1, D x> = 0;
    dissipative, <x, D x> < 0.  A gradient taken at the wrong face, a metric
    factor of the wrong stagger, a sign, or a flux leaking through a wall
    breaks one of the three."""
    cfg = configs.box() if name == "box" else _small_global()
    o = Oracle(cfg)
    rng = np.random.default_rng(14)
    ni, nj, nk = o.ni, o.nj, o.nk
    o.set("kh", 1e-4 * np.abs(_periodic(rng.standard_normal((ni, nj, nk + 1)))))
    ch = 1.0 + 0.1 * _periodic(rng.standard_normal((ni, nj)))
    x = _periodic(rng.standard_normal((ni, nj, nk)))
    y = _periodic(rng.standard_normal((ni, nj, nk)))
    dx, dy = o.op_dif(ch, x), o.op_dif(ch, y)
    vol = (o.metric(1, 0) * o.metric(1, 2) * o.metric(1, 3))[:, :, None] * o.vgrid(2, 3)[None, None, :]

    def ip(a, b):
        t = (vol * a * b)[1:-1, 1:-1, :]
        return t.sum(), np.abs(t).sum()

    s1, c1 = ip(x, dy)
    s2, c2 = ip(y, dx)
    assert abs(s1 - s2) < 1e-13 * max(c1, c2)
    s, c = ip(np.ones_like(x), dx)
    assert abs(s) < 1e-13 * c
    s, c = ip(x, dx)
    assert s < -0.5 * c


def test_op_dif_vertical_part_dissipates():
    """A tracer uniform in the horizontal feels only the vertical diffusion
    d/dp(g rho0 kh dT/dz): self-adjoint and dissipative on its own."""
    o = Oracle(_small_global())
    rng = np.random.default_rng(22)
    ni, nj, nk = o.ni, o.nj, o.nk
    o.set("kh", 1e-4 * np.abs(_periodic(rng.standard_normal((ni, nj, nk + 1)))))
    ch = 1.0 + 0.1 * _periodic(rng.standard_normal((ni, nj)))
    x = np.broadcast_to(rng.standard_normal(nk)[None, None, :], (ni, nj, nk)).copy()
    y = np.broadcast_to(rng.standard_normal(nk)[None, None, :], (ni, nj, nk)).copy()
    dx, dy = o.op_dif(ch, x), o.op_dif(ch, y)
    dpr = o.vgrid(2, 3)[None, None, :]
    I = (slice(1, ni - 1), slice(1, nj - 1))
    s_xy, s_yx, s_xx = ((dpr * a * b)[I].sum(2) for a, b in ((x, dy), (y, dx), (x, dx)))
    sc = (dpr * np.abs(x * dy))[I].sum(2) + (dpr * np.abs(y * dx))[I].sum(2) + 1e-300
    assert np.all(np.abs(s_xy - s_yx) < 1e-13 * sc)
    deep = o.igrid(2)[I] > 1
    assert np.all(s_xx[deep] < 0)


@pytest.mark.parametrize("field", ["lon", "lat"])
def test_op_dif_closed_form_on_the_sphere(field):
    """Lateral diffusion div(ch grad T) of T = cos(lon) and T = sin(lat) with
    ch = 1 + 0.3 sin(2 lat), no vertical diffusion, against the exact
    spherical operators (second-order truncation, ~1e-3 here):
    cos(lon): -ah ch cos(lon) / (a cos lat)^2;
    sin(lat): ah (d/dlat(ch cos^2 lat)) / (a^2 cos lat).
    A metric factor of the wrong stagger or ch taken at the wrong face
    is off by O(grid spacing), ~1e-2."""
    o = Oracle(_channel(144, 96, nk=2))
    ni, nj, nk = o.ni, o.nj, o.nk
    o.set("kh", np.zeros((ni, nj, nk + 1)))
    lon, lat = o.metric(1, 4), o.metric(1, 5)
    ch = 1.0 + 0.3 * np.sin(2 * lat)
    ah = C["ah_c"]
    a = A_EARTH
    if field == "lon":
        t = np.cos(lon)
        want = -ah * ch * np.cos(lon) / (a * np.cos(lat)) ** 2
    else:
        t = np.sin(lat)
        d = 0.6 * np.cos(2 * lat) * np.cos(lat) ** 2 - 2.0 * ch * np.cos(lat) * np.sin(lat)
        want = ah * d / (a * a * np.cos(lat))
    got = o.op_dif(ch, t[:, :, None])[:, :, 0]
    deg = np.rad2deg(lat[1, :])
    band = (np.abs(deg) > 10) & (np.abs(deg) < 45)
    err = np.abs(got - want)[1:-1, band] / np.abs(want)[1:-1, band].max()
    assert err.max() < 2e-3, err.max()


# --------------------------------------------------------------------------
# fri_r3d / op_fri (mod_op.f90:103-184)

def _fri_setup(o, am):
    ni, nj, nk = o.ni, o.nj, o.nk
    o.set("am", np.full((ni, nj, nk), am))
    o.set("km", np.zeros((ni, nj, nk + 1)))
    o.set("bnd_taux", np.zeros((ni, nj)))
    o.set("bnd_tauy", np.zeros((ni, nj)))
    o.set("chc", np.ones((ni, nj)))


def test_fri_rigid_rotations_feel_no_viscosity():
    """A solid-body rotation of the sphere has no strain, so the lateral
    viscous force of the Laplacian with its metric terms,
    F_u = am (lap u + (1 - tan^2)/a^2 u - 2 sin/(a^2 cos^2) du/dlon), and
    likewise F_v, vanishes up to the second-order truncation error.  The
    rotation about the polar axis (u = cos lat) tests the cv1 term; the
    rotations about equatorial axes,
    (u, v) = (-sin lat cos lon, sin lon) and (sin lat sin lon, cos lon),
    test the cv2 cross terms.

    The reference forms tan(lat) as sqrt(1 - cos^2)/cos (main.f90:667-671),
    which is |tan|, so south of the equator the cross term has the wrong
    sign: there the residual is exactly the one that sign predicts,
    am * 4 |tan|/a * d(vc)/dx."""
    o = Oracle(_channel(72, 48))
    ni, nj, nk = o.ni, o.nj, o.nk
    am = 1e4
    _fri_setup(o, am)
    lon, lat = o.metric(3, 4), o.metric(3, 5)
    mu = o.igrid(1).astype(float)
    scale = am / A_EARTH ** 2
    deg = np.rad2deg(lat[1, :])
    north = (deg > 10) & (deg < 45)
    south = (deg < -10) & (deg > -45)
    sch = np.ones((ni, nj))
    levels = slice(0, nk - 1)            # the bottom level also feels drag
    rotations = {
        "polar": (np.cos(lat), 0 * lat, 0 * lat, 0 * lat),
        "x": (-np.sin(lat) * np.cos(lon), np.sin(lon), np.cos(lon), np.sin(lat) * np.sin(lon)),
        "y": (np.sin(lat) * np.sin(lon), np.cos(lon), -np.sin(lon), np.sin(lat) * np.cos(lon)),
    }
    for name, (u, v, dvdlon, dudlon) in rotations.items():
        o.set("uc", np.repeat(u[:, :, None], nk, 2) * mu)
        o.set("vc", np.repeat(v[:, :, None], nk, 2) * mu)
        fx, fy = o.op_fri(sch)
        for f in (fx, fy):
            r = f[1:-1, north, levels]
            assert np.abs(r).max() < 5e-3 * scale, (name, np.abs(r).max() / scale)
        # fx's cross term differentiates vc = -v, fy's differentiates u
        tq = 4.0 * np.abs(np.tan(lat)) / A_EARTH / (A_EARTH * np.cos(lat))
        pred_x = am * tq * (-dvdlon)
        pred_y = am * tq * dudlon
        for f, pred in ((fx, pred_x), (fy, pred_y)):
            r = f[1:-1, south, levels] - pred[1:-1, south, None]
            assert np.abs(r).max() < 1e-2 * scale, (name, np.abs(r).max() / scale)
        if name == "polar":
            # no cross term: the southern rows vanish as well
            assert np.abs(fx[1:-1, south, levels]).max() < 5e-3 * scale


def test_fri_rigid_rotation_error_is_second_order():
    """The polar-axis rotation's residual falls by ~4x per halving of the
    grid spacing (second-order operators and metric terms)."""
    res = []
    for ni, nj in ((36, 24), (72, 48)):
        o = Oracle(_channel(ni, nj))
        _fri_setup(o, 1e4)
        lat = o.metric(3, 5)
        mu = o.igrid(1).astype(float)
        o.set("uc", np.repeat(np.cos(lat)[:, :, None], o.nk, 2) * mu)
        o.set("vc", np.zeros((o.ni, o.nj, o.nk)))
        fx, _ = o.op_fri(np.ones((o.ni, o.nj)))
        deg = np.rad2deg(lat[1, :])
        band = (np.abs(deg) > 10) & (np.abs(deg) < 45)
        res.append(np.abs(fx[1:-1, band, 0]).max())
    assert 3.5 < res[0] / res[1] < 4.5, res


def test_fri_column_budget_is_wind_minus_drag():
    """With no lateral viscosity the friction is the divergence of the
    vertical stress, so its dpr-weighted column sum is the surface stress
    minus the quadratic bottom drag, whatever km and the velocity profile:
    sum_k dpr_k fx_k = g (taux - cdbot |V_b| u_b) / sch,
    u_b, V_b = the unweighted velocity at the bottom level."""
    o = Oracle(_small_global())
    rng = np.random.default_rng(15)
    ni, nj, nk = o.ni, o.nj, o.nk
    g = C["g"]
    mu = o.igrid(1).astype(float)
    lev_u = o.igrid(3)
    _fri_setup(o, 0.0)
    o.set("km", 1e-2 * np.abs(_periodic(rng.standard_normal((ni, nj, nk + 1)))))
    taux = 0.1 * _periodic(rng.standard_normal((ni, nj)))
    tauy = 0.1 * _periodic(rng.standard_normal((ni, nj)))
    o.set("bnd_taux", taux)
    o.set("bnd_tauy", tauy)
    uc = _periodic(rng.standard_normal((ni, nj, nk))) * mu
    vc = _periodic(rng.standard_normal((ni, nj, nk))) * mu
    o.set("uc", uc)
    o.set("vc", vc)
    sch = 1.0 + 0.05 * _periodic(rng.uniform(size=(ni, nj)))
    fx, fy = o.op_fri(sch)
    dpr = o.vgrid(2, 3)
    cdbot = C["cdbot"]
    for i in range(1, ni - 1):
        for j in range(1, nj - 1):
            L = lev_u[i, j]
            if L == 0:
                continue
            ub, vb = uc[i, j, L - 1] / sch[i, j], vc[i, j, L - 1] / sch[i, j]
            mag = np.hypot(ub, vb)
            for f, tau, wb in ((fx, taux, ub), (fy, tauy, vb)):
                got = (f[i, j, :L] * dpr[:L]).sum()
                want = g * (tau[i, j] - cdbot * mag * wb) / sch[i, j]
                assert abs(got - want) < 1e-12 * (abs(g * tau[i, j]) + abs(g * cdbot * mag * wb)), (i, j)


def test_fri_vertical_viscosity_self_adjoint_and_dissipative():
    """Without lateral viscosity, wind or a bottom velocity, the friction is
    d/dp(g rho0 km d(u)/dz) / sch^2 in the weighted velocity: self-adjoint
    and dissipative under the column inner product sum_k dpr_k a_k b_k."""
    o = Oracle(_small_global())
    rng = np.random.default_rng(16)
    ni, nj, nk = o.ni, o.nj, o.nk
    mu = o.igrid(1).astype(float)
    lev_u = o.igrid(3)
    _fri_setup(o, 0.0)
    o.set("km", 1e-2 * np.abs(_periodic(rng.standard_normal((ni, nj, nk + 1)))))
    sch = 1.0 + 0.05 * _periodic(rng.uniform(size=(ni, nj)))
    fields = []
    for _ in range(2):
        u = _periodic(rng.standard_normal((ni, nj, nk))) * mu
        for k in range(nk):               # no bottom velocity: no drag
            u[:, :, k] *= lev_u != k + 1
        o.set("uc", u)
        o.set("vc", np.zeros((ni, nj, nk)))
        fx, _ = o.op_fri(sch)
        fields.append((u, fx))
    (a, fa), (b, fb) = fields
    dpr = o.vgrid(2, 3)[None, None, :]
    s_ab = (dpr * a * fb)[1:-1, 1:-1].sum(2)
    s_ba = (dpr * b * fa)[1:-1, 1:-1].sum(2)
    s_aa = (dpr * a * fa)[1:-1, 1:-1].sum(2)
    sc = (dpr * np.abs(a * fb))[1:-1, 1:-1].sum(2) + (dpr * np.abs(b * fa))[1:-1, 1:-1].sum(2) + 1e-300
    assert np.all(np.abs(s_ab - s_ba) < 1e-13 * sc)
    assert np.all(s_aa <= 1e-13 * sc)
    assert (s_aa < 0).sum() > 10


def test_fri_linear_shear_with_matching_wind_is_steady():
    """Without lateral viscosity, a velocity growing linearly with depth,
    u = s z, under a uniform km carries the stress rho0 km du/dz_up =
    -rho0 km s at every interface; with that same stress as the wind,
    every layer above the bottom one is in equilibrium (the bottom one also
    feels the drag).  Pins the conversion of the vertical stress between
    the weighted velocity (u sch) and the wind (1/sch^2 against 1/sch)."""
    o = Oracle(configs.box())
    rng = np.random.default_rng(23)
    ni, nj, nk = o.ni, o.nj, o.nk
    mu = o.igrid(1).astype(float)
    _fri_setup(o, 0.0)
    km, s = 2e-3, 1e-4
    o.set("km", np.full((ni, nj, nk + 1), km))
    tau = np.full((ni, nj), -C["rho0"] * km * s)
    o.set("bnd_taux", tau)
    o.set("bnd_tauy", -tau)
    sch = 1.0 + 0.2 * _periodic(rng.uniform(size=(ni, nj)))
    z = o.vgrid(2, 0)
    prof = s * z[None, None, :] * sch[:, :, None] * mu
    o.set("uc", prof)
    o.set("vc", -prof)
    fx, fy = o.op_fri(sch)
    wet = o.igrid(3) > 0
    g = C["g"]
    scale = g * abs(tau[0, 0]) / o.vgrid(2, 3).min()
    for f in (fx, fy):
        assert np.abs(f[1:-1, 1:-1, :nk - 1][wet[1:-1, 1:-1]]).max() < 1e-12 * scale
        assert np.abs(f[1:-1, 1:-1, nk - 1][wet[1:-1, 1:-1]]).max() > 1e-3 * scale


# --------------------------------------------------------------------------
# upwelling (mod_int.f90:626-647)

def test_upwelling_of_barotropic_flow_is_zero():
    """A depth-independent horizontal flow has no baroclinic divergence:
    w(k) = w(1) at every wet level (flat-bottomed basin, where every column
    sees the same flow at every level)."""
    o = Oracle(configs.box())
    rng = np.random.default_rng(17)
    ni, nj, nk = o.ni, o.nj, o.nk
    mu = o.igrid(1).astype(float)
    mt = o.igrid(0).astype(float)
    lev = o.igrid(2)
    u2 = _periodic(rng.standard_normal((ni, nj)))
    v2 = _periodic(rng.standard_normal((ni, nj)))
    o.set("uc", u2[:, :, None] * mu)
    o.set("vc", v2[:, :, None] * mu)
    o.set("chp", 1.0 + 0.1 * _periodic(rng.standard_normal((ni, nj))))
    wm = np.zeros((ni, nj, nk + 1))
    wm[:, :, 0] = 0.3 * rng.standard_normal((ni, nj)) * (lev > 0)
    o.set("wm", wm)
    w = o.upwelling()
    assert np.abs(w[:, :, 0] - wm[:, :, 0]).max() == 0.0        # w(1) is not written
    d = np.abs(w[:, :, 1:nk] - w[:, :, :1]) * mt[:, :, 1:nk]
    # a flow that differs between levels moves w by O(1) here
    assert d.max() < 1e-11
    assert np.all(w[:, :, 1:nk][mt[:, :, 1:nk] == 0] == 0.0)


def test_upwelling_of_one_level_flow():
    """Flow at a single level k0 with divergence D there: continuity gives a
    w that changes by c dpr_k across every other wet level (c = -D dpr_k0 /
    P, the barotropic divergence spread over the column) and by (D + c)
    dpr_k0 across level k0, so that w returns to w(1) at the bottom:
    w(lev) + c dpr_lev = w(1).  Checked from the output alone: the
    increments of every other level share one c, and the k0 increment
    closes the column."""
    o = Oracle(_small_global())
    rng = np.random.default_rng(18)
    ni, nj, nk = o.ni, o.nj, o.nk
    mu = o.igrid(1).astype(float)
    lev = o.igrid(2)
    dpr = o.vgrid(2, 3)
    o.set("chp", 1.0 + 0.1 * _periodic(rng.standard_normal((ni, nj))))
    checked = 0
    for k0 in range(nk - 1):
        uc = np.zeros((ni, nj, nk))
        vc = np.zeros((ni, nj, nk))
        uc[:, :, k0] = _periodic(rng.standard_normal((ni, nj))) * mu[:, :, k0]
        vc[:, :, k0] = _periodic(rng.standard_normal((ni, nj))) * mu[:, :, k0]
        o.set("uc", uc)
        o.set("vc", vc)
        wm = np.zeros((ni, nj, nk + 1))
        wm[:, :, 0] = 0.3 * rng.standard_normal((ni, nj)) * (lev > 0)
        o.set("wm", wm)
        w = o.upwelling()
        for i in range(1, ni - 1):
            for j in range(1, nj - 1):
                L = lev[i, j]
                if L < 2 or k0 >= L - 1:
                    continue
                inc = np.diff(w[i, j, :L])                  # across levels 1 .. L-1
                c = (w[i, j, 0] - w[i, j, L - 1]) / dpr[L - 1]  # from the bottom closure
                tol = 1e-9 * (abs(c) + np.abs(inc / dpr[:L - 1]).max()) + 1e-300
                for k in range(L - 1):
                    if k != k0:
                        assert abs(inc[k] / dpr[k] - c) < tol, (k0, i, j, k)
                assert abs(inc[k0] + c * (dpr[:L].sum() - dpr[k0])) / dpr[k0] < tol, (k0, i, j)
                checked += 1
    assert checked > 1000


# --------------------------------------------------------------------------
# int_pgra (mod_int.f90:64-88)

def _pgra(o, alpha):
    o.set("alpha", alpha)
    o.run_stage("int_pgra")
    return {n: o.get(n) for n in ("bphi_xn", "bphi_xs", "bphi_ye", "bphi_yw",
                                  "bgra_xn", "bgra_xs", "bgra_ye", "bgra_yw")}


def test_int_pgra_constant_specific_volume():
    """For a constant specific volume alpha in a flat-bottomed basin,
    adp + p alpha = alpha P_bottom at every level (adp = alpha (P - p) at
    mid-layer pressures p), so each face integral is alpha P_bottom where its
    mask is wet and 0 where it is land, the masks being the face's own row
    (xn) or the row south (xs), its own column (ye) or the column west (yw);
    the face gradients vanish."""
    cfg = configs.box()
    o = Oracle(cfg)
    ni, nj, nk = o.ni, o.nj, o.nk
    pr1 = o.vgrid(1, 1)
    np.testing.assert_array_equal(o.vgrid(2, 1), 0.5 * (pr1[:-1] + pr1[1:]))
    P = pr1[-1]
    al = 9.7e-4
    b = _pgra(o, np.full((ni, nj, nk), al))
    lev_t, lev_u = o.igrid(2), o.igrid(3)
    wet_u = lev_u > 0
    south = np.roll(wet_u, 1, axis=1)        # row j-1; row 1 takes row nj
    west = np.roll(wet_u, 1, axis=0)
    wet_t = lev_t > 0
    # faces between two wet T columns: x faces pair T(i, j) and T(i+1, j),
    # y faces pair T(i, j) and T(i, j+1)
    fx_ok = wet_t & np.roll(wet_t, -1, axis=0)
    fy_ok = wet_t & np.roll(wet_t, -1, axis=1)
    I = (slice(1, ni - 1), slice(1, nj - 1))
    for name, ok, mask in (("bphi_xn", fx_ok, wet_u), ("bphi_xs", fx_ok, south),
                           ("bphi_ye", fy_ok, wet_u), ("bphi_yw", fy_ok, west)):
        got, want = b[name][I][ok[I]], (al * P * mask)[I][ok[I]]
        np.testing.assert_allclose(got, want, rtol=1e-15, atol=0, err_msg=name)
        assert (want == 0).any() and (want != 0).any(), name
    for name in ("bgra_xn", "bgra_xs", "bgra_ye", "bgra_yw"):
        assert np.abs(b[name][I][(fx_ok & fy_ok)[I]]).max() == 0.0, name


@pytest.mark.parametrize("axis", ["lon", "lat"])
def test_int_pgra_linear_specific_volume(axis):
    """alpha = alpha0 + beta x with x the longitude (or latitude) in
    radians: the depth-mean of grad(adp) over the face is
    beta/(a rh) * sum dpr (P - p) / P = beta P / (2 a rh) along x
    (beta P / (2 a) along y; the midpoint sum of a linear integrand is
    exact) and 0 across."""
    cfg = configs.box()
    o = Oracle(cfg)
    ni, nj, nk = o.ni, o.nj, o.nk
    P = o.vgrid(1, 1)[-1]
    beta = 1e-5
    x = o.metric(1, 4) if axis == "lon" else o.metric(1, 5)
    b = _pgra(o, np.repeat((9.7e-4 + beta * x)[:, :, None], nk, 2))
    wet_t = o.igrid(2) > 0
    lev_u = o.igrid(3)
    # interior faces of the basin whose whole neighbourhood is wet
    ok = wet_t.copy()
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            ok &= np.roll(np.roll(wet_t, di, 0), dj, 1)
    ok &= lev_u > 0
    I = (slice(2, ni - 2), slice(2, nj - 2))
    if axis == "lon":
        want = beta * P / (2 * A_EARTH * o.metric(2, 0))
        np.testing.assert_allclose(b["bgra_xn"][I][ok[I]], want[I][ok[I]], rtol=1e-11)
        assert np.abs(b["bgra_ye"][I][ok[I]]).max() < 1e-12 * np.abs(want).max()
    else:
        want = beta * P / (2 * A_EARTH)
        np.testing.assert_allclose(b["bgra_ye"][I][ok[I]], want, rtol=1e-11)
        assert np.abs(b["bgra_xn"][I][ok[I]]).max() < 1e-12 * want


# --------------------------------------------------------------------------
# int_trop (mod_int.f90:153-296)

def _trop_quiet(o):
    """No lateral viscosity, no baroclinic forcing of the barotropic mode."""
    ni, nj, nk = o.ni, o.nj, o.nk
    for n in ("but", "bvt"):
        o.set(n, np.zeros((ni, nj)))
    o.set("am", np.zeros((ni, nj, nk)))


def test_int_trop_ocean_at_rest_stays_at_rest():
    """The reference stratification (horizontal-mean T/S everywhere) at
    rest with a flat surface (ch = 1) over the seeded bathymetry: the
    bottom-geopotential gradient graphih and the face integrals of the
    baroclinic pressure (ch * bgra) cancel exactly, so the subcycle leaves
    the ocean at rest and ch unchanged.  Pins the signs and the face
    pairing (bgra_xn at row j with bgra_xs at row j+1) against the
    topography term."""
    cfg = _small_global()
    o = Oracle(cfg)
    ni, nj, nk = o.ni, o.nj, o.nk
    t = np.broadcast_to(cfg.tmn[None, None, :], (ni, nj, nk)).copy()
    sal = np.broadcast_to(cfg.smn[None, None, :], (ni, nj, nk)).copy()
    for n, v in (("tc", t), ("sc", sal), ("tp", t), ("sp", sal)):
        o.set(n, v)
    for n in ("chc", "chp"):
        o.set(n, np.ones((ni, nj)))
    for n in ("buc", "bvc", "bup", "bvp", "but", "bvt"):
        o.set(n, np.zeros((ni, nj)))
    o.run_stage("den_alpha_main")
    o.run_stage("int_pgra")
    wet = o.igrid(3) > 0
    force = np.abs(o.get("graphihx"))[wet].max()
    assert force > 1e-3                       # the topography term is not small
    o.run_stage("int_trop")
    # an imbalance of 1e-3 of the topography term would reach ~force * 1e-3 * dtbc
    assert np.abs(o.get("buc"))[wet].max() < 1e-12
    assert np.abs(o.get("bvc"))[wet].max() < 1e-12
    assert np.abs(o.get("chc") - 1.0)[o.igrid(2) > 0].max() < 1e-14


def test_int_trop_geostrophic_balance_is_steady():
    """On an f-plane (uniform cor) in a flat channel with constant alpha,
    a zonal flow in discrete geostrophic balance with a meridional surface
    slope, f u = -alpha P sch d(ch)/dy / a, is steady: it is non-divergent,
    so ch keeps its value, and the predictor-corrector Coriolis step leaves
    it unchanged."""
    o = Oracle(_channel(48, 32, nk=3))
    ni, nj, nk = o.ni, o.nj, o.nk
    al = 9.7e-4
    o.set("alpha", np.full((ni, nj, nk), al))
    o.run_stage("int_pgra")
    for n in ("bgra_xn", "bgra_xs", "bgra_ye", "bgra_yw", "graphihx", "graphihy", "bvc", "bvp"):
        o.set(n, np.zeros((ni, nj)))
    _trop_quiet(o)
    f0 = 1e-4
    o.set("cor", np.full((ni, nj), f0))
    ch = 1.0 + 0.01 * np.sin(3 * o.metric(1, 5))
    o.set("chc", ch)
    o.set("chp", ch)
    P = o.vgrid(1, 1)[-1]
    gy = o.op_px2(1, 4, ch) / A_EARTH
    sch = np.sqrt(o.op_ter(1, 3, ch))
    wet = o.igrid(3) > 0
    u = np.where(wet, -al * P * gy * sch / f0, 0.0)
    assert np.abs(u).max() > 0.1
    o.set("buc", u)
    o.set("bup", u)
    o.run_stage("int_trop")
    assert np.abs(o.get("buc") - u)[wet].max() < 1e-13
    assert np.abs(o.get("bvc"))[wet].max() < 1e-13
    assert np.abs(o.get("chc") - ch)[o.igrid(2) > 0].max() < 1e-13


def test_int_trop_inertial_oscillation():
    """A uniform flow with no pressure gradient turns under Coriolis only:
    each substep's predictor (Crank-Nicolson) and corrector (f times the
    half-step velocity) together rotate u + i v by exp(-i theta), with
    cos(theta) = (1 - x^2)/(1 + x^2), x = f dtbt / 2, and the step returns
    the mean of the initial and the nstep substep velocities."""
    o = Oracle(_channel(48, 32, nk=2))
    ni, nj = o.ni, o.nj
    for n in ("bphi_xn", "bphi_xs", "bphi_ye", "bphi_yw", "bgra_xn", "bgra_xs", "bgra_ye", "bgra_yw",
              "graphihx", "graphihy"):
        o.set(n, np.zeros((ni, nj)))
    _trop_quiet(o)
    wet = o.igrid(3) > 0
    u0, v0 = 0.3, -0.2
    for n, v in (("buc", u0), ("bup", u0), ("bvc", v0), ("bvp", v0)):
        o.set(n, np.where(wet, v, 0.0))
    o.run_stage("int_trop")
    f = o.get("cor")
    bt = o.cfg.dtbt
    nstep = 2 * o.cfg.dtbc // bt
    x = 0.5 * f * bt
    theta = np.arctan2(2 * x, 1 - x * x)
    n = np.arange(nstep + 1)
    want = (u0 + 1j * v0) * np.exp(-1j * theta[:, :, None] * n[None, None, :]).mean(2)
    got = o.get("buc") + 1j * o.get("bvc")
    assert np.abs(got - want)[wet].max() < 1e-14
    # the flow has turned measurably over the step
    assert np.abs(want - (u0 + 1j * v0))[wet].max() > 1e-3


# --------------------------------------------------------------------------
# int_clin (mod_int.f90:298-401)

def test_int_clin_depth_mean_is_barotropic_velocity():
    """After every step the baroclinic velocity's dpr-weighted depth mean
    over wet levels equals the barotropic velocity (mod_int.f90:371-383
    replaces the mean by it)."""
    o = Oracle(configs.global_1deg(glo_ni=36, glo_nj=18, nk=6))
    o.step(3)
    uc, vc, buc, bvc = o.get("uc"), o.get("vc"), o.get("buc"), o.get("bvc")
    mu = o.igrid(1).astype(float)
    dpr = o.vgrid(2, 3)[None, None, :]
    prh = o.get("prh_u")
    wet = o.igrid(3) > 0
    for v, bv in ((uc, buc), (vc, bvc)):
        mean = (v * dpr * mu).sum(2) / prh
        assert np.abs(mean - bv)[wet].max() < 1e-12 * (np.abs(bv).max() + 1e-3)
        assert np.abs(bv[wet]).max() > 0


def _clin_quiet(o):
    """A flat-bottomed basin with every forcing of int_clin switched off:
    no pressure gradient (alpha = adp = 0, uniform mch), no friction, no
    barotropic velocity."""
    ni, nj, nk = o.ni, o.nj, o.nk
    z3 = np.zeros((ni, nj, nk))
    z2 = np.zeros((ni, nj))
    for n in ("auc", "aup", "aupp", "avc", "avp", "avpp", "fx", "fy", "alpha", "adp"):
        o.set(n, z3)
    for n in ("pax", "pay", "buc", "bvc"):
        o.set(n, z2)
    o.set("mch", np.ones((ni, nj)))
    o.set("mch2", np.ones((ni, nj)))


def _baroclinic(o, rng):
    """A random velocity with zero depth mean at every U column."""
    mu = o.igrid(1).astype(float)
    dpr = o.vgrid(2, 3)[None, None, :]
    prh = o.get("prh_u")
    v = _periodic(rng.standard_normal((o.ni, o.nj, o.nk))) * mu
    mean = (v * dpr * mu).sum(2) / prh
    return (v - mean[:, :, None]) * mu


def test_int_clin_coriolis_is_crank_nicolson_rotation():
    """With only the Coriolis force acting, the semi-implicit step rotates
    (u, v) by the Crank-Nicolson angle: speed is kept exactly and
    cos(angle) = (1 - c^2)/(1 + c^2), c = f dt / 2."""
    o = Oracle(configs.box())
    rng = np.random.default_rng(19)
    _clin_quiet(o)
    u, v = _baroclinic(o, rng), _baroclinic(o, rng)
    for n, x in (("uc", u), ("up", u), ("vc", v), ("vp", v)):
        o.set(n, x)
    o.run_stage("int_clin")
    u2, v2 = o.get("uc"), o.get("vc")
    wet = o.igrid(1) > 0
    s0 = (u * u + v * v)[wet]
    np.testing.assert_allclose((u2 * u2 + v2 * v2)[wet], s0, rtol=1e-13)
    c = 0.5 * o.get("cor")[:, :, None] * o.cfg.dtbc
    cosang = np.broadcast_to((1 - c * c) / (1 + c * c), u.shape)[wet]
    np.testing.assert_allclose((u * u2 + v * v2)[wet], cosang * s0, rtol=1e-12, atol=1e-13 * s0.max())
    sinang = np.broadcast_to(2 * c / (1 + c * c), u.shape)[wet]
    # f > 0 in the northern basin: the rotation is clockwise (to the right)
    np.testing.assert_allclose((u * v2 - v * u2)[wet], -sinang * s0, rtol=1e-12, atol=1e-13 * s0.max())


def test_int_clin_ab3_integrates_quadratic_history():
    """Without Coriolis, the momentum tendency enters through the
    third-order Adams-Bashforth extrapolation of the advection history
    (auc, aup, aupp at steps 0, -1, -2), which integrates any quadratic
    history q(s) exactly over the next step: u' = up - dt (q0 + q1/2 + q2/3)
    for q(s) = q0 + q1 s + q2 s^2.  The reference's coefficients are the
    float32 values of 23/12, -16/12, 5/12 (default-REAL literals), so the
    identity holds to 1e-7."""
    o = Oracle(configs.box())
    rng = np.random.default_rng(20)
    _clin_quiet(o)
    o.set("cor", np.zeros((o.ni, o.nj)))
    q0, q1, q2 = (1e-6 * _baroclinic(o, rng) for _ in range(3))
    up = _baroclinic(o, rng)
    o.set("up", up)
    o.set("uc", up)
    o.set("vp", np.zeros_like(up))
    o.set("vc", np.zeros_like(up))
    o.set("auc", q0)
    o.set("aup", q0 - q1 + q2)
    o.set("aupp", q0 - 2 * q1 + 4 * q2)
    o.run_stage("int_clin")
    dt = o.cfg.dtbc
    want = up - dt * (q0 + q1 / 2 + q2 / 3)
    wet = o.igrid(1) > 0
    err = np.abs(o.get("uc") - want)[wet].max()
    step = dt * np.abs(q0 + q1 / 2 + q2 / 3)[wet].max()
    assert err < 1e-6 * step
    # a second-order (AB2) or first-order scheme misses by O(q2) dt
    assert dt * np.abs(q2 / 3)[wet].max() > 1e3 * err


# --------------------------------------------------------------------------
# int_ssh (mod_int.f90:602-624) and int_ts (mod_int.f90:442-507)

def test_int_ssh_of_reference_stratification_is_flat():
    """With the horizontal-mean T/S profile everywhere and mch = 1, the
    specific volume at each level is the one that built the vertical grid
    (pr1(k+1) - pr1(k) = rho g dz at the mid-layer pressure), so the
    hydrostatic column height sum(alpha dpr)/g equals the depth and the sea
    surface phih/g + sum(alpha dpr mch)/g is flat (0) to rounding."""
    cfg = _small_global()
    o = Oracle(cfg)
    ni, nj, nk = o.ni, o.nj, o.nk
    o.set("tp", np.broadcast_to(cfg.tmn[None, None, :], (ni, nj, nk)).copy())
    o.set("sp", np.broadcast_to(cfg.smn[None, None, :], (ni, nj, nk)).copy())
    o.set("mch", np.ones((ni, nj)))
    o.run_stage("int_ssh")
    ssh = o.get("ssh")
    wet = o.igrid(2) > 0
    depth = -o.get("phih_t") / C["g"]
    assert depth[wet].min() > 49.0          # phih is the bottom geopotential -g H
    assert np.abs(ssh[wet]).max() < 1e-14 * depth[wet].max() * 100


def test_int_ts_surface_restoring_budget():
    """At rest (no advection) with a horizontally uniform stable column and
    no vertical diffusion, int_ts changes only the top layer:
    the column heat content sum(dpr mch (T' - T)) changes by exactly the
    restoring flux dt gamma_t (T* - T_1), and the surface salinity relaxes
    by the fraction dt gamma_s toward S* (the reference's salt restoring is
    not divided by the layer thickness, mod_int.f90:470-478)."""
    cfg = configs.box()
    o = Oracle(cfg)
    rng = np.random.default_rng(24)
    ni, nj, nk = o.ni, o.nj, o.nk
    z3 = np.zeros((ni, nj, nk))
    for n in ("uc", "vc"):
        o.set(n, z3)
    o.set("wm", np.zeros((ni, nj, nk + 1)))
    o.set("kh", np.zeros((ni, nj, nk + 1)))
    t = np.broadcast_to(np.linspace(20.0, 2.0, nk)[None, None, :], (ni, nj, nk)).copy()
    sal = np.full((ni, nj, nk), 35.0)
    for n, v in (("tc", t), ("tp", t), ("sc", sal), ("sp", sal)):
        o.set(n, v)
    mch = 1.0 + 0.05 * _periodic(rng.uniform(size=(ni, nj)))
    o.set("mch", mch)
    o.set("mch2", mch)
    tstar = 20.0 + _periodic(rng.uniform(-1.0, 0.5, size=(ni, nj)))
    sstar = 35.0 + _periodic(rng.uniform(-0.5, 0.5, size=(ni, nj)))
    o.set("bnd_t", tstar)
    o.set("bnd_s", sstar)
    o.run_stage("int_ts")
    tn, sn = o.get("tc"), o.get("sc")
    wet = o.igrid(2) > 0
    dt = o.cfg.dtbc
    dpr = o.vgrid(2, 3)
    np.testing.assert_array_equal(tn[:, :, 1:][wet], t[:, :, 1:][wet])
    np.testing.assert_array_equal(sn[:, :, 1:][wet], sal[:, :, 1:][wet])
    heat = ((tn - t) * dpr[None, None, :]).sum(2) * mch
    want = dt * C["gamma_t"] * (tstar - t[:, :, 0])
    # T' - T loses the digits of T itself: a few ulp(20 K) per layer
    np.testing.assert_allclose(heat[wet], want[wet], rtol=1e-12, atol=4e-14 * 20.0 * dpr[0] * mch.max())
    np.testing.assert_allclose((sn[:, :, 0] - 35.0)[wet], (dt * C["gamma_s"] * (sstar - 35.0))[wet],
                               rtol=1e-9, atol=4 * np.spacing(35.0))
