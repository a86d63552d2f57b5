"""Learned barotropic surrogate: the CPU network restatement pinned by a
brute-force convolution, BF16 rounding cases and finite differences; the
HIP forward/backward (BF16 MFMA) against it; training and use in the time
stepper, checked by the plain subcycle."""
import numpy as np
import pytest

from oracle import surrogate_ref as ref
from pcom_amd import configs


def _random_params(rng, scale=1.0):
    parts = []
    for ci, co in ref.DIMS:
        used_ci = ref.IN_USED if ci == ref.IN_C and len(parts) == 0 else ci
        w = np.zeros((co, 3, 3, ci))
        w[:, :, :, :used_ci] = rng.standard_normal((co, 3, 3, used_ci)) * scale * np.sqrt(2.0 / (9 * used_ci))
        parts += [w.ravel(), 0.1 * rng.standard_normal(co)]
    parts.append(np.ones(ref.N_NORM))
    return np.concatenate(parts)


def test_conv_matches_brute_force():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 3, 5, 4))
    w = rng.standard_normal((3, 3, 3, 4))
    b = rng.standard_normal(3)
    out = ref.conv(x, w, b)
    bf = np.zeros_like(out)
    for n in range(2):
        for y in range(3):
            for xx in range(5):
                for co in range(3):
                    s = b[co]
                    for dy in range(3):
                        for dx in range(3):
                            yy = y + dy - 1
                            if 0 <= yy < 3:
                                s += np.dot(w[co, dy, dx], x[n, yy, (xx + dx - 1) % 5])
                    bf[n, y, xx, co] = s
    np.testing.assert_allclose(out, bf, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dil", [2, 4])
def test_dilated_conv_matches_brute_force(dil):
    """Taps dil pixels apart, periodic in x (also past a full period) and
    zero outside the rows."""
    rng = np.random.default_rng(dil)
    x = rng.standard_normal((1, 9, 5, 2))
    w = rng.standard_normal((2, 3, 3, 2))
    b = np.zeros(2)
    out = ref.conv(x, w, b, dil)
    bf = np.zeros_like(out)
    for y in range(9):
        for xx in range(5):
            for co in range(2):
                s = 0.0
                for dy in range(3):
                    for dx in range(3):
                        yy = y + (dy - 1) * dil
                        if 0 <= yy < 9:
                            s += np.dot(w[co, dy, dx], x[0, yy, (xx + (dx - 1) * dil) % 5])
                bf[0, y, xx, co] = s
    np.testing.assert_allclose(out, bf, rtol=1e-12, atol=1e-12)
    # the data gradient is the adjoint of the convolution
    d = rng.standard_normal(out.shape)
    lhs = np.sum(ref.conv(x, w, b, dil) * d)
    rhs = np.sum(x * ref.conv_data_grad(d, w, dil))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    g = ref.conv_weight_grad(x, d, dil)
    assert abs(np.sum(g * w) - lhs) <= 1e-10 * abs(lhs)


def test_bf16_rounding_cases():
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 3.0e-3, 65504.0], np.float32)
    got = ref.bf16(v)
    np.testing.assert_array_equal(got, np.array([1.0, 1.0, 1 + 2 ** -6, -2.5, 0.0030059814453125, 65536.0],
                                                np.float32))


def test_gradients_match_finite_differences():
    """The float64 backward pass against central differences of the loss."""
    rng = np.random.default_rng(3)
    p = _random_params(rng)
    x = rng.standard_normal((1, 3, 4, ref.IN_USED))
    x[..., 5] = (rng.uniform(size=(1, 3, 4)) > 0.3)
    x[..., 6] = (rng.uniform(size=(1, 3, 4)) > 0.3)
    t = rng.standard_normal((1, 3, 4, ref.OUT_USED))
    loss, _, g = ref.loss_and_grads(p, x, t, rnd=False)
    off = 0
    checked = 0
    for ci, co in ref.DIMS:
        nw = 9 * ci * co
        cand = [off + int(k) for k in rng.choice(nw, 40, replace=False)] + [off + nw + int(rng.integers(co))]
        for k in cand:
            if p[k] == 0 and g[k] == 0:
                continue
            h = 1e-6 * max(1.0, abs(p[k]))
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            fd = (ref.loss_and_grads(pp, x, t, rnd=False)[0] - ref.loss_and_grads(pm, x, t, rnd=False)[0]) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(abs(fd), 1e-3 * loss), (k, fd, g[k])
            checked += 1
        off += nw + co
    assert checked > 100


def _inputs(cfg, rng, batch):
    m_h, m_w = cfg.glo_nj, cfg.glo_ni
    x = rng.standard_normal((batch, m_h, m_w, ref.IN_USED)).astype(np.float32)
    sea_t = (cfg.itn.T > 0).astype(np.float32)
    x[..., 5] = sea_t
    x[..., 6] = sea_t * np.roll(sea_t, -1, axis=1)
    x[..., 7] = np.linspace(0.3, 1.0, m_h)[None, :, None]
    t = rng.standard_normal((batch, m_h, m_w, ref.OUT_USED)).astype(np.float32)
    return x, t


@pytest.mark.gpu
def test_gpu_forward_backward_matches_oracle(gpu_lib):
    from pcom_amd.lib import Model
    cfg = configs.box()
    m = Model(cfg)
    m.nn_create(capacity=0, seed=5)
    rng = np.random.default_rng(11)
    p = _random_params(rng).astype(np.float32)
    m.nn_params(p)
    x, t = _inputs(cfg, rng, 2)
    loss, out, g = m.nn_forward_backward(x, t)
    rl, rout, rg = ref.loss_and_grads(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), rnd=True)
    # BF16 activations rounded from FP32 (device) or float64 (here) sums
    # differ by one BF16 ulp now and then; tolerances are set by that
    assert abs(loss - rl) <= 2e-3 * rl
    assert np.sqrt(np.mean((out - rout) ** 2)) <= 1e-2 * np.sqrt(np.mean(rout ** 2))
    off = 0
    for li, (ci, co) in enumerate(ref.DIMS):
        n = 9 * ci * co + co
        gg, rr = g[off:off + n], rg[off:off + n]
        assert np.linalg.norm(gg - rr) <= 3e-2 * np.linalg.norm(rr), f"layer {li}"
        off += n


@pytest.mark.gpu
def test_gpu_surrogate_trains_and_beats_persistence(gpu_lib):
    """Snapshots from plain box steps, a short training run, then steps
    with the network checked by the plain subcycle every 4th step; its
    first step is a check too (which sets the limit of its increments)."""
    from pcom_amd.lib import Model
    cfg = configs.box()
    m = Model(cfg)
    m.nn_create(capacity=96, seed=2)
    m.nn_record(True)
    m.step(96)
    m.nn_record(False)
    assert m.nn_stats().recorded == 96
    loss = m.nn_train(3000, batch=4, lr=2e-3, seed=3)
    assert np.all(np.isfinite(loss))
    assert loss[-100:].mean() < 0.2 * loss[:10].mean()
    plain = Model(cfg)
    plain.step(96)
    m.nn_use(4)
    m.step(16)
    plain.step(16)
    st = m.nn_stats()
    assert (st.surrogate_steps, st.checks) == (11, 5)
    assert m.check_finite()
    err = np.array(st.last_check_rmse[:])
    per = np.array(st.last_persist_rmse[:])
    assert np.all(np.isfinite(err))
    # the network's free surface and transport beat "no change over the step"
    assert err[0] < per[0] and err[2] < per[2] and err[3] < per[3], (err, per)
    # mass: the surrogate removes its mean free-surface increment
    s_nn, s_pl = m.global_sums()[2], plain.global_sums()[2]
    assert abs(s_nn - s_pl) <= 1e-9 * abs(s_pl)


@pytest.mark.gpu
def test_gpu_snapshot_scale_is_sea_rms_of_velocity_increment(gpu_lib):
    """Each snapshot's first-substep increment r is divided by s, the RMS
    of (r_u, r_v) over sea U points, so over sea mean(r_u^2 + r_v^2) / s^2
    is 2 in every snapshot; the normalisation's root-mean-square scales of
    r_u / s and r_v / s over all interior pixels and snapshots therefore
    satisfy in_std(u)^2 + in_std(v)^2 = 2 x (sea fraction of U points).
    Parameters end with the normalisation: input means (12), input scales
    (12), output scales (4), Coriolis and bottom-pressure scales, the scale
    floor (a fifth of the median scale, which these snapshots stay above)."""
    from pcom_amd.lib import Model
    cfg = configs.box()
    m = Model(cfg)
    m.step(4)
    m.nn_create(capacity=8, seed=2)
    m.nn_record(True)
    m.step(6)
    m.nn_record(False)
    m.nn_train(1, batch=2, lr=1e-4, seed=3)
    norm = m.nn_params()[-31:]
    in_std = norm[12:24]
    assert norm[30] > 0.0  # a scale floor was set
    sea_u = m.get("prh_u")[1:-1, 1:-1] != np.float64(np.float32(0.1))
    frac = sea_u.mean()
    assert 0.1 < frac < 1.0
    np.testing.assert_allclose(in_std[1] ** 2 + in_std[2] ** 2, 2.0 * frac, rtol=1e-4)


@pytest.mark.gpu
def test_gpu_recording_in_use_leaves_the_network_run_unchanged(gpu_lib):
    """pcom_nn_record while the network is in use: every step records a
    snapshot (the network's steps with the plain subcycle run beside them
    and undone, the checks as plain steps), and the run keeps the bits of
    the same network stepping without recording."""
    from pcom_amd.lib import Model
    cfg = configs.box()
    a = Model(cfg)
    a.nn_create(capacity=48, seed=2)
    a.nn_record(True)
    a.step(24)
    a.nn_record(False)
    a.nn_train(200, batch=4, lr=2e-3, seed=3)
    b = Model(cfg)
    b.step(24)
    b.nn_create(capacity=0, seed=5)
    b.nn_params(a.nn_params())
    a.nn_use(4, tol=0.0)
    b.nn_use(4, tol=0.0)
    a.nn_record(True)
    a.step(16)
    a.nn_record(False)
    b.step(16)
    sa, sb = a.nn_stats(), b.nn_stats()
    assert sa.recorded == 24 + 16
    assert (sa.surrogate_steps, sa.checks) == (sb.surrogate_steps, sb.checks) == (11, 5)
    for n in ("tc", "sc", "uc", "vc", "chc", "chp", "mch", "mch2", "buc", "bvc", "bup", "bvp", "acch", "am"):
        assert np.array_equal(a.get(n), b.get(n)), n


@pytest.mark.gpu
def test_gpu_failed_check_holds_then_resumes(gpu_lib):
    """pcom_nn_resume: with a bound no prediction meets (tol = 1e-300), every
    check fails; each failure runs 3 plain steps and the network resumes
    with a check.  Checks every 2nd step, and the network's first step is a
    check (it sets the limit of the increments): checks at steps 1, 5 and 9,
    all 10 steps plain.  Later, with the limit set, the cadence alone rules
    (network at step 11, check at 12); without resume the first failure
    ends the surrogate."""
    from pcom_amd.lib import Model
    cfg = configs.box()
    m = Model(cfg)
    m.nn_create(capacity=8, seed=2)
    m.nn_record(True)
    m.step(8)
    m.nn_record(False)
    m.nn_train(20, batch=2, lr=1e-3, seed=3)
    m.nn_use(2, tol=1e-300, resume=3)
    m.step(10)
    st = m.nn_stats()
    assert (st.surrogate_steps, st.checks, st.fallbacks, st.plain_steps) == (0, 3, 3, 10 + 8)
    assert m.check_finite()
    m.nn_use(2, tol=1e-300, resume=0)
    m.step(6)
    st = m.nn_stats()
    assert (st.surrogate_steps, st.checks, st.fallbacks, st.in_use) == (1, 4, 4, 0)
