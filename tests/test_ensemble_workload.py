"""The ensemble workload's members (CPU only: inputs, no stepping)."""
import numpy as np

from pcom_amd import configs, ensemble


def test_members_are_seeded_perturbations():
    cfgs = ensemble.member_configs(n2=3, n1=2)
    assert [d for d, _ in cfgs] == [2, 2, 2, 1, 1]
    b2 = configs.global_1deg(glo_ni=180, glo_nj=90)
    b1 = configs.global_1deg()
    for m, (deg, c) in enumerate(cfgs):
        base = b2 if deg == 2 else b1
        assert (c.glo_ni, c.glo_nj, c.nk) == (base.glo_ni, base.glo_nj, 30)
        assert np.array_equal(c.itn, base.itn) and np.array_equal(c.sa, base.sa)
        d = c.pt - base.pt
        rng = np.random.default_rng(100 + m)
        assert np.array_equal(d, (base.pt + 0.01 * rng.standard_normal(base.pt.shape)) - base.pt)
        assert 0.009 < d.std() < 0.011 and abs(d.mean()) < 1e-3
        assert c.pt.flags.f_contiguous
    # the base configurations are left as they were
    assert np.array_equal(b2.pt, configs.global_1deg(glo_ni=180, glo_nj=90).pt)


def test_sypd():
    # one day of 3600 s steps in 86400 s of wall time is one simulated day per day
    assert abs(ensemble.sypd(24, 3600, 86400.0) - 1.0 / 365.0) < 1e-15
