"""Ensembles (packed subcycle launches) and the subcycle autotuner: both
give the bits of plain per-context stepping."""
import numpy as np
import pytest

from pcom_amd import abi, configs

pytestmark = pytest.mark.gpu

FIELDS = ["tc", "sc", "uc", "vc", "chc", "chp", "mch", "mch2", "buc", "bvc", "acch", "am", "ssh"]


def _perturbed(cfg, seed, sd=0.01):
    rng = np.random.default_rng(seed)
    cfg.pt = np.asfortranarray(cfg.pt + sd * rng.standard_normal(cfg.pt.shape))
    return cfg


def _same(a, b, names=FIELDS):
    bad = []
    for n in names:
        x, y = a.get(n), b.get(n)
        if not np.array_equal(x, y):
            bad.append(f"{n}: max |d| = {np.nanmax(np.abs(x - y))}")
    return bad


def test_ensemble_matches_individual_runs(gpu_lib):
    """Members of three grid sizes and two forms, packed together."""
    from pcom_amd.lib import Ensemble, Model
    cfgs = [_perturbed(configs.box(), 100), _perturbed(configs.box(), 101), configs.small(),
            _perturbed(configs.global_1deg(glo_ni=180, glo_nj=90), 102), _perturbed(configs.box(), 103)]
    forms = [abi.PCOM_FP64, abi.PCOM_MIXED, abi.PCOM_FP64, abi.PCOM_FP64, abi.PCOM_FP64]
    modes = [abi.PCOM_TROP_AUTO] * 4 + [abi.PCOM_TROP_LAUNCHES]  # the last runs its subcycle alone
    members = [Model(c, precision=f, trop_mode=md) for c, f, md in zip(cfgs, forms, modes)]
    ens = Ensemble(members)
    ens.step(5)
    npk, of = ens.packs()
    assert npk < len(members) - 1, (npk, of)  # small members share launches
    assert of[-1] == -1
    for c, f, m in zip(cfgs, forms, members):
        solo = Model(c, precision=f)
        solo.step(5)
        bad = _same(m, solo)
        assert not bad, f"{c.name}: " + "; ".join(bad)


def test_pack_of_both_workgroup_sizes(gpu_lib):
    """A pack whose members' regions need the 768- and the 1024-thread
    subcycle kernels runs in the larger one; every member keeps its bits."""
    from pcom_amd.lib import Ensemble, Model
    cfgs = [_perturbed(configs.box(), 110), _perturbed(configs.box(), 111)]
    tiles = [(8, 8, 1), (20, 14, 3)]  # regions of 169 and 1015 points
    members = [Model(c, trop_tile=t) for c, t in zip(cfgs, tiles)]
    ens = Ensemble(members)
    ens.step(4)
    npk, of = ens.packs()
    assert npk == 1 and list(of) == [0, 0], (npk, of)
    for c, t, m in zip(cfgs, tiles, members):
        solo = Model(c, trop_tile=t)
        solo.step(4)
        bad = _same(m, solo)
        assert not bad, f"{t}: " + "; ".join(bad)


def test_autotuner_keeps_state_and_bits(gpu_lib):
    from pcom_amd.lib import Model
    cfg = configs.global_1deg()
    a, b = Model(cfg), Model(cfg)
    a.step(3)
    b.step(3)
    tile, ms = a.tune_subcycle(trials=4)
    assert ms > 0 and all(t > 0 for t in tile)
    assert not _same(a, b)  # tuning left the state as it was
    a.step(3)
    b.step(3)
    bad = _same(a, b)
    assert not bad, "; ".join(bad)


def test_tiling_set_and_read_back(gpu_lib):
    """pcom_subcycle_tiling: a tiling set on a live context gives the bits
    of a context created with it; a tiling that does not fit is refused and
    the old one kept."""
    from pcom_amd.lib import Model, PcomError
    cfg = configs.global_1deg(glo_ni=180, glo_nj=90)
    a = Model(cfg)
    b = Model(cfg, trop_tile=(26, 16, 2))
    default = a.tiling()
    assert all(t > 0 for t in default)
    a.step(2)
    b.step(2)
    assert a.tiling((26, 16, 2)) == (26, 16, 2)
    with pytest.raises(PcomError):
        a.tiling((40, 40, 4))  # tile plus halo above one workgroup's points
    assert a.tiling() == (26, 16, 2)
    a.step(3)
    b.step(3)
    assert not _same(a, b)
    assert a.tiling((0, 0, 0)) == default
