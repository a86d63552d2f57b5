"""pcom_output_means_noalloc, the output entry point that allocates
nothing (its staging is the library's static device buffer), gives the
bits of pcom_output_means: on two calls in a row, the second on a smaller
grid, and on a second interval after the accumulators' reset; a grid above
the buffer's limit is refused."""
import numpy as np
import pytest

from pcom_amd import abi, configs

pytestmark = pytest.mark.gpu


def test_output_means_noalloc_matches_output_means(gpu_lib):
    from pcom_amd.lib import Model
    for cfg, form in ((configs.global_1deg(), abi.PCOM_MIXED), (configs.box(), abi.PCOM_FP64)):
        print(f"noalloc means: {cfg.name} {cfg.glo_ni}x{cfg.glo_nj}x{cfg.nk}", flush=True)
        a, b = Model(cfg, precision=form), Model(cfg, precision=form)
        for n in (5, 3):
            a.step(n)
            b.step(n)
            ma = a.output_means()
            mb = b.output_means(noalloc=True)
            for k in abi.MEAN_NAMES:
                assert ma[k].shape == mb[k].shape
                np.testing.assert_array_equal(mb[k], ma[k], err_msg=f"{cfg.name} after {n} steps: {k}")
            assert np.isfinite(mb["pt"]).all()
        a.close()
        b.close()


def test_output_means_noalloc_refuses_grids_above_its_buffer(gpu_lib):
    from pcom_amd.lib import Model, PcomError
    cfg = configs.global_1deg(glo_ni=1440, glo_nj=720, nk=3)
    cfg.dtbt, cfg.dtbc = 15, 900
    m = Model(cfg)
    m.step(1)
    with pytest.raises(PcomError) as e:
        m.output_means(noalloc=True)
    assert e.value.code == abi.PCOM_ERR_ARG and "static buffer" in str(e.value)
    m.output_means()  # the allocating entry point still serves this grid
