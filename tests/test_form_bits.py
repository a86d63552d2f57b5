"""The mixed forms' bits, pinned.

The FP64 form is pinned to the FP64 oracle bit for bit (test_gpu_parity.py).
The mixed forms have no bitwise reference: their accuracy is measured by the
one-year RMSE against FP64.  Kernel rewrites that keep every operation and
its order must keep these bits; this test records them (SHA-1 of nine fields
after 30 steps of the 1-degree configuration, measured when the 1-year RMSE
tables were made) so that such a rewrite is checked, not assumed.  A change
of the mixed forms' arithmetic on purpose updates the recorded values and
the RMSE tables together."""
import hashlib

import pytest

from pcom_amd import abi, configs

pytestmark = pytest.mark.gpu

FIELDS = ["tc", "sc", "uc", "vc", "chc", "mch2", "buc", "am", "ssh"]
EXPECTED = {
    "mixed": "cd35c12b6c63d32f",
    "mixed_eos32": "c3e441d83e6f1bb1",
    "fp64": "7aff16207ed32fb4",
}


@pytest.mark.parametrize("form", list(EXPECTED))
def test_form_bits_after_30_steps(gpu_lib, form):
    from pcom_amd.lib import Model
    m = Model(configs.global_1deg(), precision=abi.FORMS[form])
    m.step(30)
    h = hashlib.sha1(b"".join(m.get(n).tobytes() for n in FIELDS)).hexdigest()[:16]
    assert h == EXPECTED[form]
