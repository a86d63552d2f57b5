"""The oracle's OpenMP loops change nothing: the build used everywhere
(liboracle.so, -fopenmp, several host threads) gives the bits of the same
source built serially, over steps that run every routine (daily forcing,
convection, the subcycle) on three grids."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle.oracle import Oracle, threads
from pcom_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIELDS = ["tc", "sc", "uc", "vc", "auc", "avc", "fx", "fy", "am", "alpha", "adp", "wm", "prho_t", "prho_s",
          "act", "acr", "acu", "acw", "chc", "mch", "mch2", "buc", "bvc", "but", "bvt", "bphi_xn", "bgra_yw", "ssh"]

_RUN = """
import sys
import numpy as np
sys.path.insert(0, {root!r})
import oracle.oracle as oo
oo._HERE = {here!r}
from pcom_amd import configs
cfg = {cfg}
o = oo.Oracle(cfg)
o.step({n})
np.savez({out!r}, **{{f: o.get(f) for f in {fields!r}}})
"""


@pytest.mark.parametrize("name,cfg,n", [
    ("small", "configs.small()", 6),
    ("box_pa", "configs.box_pa()", 26),
    ("global", "configs.global_1deg(glo_ni=72, glo_nj=36, nk=12)", 4),
])
def test_threaded_oracle_is_bitwise_serial(tmp_path, name, cfg, n):
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-std=c11",
                           "-Wno-unknown-pragmas", "-I", os.path.join(ROOT, "oracle"), "-shared", "-o",
                           str(tmp_path / "liboracle.so"), os.path.join(ROOT, "oracle", "pcom_oracle.c"), "-lm"])
    out = str(tmp_path / "serial.npz")
    code = _RUN.format(root=ROOT, here=str(tmp_path), cfg=cfg, n=n, out=out, fields=FIELDS)
    subprocess.run([sys.executable, "-c", code], check=True, cwd=ROOT, timeout=600)
    serial = np.load(out)
    threads(4)
    assert threads() == 4
    o = Oracle(eval(cfg))
    o.step(n)
    for f in FIELDS:
        np.testing.assert_array_equal(o.get(f), serial[f], err_msg=f"{name}: {f}")
