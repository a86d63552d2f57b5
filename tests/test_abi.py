"""The C ABI: the built library loads without a GPU and exports every
symbol include/pcom_gpu.h declares; the ctypes mirror matches the header."""
import ctypes as C
import os
import re

from pcom_amd import abi
from pcom_amd import lib as plib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pcom_gpu.h")


def _header():
    with open(HEADER) as f:
        return f.read()


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", _header(), flags=re.S)
    return sorted(set(re.findall(r"\b(pcom_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(plib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(plib.EXPORTS) == declared_functions()


def test_field_ids_match_header():
    text = _header()
    body = text[text.index("PCOM_F_TC = 0"):text.index("PCOM_F_COUNT")]
    names = re.findall(r"PCOM_F_([A-Z0-9_]+)", body)
    assert [n.lower() for n in names] == abi.FIELD_NAMES


def test_struct_layouts():
    # pcom_inputs: 3 ints, then 15 pointers (8-byte aligned)
    assert C.sizeof(abi.PcomInputs) == 16 + 15 * 8
    assert C.sizeof(abi.PcomOpts) == 13 * 4  # 13 ints, trop_mode last
    assert C.sizeof(abi.PcomCounters) == 8 + 11 * 4 + 4
    assert C.sizeof(abi.PcomForcing) == 4 * 8


def test_no_gpu_error_is_reported_not_aborted():
    """Without a device pcom_create returns an error code and a message
    (the library never aborts, never falls back to the CPU)."""
    import pytest
    from pcom_amd import configs
    try:
        m = plib.Model(configs.small())
    except plib.PcomError as e:
        assert e.code in (abi.PCOM_ERR_NOGPU, abi.PCOM_ERR_HIP)
        assert "device" in str(e).lower() or "hip" in str(e).lower()
        return
    m.close()
    pytest.skip("a GPU is visible")
