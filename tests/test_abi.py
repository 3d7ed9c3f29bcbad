"""The C-ABI library loads and exports every entry point include/edus_gpu.h
declares; config composition follows the reference's schema and unit rules.
No GPU calls here."""
import ctypes
import re

import pytest

import edus
from edus import config


def _declared_functions():
    text = open(edus.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(edus_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    lib = ctypes.CDLL(edus.LIB_PATH)
    names = _declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(edus.EXPORTS) <= set(names)


def test_no_oracle_in_product():
    """The product library must not link or name the CPU oracle."""
    data = open(edus.LIB_PATH, "rb").read()
    assert b"liboracle" not in data and b"oracle_create" not in data


def test_config_defaults_and_unit_quirk():
    p = config.compose({"lasers": [{"intensity": 1.0}]})
    assert p["grid"] == [10, 10, 1] and p["solver"] == "RK" and p["order"] == 4
    assert p["printresolution_pulse"] == 0 and p["dt_units"] == "autime"
    # schema default "Wcm2" is kept verbatim; the C++ unit() maps it to the null unit
    assert p["lasers"][0]["intensity_units"] == "Wcm2"
    assert p["r0"] == [10.0]
    assert config.compose({"r0": 40})["r0"] == [40.0]
    assert config.compose({"peierls": True})["peierls"] is True
    assert config.compose({"gradient_space": "k"})["gradient_space"] == "k"
    with pytest.raises(ValueError):
        config.compose({"gradient_space": "x"})


def test_lindhard_rejects_bad_shapes_before_any_device_call():
    """edus_lindhard validates its arguments on the host first (nb in [2, 16],
    0 <= filledbands <= nb, matching array shapes); no GPU is touched."""
    import numpy as np

    U = np.zeros((6, 2, 2), dtype=np.complex128)
    e = np.zeros((6, 2))
    with pytest.raises(ValueError):
        edus.lindhard(U, e, 4, 2, 1)  # 6 k-points for a 4 x 2 grid
    with pytest.raises(edus.EdusError):
        edus.lindhard(np.zeros((6, 1, 1), dtype=np.complex128), np.zeros((6, 1)), 3, 2, 0)  # nb = 1
    with pytest.raises(edus.EdusError):
        edus.lindhard(U, e, 3, 2, 3)  # filledbands > nb
    big = np.zeros((2, 17, 17), dtype=np.complex128)
    with pytest.raises(edus.EdusError):
        edus.lindhard(big, np.zeros((2, 17)), 2, 1, 8)  # nb = 17


def test_toprint_keys():
    """The reference's "toprint" flags are strings; the density matrices are
    written, fullH and SelfEnergy are refused."""
    p = config.compose({"toprint": {"DMk_bloch": "true"}})
    assert p["toprint"] == {"DMk_wannier": "false", "DMk_bloch": "true", "fullH": "false", "SelfEnergy": "false"}
    assert p["write_h5"] is False
    for key in ("fullH", "SelfEnergy"):
        with pytest.raises(ValueError):
            config.compose({"toprint": {key: "true"}})
