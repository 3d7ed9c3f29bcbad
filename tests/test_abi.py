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
