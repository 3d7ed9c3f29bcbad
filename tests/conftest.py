import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "edus_2.0_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def hbn_params(coulomb=False, **overrides):
    """The reference's ci-test input (ci-test/inputs/hBN.json.in) with the
    tight-binding path pointed at the committed copy of the model."""
    from edus import config

    name = "hBN_coulomb/hBN_coulomb.json.in" if coulomb else "hBN/hBN.json.in"
    with open(os.path.join(GOLDEN, name)) as f:
        user = json.load(f)
    user["tb_file"] = os.path.join(GOLDEN, "models", "hBN_gap7.25eV_a2.5A")
    user.update(overrides)
    return config.compose(user)


@pytest.fixture(scope="session")
def model_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("models"))
