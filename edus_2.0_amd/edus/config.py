"""Reference-format simulation input -> normalised parameter dict.

The reference reads one JSON file whose keys and defaults are given by its
schema (src/InputVariables/input_schema.json; defaults composed by
src/InputVariables/simulation_parameters.cpp:17-90).  This module applies the
same defaults and keeps the unit strings verbatim, so the C++ side can resolve
them with the reference's own unit() rules (src/ConvertUnits.cpp:22-37),
including its quirk that only lower-case names are known ("eV", "Wcm2", "au"
resolve to a null unit of value 0).
"""
from __future__ import annotations

import copy
import json
import os

LASER_DEFAULTS = {
    "intensity": 0.0, "intensity_units": "Wcm2",
    "frequency": 0.0, "frequency_units": "eV",
    "wavelength": 0.0, "wavelength_units": "nm",
    "polarization": [0.0, 0.0, 0.0],
    "t0": 0.0, "t0_units": "fs",
    "cycles": 0.0, "phase": 0.0,
}

DEFAULTS = {
    "tb_file": "",
    "coulomb": False,
    "opengap": 0, "opengap_units": "electronvolt",
    "epsilon": 1.0,
    "r0": [10.0], "r0_units": "angstrom",
    "grid": [10, 10, 1],
    "filledbands": 0,
    "dt": 0.01, "dt_units": "autime",
    "gradient_space": "R",
    "printresolution": 1, "printresolution_pulse": 0,
    "initialtime": 0.0, "initialtime_units": "autime",
    "finaltime": 0.0, "finaltime_units": "autime",
    "solver": "RK", "order": 4,
    "peierls": False,
    "toprint": {"DMk_wannier": "false", "DMk_bloch": "false", "fullH": "false", "SelfEnergy": "false"},
    "lasers": [],
    # engine extensions (not reference keys)
    "coulomb_center_quirk": True,   # keep ModelCoulomb.cpp:76-79 behaviour
    "write_h5": False,              # write output.h5 next to the text outputs
}


def compose(user: dict, base_dir: str | None = None) -> dict:
    """Defaults overlaid with the user's keys (simulation_parameters.cpp:94-103)."""
    out = copy.deepcopy(DEFAULTS)
    for k, v in user.items():
        if k == "lasers":
            out["lasers"] = []
            for las in v:
                d = copy.deepcopy(LASER_DEFAULTS)
                d.update(las)
                out["lasers"].append(d)
        else:
            out[k] = copy.deepcopy(v)
    if not isinstance(out["r0"], (list, tuple)):
        out["r0"] = [float(out["r0"])]
    if base_dir and out["tb_file"] and not os.path.isabs(out["tb_file"]):
        out["tb_file"] = os.path.join(base_dir, out["tb_file"])
    if out["gradient_space"] not in ("R", "k"):
        raise ValueError("gradient_space must be 'R' or 'k'")
    tp = dict(DEFAULTS["toprint"])
    tp.update(out.get("toprint") or {})
    out["toprint"] = tp
    for key in ("fullH", "SelfEnergy"):
        if str(tp[key]) == "true":
            raise ValueError(f"toprint {key} is not written by this engine (only DMk_wannier and DMk_bloch)")
    if len(out["lasers"]) > 10:
        raise ValueError("at most 10 lasers (input_schema.json maxItems)")
    return out


def load(path: str, **overrides) -> dict:
    with open(path) as f:
        user = json.load(f)
    user.update(overrides)
    return compose(user, base_dir=os.path.dirname(os.path.abspath(path)))
