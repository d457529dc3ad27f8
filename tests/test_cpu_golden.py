"""The oracle reproduces the committed golden fixtures of the wave-equation step bit for bit."""
from __future__ import annotations

import importlib.util
import os

import numpy as np
import pytest

from conftest import assert_bitwise

_HERE = os.path.dirname(os.path.abspath(__file__))


def load_golden_module():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(_HERE, "golden", "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


G = load_golden_module()


def load_fixture(name):
    with np.load(G.fixture_path(name), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.mark.parametrize("name", G.CASES)
def test_fixture_inputs_are_regenerated_exactly(name):
    fx = load_fixture(name)
    inp, _ = G.case_inputs(name)
    for k, v in inp.items():
        assert_bitwise(fx[k], v, f"{name} input {k}")


@pytest.mark.parametrize("name", G.CASES)
def test_oracle_reproduces_fixture(name):
    fx = load_fixture(name)
    inp = {k: fx[k] for k in ("xy", "elnode", "open_bnd", "depth", "kb", "sigma", "eta", "uvw", "su2")}
    out = G.run_oracle(inp, G.params_from(fx["params"]))
    assert set(out) <= set(fx)
    for k, v in out.items():
        assert_bitwise(v, fx[k], f"{name} {k}")


def test_fixtures_exercise_the_boundaries():
    """The fixtures cover open-boundary exits and land slides, dry entities and bottom levels."""
    isl = load_fixture("island-20x14")
    bits = np.bitwise_or.reduce(isl["bstatus_0"].reshape(-1))
    assert bits & 8 and bits & 16
    assert (isl["foot_0"] < 0).any() and (isl["kb"] > 0).any() and (isl["depth"] < 0).any()
