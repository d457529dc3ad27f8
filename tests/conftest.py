"""Test configuration: the `gpu` marker and shared fixtures."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ctx():
    from gschism_amd import core

    print("[gpu] creating context on device 0", flush=True)
    c = core.Context(0)
    yield c
    c.close()


def oracle_mesh_for(mesh, sigma):
    """Oracle mesh with exactly the device mesh's (renumbered) geometry."""
    from gschism_amd import _lib as L
    from oracle import pyoracle as O

    kb = mesh.topology(L.T_KB)
    return O.OracleMesh(mesh.topology(L.T_XY), mesh.topology(L.T_ELNODE), mesh.topology(L.T_DEPTH),
                        mesh.topology(L.T_OPEN_BND), sigma, kb=kb)


def assert_bitwise(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    if not np.array_equal(a, b):
        bad = np.flatnonzero(a.reshape(-1) != b.reshape(-1))
        i = bad[0]
        raise AssertionError(f"{what}: {bad.size} of {a.size} entries differ; first at {i}: "
                             f"{a.reshape(-1)[i]!r} vs {b.reshape(-1)[i]!r}")
