"""CPU-side checks of the C ABI: the library loads, exports every declared entry point, and its
host-side logic (generators, topology, argument checking) behaves."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

from gschism_amd import _lib as L
from gschism_amd import cases, core

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "gsh.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gsh_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(str(L.LIB_PATH))
    names = _declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes signature table covers them all
    assert set(names) <= set(L.SIGNATURES), sorted(set(names) - set(L.SIGNATURES))


def test_lib_loads_with_signatures():
    lib = L.lib()
    assert lib.gsh_version() == 100
    # the message is per thread and kept until the next entry point runs; a successful call clears it
    assert lib.gsh_grid_get(None, None, None, None, None, None, None) == L.GSH_E_ARG
    assert b"NULL" in lib.gsh_last_error()
    ncol = C.c_int32()
    ip = np.array([0, 1], dtype=np.int32)
    ix = np.array([0], dtype=np.int32)
    col = np.empty(1, dtype=np.int32)
    assert lib.gsh_colour_csr(1, ip.ctypes.data_as(C.POINTER(C.c_int32)), ix.ctypes.data_as(C.POINTER(C.c_int32)),
                              col.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(ncol)) == L.GSH_OK
    assert lib.gsh_last_error() == b""


def test_channel_generator_is_delaunay_and_ccw():
    xy, en, ob = cases.channel(30, 20, 100.0, 0.2, seed=0)
    assert xy.shape == (600, 2) and en.shape == (2 * 29 * 19, 3)
    p = xy[en]
    area = 0.5 * ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    assert (area > 0).all()
    # Delaunay: no node lies strictly inside the circumcircle of a neighbouring triangle
    edges = {}
    for t, tri in enumerate(en):
        for k in range(3):
            a, b = tri[(k + 1) % 3], tri[(k + 2) % 3]
            edges.setdefault((min(a, b), max(a, b)), []).append((t, tri[k]))
    worst = 0.0
    for (a, b), lst in edges.items():
        if len(lst) != 2:
            continue
        (t, _), (_, d) = lst
        A, B, Cc = xy[en[t]]
        D = xy[d]
        m = np.array([[A[0] - D[0], A[1] - D[1], (A - D) @ (A - D)], [B[0] - D[0], B[1] - D[1], (B - D) @ (B - D)],
                      [Cc[0] - D[0], Cc[1] - D[1], (Cc - D) @ (Cc - D)]])
        worst = max(worst, np.linalg.det(m))
    assert worst <= 1e-6 * 100.0 ** 4
    # open boundary is the west edge
    assert np.array_equal(ob.astype(bool), np.isclose(xy[:, 0], 0.0))


def test_generators_deterministic():
    a = cases.normal(7, 1000)
    b = cases.normal(7, 1000)
    assert np.array_equal(a, b) and abs(a.mean()) < 0.15 and abs(a.std() - 1) < 0.1
    h = cases.bathymetry(64, 48, 20.0, 5.0, 4000.0)
    assert np.isclose(h.min(), 5.0) and np.isclose(h.max(), 4000.0)
    assert np.array_equal(h, cases.bathymetry(64, 48, 20.0, 5.0, 4000.0))


def test_velocity_is_divergence_free_and_scaled():
    sig = cases.sigma_levels(5)
    g = np.stack(np.meshgrid(np.linspace(0, 5e4, 60), np.linspace(0, 3e4, 40)), -1).reshape(-1, 2)
    u = cases.velocity(g, sig, 5e4, 3e4, 500.0, 2)
    top = u[:, -1, :]
    assert np.sqrt((top ** 2).sum(1)).max() <= 1.0 + 1e-12
    assert np.allclose(u[:, 0, :], 0.5 * top)  # shear factor 1 + 0.5 sigma at sigma = -1
    # central differences with a 1 m step: the divergence vanishes to truncation error
    rng = np.random.default_rng(1)
    c = rng.uniform([0, 0], [5e4, 3e4], size=(50, 2))
    h = 1.0
    off = np.array([[h, 0], [-h, 0], [0, h], [0, -h]])
    pts = (c[:, None, :] + off[None]).reshape(-1, 2)
    w = cases.velocity(pts, sig, 5e4, 3e4, 500.0, 2)[:, -1, :].reshape(50, 4, 2)
    dudx = (w[:, 0, 0] - w[:, 1, 0]) / (2 * h)
    dvdy = (w[:, 2, 1] - w[:, 3, 1]) / (2 * h)
    scale = np.abs(dudx).max() + np.abs(dvdy).max()
    assert np.abs(dudx + dvdy).max() < 1e-6 * scale


def test_host_mesh_topology_errors():
    xy = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], dtype=np.float64)
    sig = cases.sigma_levels(3)
    # clockwise element
    with pytest.raises(L.GshError) as e:
        core.Mesh(None, xy, np.array([[0, 2, 1]]), np.ones(4), sig, renumber=False)
    assert e.value.code == L.GSH_E_TOPOLOGY
    # hanging node 3
    with pytest.raises(L.GshError) as e:
        core.Mesh(None, xy, np.array([[0, 1, 2]]), np.ones(4), sig, renumber=False)
    assert e.value.code == L.GSH_E_TOPOLOGY
    # bad kb
    with pytest.raises(L.GshError) as e:
        core.Mesh(None, xy, np.array([[0, 1, 2], [1, 3, 2]]), np.ones(4), sig, kb=np.array([0, 0, 5, 0]))
    assert e.value.code == L.GSH_E_ARG


def test_host_only_mesh_rejects_device_ops():
    xy, en, ob = cases.channel(8, 6, 100.0)
    m = core.Mesh(None, xy, en, np.full(48, 5.0), cases.sigma_levels(4), open_bnd=ob)
    with pytest.raises(L.GshError) as e:
        m.get(L.F_ETA2)
    assert e.value.code == L.GSH_E_STATE


def test_rcm_reduces_bandwidth():
    xy, en, ob = cases.channel(60, 40, 100.0)
    perm = np.random.default_rng(0).permutation(xy.shape[0])
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    xy_s, en_s = xy[perm], inv[en]
    m = core.Mesh(None, xy_s, en_s, np.full(xy.shape[0], 5.0), cases.sigma_levels(3), open_bnd=ob[perm])
    ip, ix = m.topology(L.T_CSR_INDPTR), m.topology(L.T_CSR_INDICES)
    rows = np.repeat(np.arange(m.n), np.diff(ip))
    bw = np.abs(rows - ix).max()
    assert bw <= 2 * 40 + 5
    # permutation maps back to the input coordinates
    assert np.array_equal(m.topology(L.T_XY), xy_s[m.node_perm])


def test_ssor_colouring_matches_oracle():
    """The library's SSOR colouring (host code) equals the oracle's on mesh and random patterns."""
    import scipy.sparse as sp

    from gschism_amd import cases, core
    from oracle import pyoracle as O

    xy, en, ob = cases.channel(40, 30, 100.0)
    m = core.Mesh(None, xy, en, np.full(xy.shape[0], 5.0), cases.sigma_levels(3), open_bnd=ob)
    pats = [(m.topology(L.T_CSR_INDPTR), m.topology(L.T_CSR_INDICES))]
    A = sp.random(3000, 3000, density=4e-3, random_state=3, format="csr")
    A = (A + A.T + sp.eye(3000)).tocsr()
    pats.append((A.indptr, A.indices))
    for ip, ix in pats:
        col, nc = core.colour_csr(ip, ix)
        ocol, onc = O.colour(ip, ix)
        assert nc == onc
        assert np.array_equal(col, ocol)
    with pytest.raises(L.GshError):
        core.colour_csr(np.array([0, 1], np.int32), np.array([5], np.int32))


def test_mirror_constructor_checks():
    """PCGSolverMultiGPU rejects inconsistent partition arguments before touching a device, as
    the reference constructor does (pcg_solver.py:547-560)."""
    from gschism_amd import pcg_solver as ps

    cfg = ps.CudaDeviceConfig([0])
    with pytest.raises(ValueError):
        ps.PCGSolverMultiGPU(cfg, [10], [20, 30])
    with pytest.raises(ValueError):
        ps.PCGSolverMultiGPU(ps.CudaDeviceConfig([0, 1]), [10], [20])
    with pytest.raises(ValueError):
        ps.PCGSolverMultiGPU(ps.CudaDeviceConfig([0, 1]), [10, 10], [20, 20])


def test_c_harness_builds_and_links():
    """The plain-C caller of the C ABI compiles with -Werror against include/gsh.h, links to the
    in-tree library and runs its device-free query."""
    import importlib.util
    import subprocess
    from pathlib import Path

    here = Path(__file__).resolve().parent / "c_abi"
    spec = importlib.util.spec_from_file_location("c_abi_build", here / "build.py")
    hb = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(hb)
    exe = hb.build()
    r = subprocess.run([str(exe), "--version"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.startswith("gsh "), (r.stdout, r.stderr)
