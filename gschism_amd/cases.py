"""Synthetic workloads: jittered Delaunay channels with bathymetry, elevation and velocity.

Seeds follow one plan: 0 mesh jitter, 1 elevation noise, 2 velocity modes, 3 bathymetry,
100+m ensemble member m, 7 request arrivals.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from ._lib import check, dptr

SEED_MESH, SEED_ETA, SEED_VEL, SEED_BATHY, SEED_ARRIVALS = 0, 1, 2, 3, 7


def channel(nx: int, ny: int, dx: float, jitter: float = 0.2, seed: int = SEED_MESH):
    """Jittered-grid Delaunay channel: (xy [n,2], elnode [ne,3], open_bnd [n]); west edge open."""
    g = C.c_void_p()
    check(L.lib().gsh_gen_channel(nx, ny, dx, jitter, seed, C.byref(g)))
    try:
        n, ne = C.c_int32(), C.c_int32()
        check(L.lib().gsh_gen_sizes(g, C.byref(n), C.byref(ne)))
        xy = np.empty((n.value, 2), dtype=np.float64)
        en = np.empty((ne.value, 3), dtype=np.int32)
        ob = np.empty(n.value, dtype=np.int8)
        check(L.lib().gsh_gen_get(g, dptr(xy), dptr(en, C.c_int32), ob.ctypes.data_as(C.POINTER(C.c_int8))))
    finally:
        L.lib().gsh_gen_destroy(g)
    return xy, en, ob


def bathymetry(nx: int, ny: int, corr_cells: float = 20.0, hmin: float = 5.0, hmax: float = 4000.0,
               seed: int = SEED_BATHY) -> np.ndarray:
    out = np.empty(nx * ny, dtype=np.float64)
    check(L.lib().gsh_gen_bathymetry(nx, ny, corr_cells, hmin, hmax, seed, dptr(out)))
    return out


def velocity(pts, sigma, lx, ly, dx, ncomp, nmodes=8, umax=1.0, shear=0.5, seed=SEED_VEL) -> np.ndarray:
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    out = np.empty((pts.shape[0], sigma.size, ncomp), dtype=np.float64)
    check(L.lib().gsh_gen_velocity(pts.shape[0], dptr(pts), sigma.size, dptr(sigma), lx, ly, dx, nmodes, umax,
                                   shear, seed, ncomp, dptr(out)))
    return out


def normal(seed: int, count: int, mean: float = 0.0, std: float = 1.0) -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    check(L.lib().gsh_gen_normal(seed, count, mean, std, dptr(out)))
    return out


def sigma_levels(nvrt: int) -> np.ndarray:
    return -1.0 + np.arange(nvrt, dtype=np.float64) / (nvrt - 1)


@dataclass
class CaseSpec:
    """One synthetic configuration of the hot path."""
    name: str
    nx: int
    ny: int
    dx: float
    nvrt: int = 31
    depth: float | None = 10.0        # uniform depth; None = log-normal random field
    hmin: float = 5.0
    hmax: float = 4000.0
    corr_cells: float = 20.0
    dt: float = 150.0
    theta: float = 0.6
    substeps: int = 5
    eta_amp: float = 0.1
    eta_noise: float = 1e-3
    jitter: float = 0.2
    renumber: bool = True
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.nx * self.ny


def spec_cpu_reference() -> CaseSpec:
    """200 x 80 channel, 10 m deep, dt 150 s, theta 0.6, 5 ELM substeps (the CPU reference run)."""
    return CaseSpec("channel-200x80-h10", 200, 80, 500.0, depth=10.0, substeps=5)


def spec_gpu_1m(nx: int = 1024, ny: int = 1024) -> CaseSpec:
    """1024 x 1024 channel (1.05M nodes), dx 1 km, log-normal bathymetry 5-4000 m, 10 substeps."""
    return CaseSpec(f"channel-{nx}x{ny}-lognormal", nx, ny, 1000.0, depth=None, substeps=10)


def spec_mixed_4m() -> CaseSpec:
    return CaseSpec("channel-2048x2048-lognormal", 2048, 2048, 1000.0, depth=None, substeps=10)


def step_params(spec: CaseSpec) -> L.StepParams:
    dtb = spec.dt / spec.substeps
    # fixed substep count: dtb_min = dtb_max = dt / substeps
    return L.default_step_params(dt=spec.dt, theta=spec.theta, dtb_min=dtb, dtb_max=dtb)


@dataclass
class CaseData:
    spec: CaseSpec
    xy: np.ndarray
    elnode: np.ndarray
    open_bnd: np.ndarray
    depth: np.ndarray
    sigma: np.ndarray


def make_case(spec: CaseSpec) -> CaseData:
    xy, en, ob = channel(spec.nx, spec.ny, spec.dx, spec.jitter)
    if spec.depth is None:
        depth = bathymetry(spec.nx, spec.ny, spec.corr_cells, spec.hmin, spec.hmax)
    else:
        depth = np.full(xy.shape[0], spec.depth, dtype=np.float64)
    return CaseData(spec, xy, en, ob, depth, sigma_levels(spec.nvrt))


def initial_state(spec: CaseSpec, xy_mesh: np.ndarray, side_centres: np.ndarray, sigma: np.ndarray):
    """eta2 [n], nodal uvw [n,nvrt,3], side u2 [ns,nvrt,2] in the mesh's (renumbered) order.

    The elevation depends on x only and the velocity on position only, so the values do not
    depend on the node numbering; the noise is drawn in the generator's node order and mapped
    by the caller when a permutation is involved.
    """
    lx = (spec.nx - 1) * spec.dx
    ly = (spec.ny - 1) * spec.dx
    eta = spec.eta_amp * np.sin(2.0 * np.pi * xy_mesh[:, 0] / lx)
    uvw = velocity(xy_mesh, sigma, lx, ly, spec.dx, 3)
    su2 = velocity(side_centres, sigma, lx, ly, spec.dx, 2)
    return eta, uvw, su2


def side_centres(mesh) -> np.ndarray:
    xy = mesh.topology(L.T_XY)
    sn = mesh.topology(L.T_ISIDENODE)
    return 0.5 * (xy[sn[:, 0]] + xy[sn[:, 1]])


def load_mesh(ctx, data: CaseData, renumber: bool | None = None):
    """Builds the device mesh and its initial state; returns (mesh, params)."""
    from .core import Mesh

    spec = data.spec
    mesh = Mesh(ctx, data.xy, data.elnode, data.depth, data.sigma, open_bnd=data.open_bnd,
                renumber=spec.renumber if renumber is None else renumber)
    xy = mesh.topology(L.T_XY)
    eta, uvw, su2 = initial_state(spec, xy, side_centres(mesh), data.sigma)
    noise = normal(SEED_ETA, spec.n, 0.0, spec.eta_noise)[mesh.node_perm]
    eta = eta + noise
    p = step_params(spec)
    if ctx is not None:
        mesh.set(L.F_ETA2, eta)
        mesh.set(L.F_NODE_UVW, uvw)
        mesh.set(L.F_SIDE_U2, su2)
        mesh.update_levels(p)
    return mesh, p, (eta, uvw, su2)
