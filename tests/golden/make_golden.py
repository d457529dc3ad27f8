"""Writes the golden fixtures of the wave-equation step from the oracle.

Each fixture holds a small mesh, its initial state and step parameters, and the oracle's
outputs of two consecutive whole steps (levels, backtracking, RHS producers, assembly with
essential BC, PCG with the extrapolated guess, elevation update, momentum, element vertical
velocity, levels, nodal velocities; the second step backtracks through the velocities the
first produced). They pin the oracle against regression and give the device library fixed
vectors to reproduce bit for bit.

    python tests/golden/make_golden.py      # rewrites tests/golden/*.npz
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from gschism_amd import _lib as L  # noqa: E402
from gschism_amd import cases  # noqa: E402
from oracle import pyoracle as O  # noqa: E402

STEPS = 2
PARAM_KEYS = ("dt", "theta", "h0", "ramp", "dtb_min", "dtb_max", "velmin_btrack", "ihhat", "iettype", "rmaxvel",
              "dfv0", "shapiro0", "ishapiro")


def _island(nx, ny, dx):
    xy, en, ob = cases.channel(nx, ny, dx, 0.2)
    cx, cy = xy[:, 0].mean(), xy[:, 1].mean()
    r2 = ((xy[:, 0] - cx) ** 2 + (xy[:, 1] - cy) ** 2) / (4.0 * dx) ** 2
    depth = 12.0 - 16.0 * np.exp(-r2)
    return xy, en, ob, depth


def case_inputs(name):
    """Mesh, initial state and step parameters of one fixture case."""
    if name == "island-20x14":
        nx, ny, dx, nvrt = 20, 14, 400.0, 6
        xy, en, ob, depth = _island(nx, ny, dx)
        kb = np.clip((6 - depth).astype(np.int32), 0, nvrt - 2)
        p = L.default_step_params(dt=150.0, theta=0.6, dtb_min=30.0, dtb_max=30.0)
    elif name == "lognormal-24x16":
        nx, ny, dx, nvrt = 24, 16, 1000.0, 5
        xy, en, ob = cases.channel(nx, ny, dx, 0.2)
        depth = cases.bathymetry(nx, ny, 6.0, 5.0, 400.0)
        kb = np.zeros(xy.shape[0], np.int32)
        p = L.default_step_params(dt=300.0, theta=0.55, dtb_min=25.0, dtb_max=60.0)
    else:
        raise KeyError(name)
    sigma = cases.sigma_levels(nvrt)
    lx, ly = (nx - 1) * dx, (ny - 1) * dx
    sn_xy = None
    om = O.OracleMesh(xy, en, depth, ob, sigma, kb=kb)
    sn = om.topology(L.T_ISIDENODE)
    sn_xy = 0.5 * (xy[sn[:, 0]] + xy[sn[:, 1]])
    eta = 0.1 * np.sin(2.0 * np.pi * xy[:, 0] / lx) + cases.normal(cases.SEED_ETA, xy.shape[0], 0.0, 1e-3)
    uvw = cases.velocity(xy, sigma, lx, ly, dx, 3)
    su2 = cases.velocity(sn_xy, sigma, lx, ly, dx, 2)
    params = np.array([float(getattr(p, k)) for k in PARAM_KEYS])
    return dict(xy=xy, elnode=en, open_bnd=ob, depth=depth, kb=kb, sigma=sigma, eta=eta, uvw=uvw,
                su2=su2, params=params), p


def params_from(arr):
    kw = {k: (int(v) if k in ("ihhat", "iettype", "ishapiro") else float(v)) for k, v in zip(PARAM_KEYS, arr)}
    return L.default_step_params(**kw)


def run_oracle(inp, p, steps=STEPS, rtol=1e-12, maxiter=2000):
    """The oracle's outputs of `steps` consecutive steps, keyed '<field>_<step>'."""
    om = O.OracleMesh(inp["xy"], inp["elnode"], inp["depth"], inp["open_bnd"], inp["sigma"], kb=inp["kb"])
    om.set(L.F_ETA2, inp["eta"])
    om.set(L.F_NODE_UVW, inp["uvw"])
    om.set(L.F_SIDE_U2, inp["su2"])
    om.levels(p)
    ip, ix = om.topology(L.T_CSR_INDPTR), om.topology(L.T_CSR_INDICES)
    hist = [np.zeros(om.n)] * 3
    out = {}
    for s in range(steps):
        om.backtrack(p)
        om.rhs(p)
        val, b = om.assemble(p)
        x, it, _, _ = O.pcg(ip, ix, val, b, hist=hist, rtol=rtol, maxiter=maxiter)
        out.update({f"foot_{s}": om.get(L.F_FOOT), f"bstatus_{s}": om.get(L.F_BSTATUS),
                    f"b_{s}": b, f"x_{s}": x, f"iters_{s}": np.int32(it)})
        if s == 0:  # the bulkier per-level and matrix outputs of the first step only
            out.update({"btu_0": om.get(L.F_BTU), "hhat_0": om.get(L.F_SIDE_HHAT), "val_0": val})
        hist = [x, hist[0], hist[1]]
        om.set(L.F_ETA1, om.get(L.F_ETA2))
        om.set(L.F_ETA2, x)
        om.momentum(p)
        om.vertical_velocity()
        out.update({f"su2_{s}": om.get(L.F_SIDE_U2), f"ew_{s}": om.get(L.F_ELEM_W)})
        om.levels(p)
        om.nodal_velocity(p)
        out[f"uvw_{s}"] = om.get(L.F_NODE_UVW)
    out["zc_end"] = om.get(L.F_ZC)
    return out


CASES = ("island-20x14", "lognormal-24x16")


def fixture_path(name):
    return os.path.join(HERE, f"step_{name}.npz")


def main():
    for name in CASES:
        inp, p = case_inputs(name)
        out = run_oracle(inp, p)
        np.savez_compressed(fixture_path(name), **inp, **out)
        bits = np.bitwise_or.reduce(out["bstatus_0"].reshape(-1))
        print(f"{fixture_path(name)}: n={inp['xy'].shape[0]} iters={[int(out[f'iters_{s}']) for s in range(STEPS)]} "
              f"status bits={bits} untracked={(out['foot_0'] < 0).sum()}")


if __name__ == "__main__":
    main()
