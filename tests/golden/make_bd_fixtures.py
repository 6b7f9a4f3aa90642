"""Write the golden fixtures of the free-boundary step (bd_sim) from the
reference's own compiled code (oracle/_ref/libstzref.so, built by
``make -C oracle ref`` where the reference tree is present):

    python tests/golden/make_bd_fixtures.py

"bd_plate96": a plate with a hole (stz_amd.problems.plate_hole_phi) on 96^2
cells with the material of bd_test.cc.  It holds the reference's narrow band
built from the analytic level set (levelset::build_band, levelset.cc:209-345),
then, from one injected state (seeded stresses and chi inside the body,
a smooth velocity, a body force), the reference's outputs of each piece of
the step: the interface speed (levelset.hh:159-205), the extrapolation
(bd_sim::post_process), the operator rows (qs_multi_bd::init, free-surface
rows only -- the others are closed form), the projection
(bd_sim::projection_step), the advection (bd_sim::advection_step), and three
whole quasi-static steps (bd_sim::step_forward_quasistatic).
"bd_bar{1,2,3}": bd_sim::init_bar (bd_sim.cc:63-123) on small grids.
"""
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.pyoracle import RefBd  # noqa: E402
from stz_amd.problems import bd_params, plate_hole_phi  # noqa: E402

DT = 0.5


def injected_state(N, cstat, seed):
    """seeded fields on the 96^2 plate: chi 600 K +- 20 K, stresses +-0.01
    inside the body; u, v a smooth field; X, Y the node coordinates; a body
    force along x, odd in x, on the plate's nodes"""
    rng = np.random.default_rng(seed)
    ins = cstat < 4
    dx = 2.0 / N
    XN, YN = np.meshgrid(-1 + np.arange(N + 1) * dx, -1 + np.arange(N + 1) * dx)
    cells = {"chi": np.where(ins, (600 + 20 * rng.standard_normal(cstat.shape)) / 21000, 0.0)}
    for k in ("p", "q", "s", "tau"):
        cells[k] = np.where(ins, 0.01 * rng.standard_normal(cstat.shape), 0.0)
    nodes = {"u": 1e-5 * np.sin(np.pi * XN) * np.cos(YN), "v": 1e-5 * np.cos(np.pi * YN) * XN, "X": XN, "Y": YN}
    F = np.zeros(XN.shape + (2,))
    F[..., 0] = 0.02 * np.sin(np.pi * XN / 2) * (np.abs(YN) < 0.9)
    return cells, nodes, F


def load(r, cells, nodes, F):
    for k, v in cells.items():
        r.set_cell(k, v)
    r.set_cell("ddev", np.zeros_like(cells["chi"]))
    for k, v in nodes.items():
        r.set_node(k, v)
    r.set_body_force(F)


def plate_case(name, N, seed, tmp):
    P = bd_params(N)
    phi0 = plate_hole_phi(N)
    out = {"params": json.dumps(P), "seed": seed, "dt": DT, "phi0": phi0}
    r = RefBd(P, tmp)
    r.set_level_set(phi0)
    c, cc = r.status()
    out["build_phi"] = r.get_cell("phi")
    out["build_status"] = c
    out["build_nstatus"] = cc
    cells, nodes, F = injected_state(N, c, seed)
    for k, v in cells.items():
        out["in_" + k] = v
    for k, v in nodes.items():
        out["in_" + k] = v
    out["in_force"] = F

    def fresh():
        rr = RefBd(P, tmp)
        rr.set_level_set(phi0)
        load(rr, cells, nodes, F)
        return rr

    # extrapolation and the pieces that read the extrapolated state
    r = fresh()
    r.post_process(True)
    for k in ("p", "q", "s", "tau", "chi"):
        out["post_" + k] = r.get_cell(k)
    for k in ("u", "v", "X", "Y"):
        out["post_" + k] = r.get_node(k)
    out["vp"] = r.interface_velocity()
    d, mats, b = r.operator(DT)
    free = np.argwhere(d.ravel() == 0).ravel()
    out["op_rowtype"] = d
    out["op_src"] = b
    out["op_free_index"] = free
    out["op_free_blocks"] = mats.reshape(-1, 9, 2, 2)[free]
    r.projection_step(DT)
    out["proj_vel"] = r.get_vel()
    for k in ("u", "v"):
        out["proj_" + k] = r.get_node(k)
    for k in ("p", "q", "s", "tau"):
        out["proj_" + k] = r.get_cell(k)
    # advection from the extrapolated state
    r = fresh()
    r.post_process(True)
    r.advection_step(DT)
    c2, cc2 = r.status()
    out["adv_status"], out["adv_nstatus"], out["adv_phi"] = c2, cc2, r.get_cell("phi")
    for k in ("p", "q", "s", "tau", "chi", "ddev"):
        out["adv_" + k] = r.get_cell(k)
    for k in ("X", "Y"):
        out["adv_" + k] = r.get_node(k)
    # three whole steps
    r = fresh()
    r.post_process(True)
    r.step_forward_quasistatic(DT, 3)
    c3, _ = r.status()
    out["steps"] = 3
    out["steps_status"] = c3
    for k in ("p", "q", "s", "tau", "chi"):
        out["steps_" + k] = r.get_cell(k)
    for k in ("u", "v", "X", "Y"):
        out["steps_" + k] = r.get_node(k)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "free rows", free.size, "solution max |u| %.3e" % np.abs(out["proj_u"]).max())


def bar_case(name, ca, m, n, box, tmp):
    P = bd_params(m, n, box=box)
    r = RefBd(P, tmp)
    r.init_bar(ca)
    c, cc = r.status()
    out = {"params": json.dumps(P), "ca": ca, "status": c, "nstatus": cc, "phi": r.get_cell("phi")}
    for k in ("p", "q", "s", "tau", "chi"):
        out[k] = r.get_cell(k)
    for k in ("u", "v", "X", "Y"):
        out[k] = r.get_node(k)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "cells inside", int((c < 4).sum()))


def main():
    tmp = tempfile.mkdtemp(prefix="bdref")
    plate_case("bd_plate96", 96, 11, tmp)
    bar_case("bd_bar1", 1, 96, 48, (-2.0, 2.0, -1.0, 1.0), tmp)
    bar_case("bd_bar2", 2, 32, 64, (-0.5, 0.5, -1.0, 1.0), tmp)
    bar_case("bd_bar3", 3, 64, 64, (-1.6, 1.6, -1.6, 1.6), tmp)


if __name__ == "__main__":
    main()
