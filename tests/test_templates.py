"""Branched lipid templates (pi, y, bola; mol_build, cel/engine/oercel.c:479-576)
against vectors written by the reference engine itself (tests/golden/templates.npz,
make_golden.py): the oracle's soups and bonded forces bit for bit, its 20-step
run of tied pi lipids against the reference CLI, the library's host-side soup,
and the Python template (bond list of the snapshots) against the template's
own geometry.  CPU only."""
import ctypes as C
import importlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, params_mod

pm = params_mod()
G = np.load(os.path.join(GOLDEN, "templates.npz"))

TEMPLATES = {
    "pi": dict(lip_arch="pi", lip_nT=2, lip_tie=1, r0_head=0.7, k_head=5, cos_head=0.3, k_angle=10),
    "pi_membranes": dict(lip_arch="pi", lip_nT=1, lip_tie=1, r0_head=0.7, k_bond=300),
    "y": dict(lip_arch="y", lip_nT=2, lip_tie=1, k_head=4, cos_head=0.2, k_angle=8),
    "bola": dict(lip_arch="bola", lip_nT=3, lip_tie=2, r0_head=0.6, k_head=3, cos_head=-0.1, k_angle=6),
}


def mk(name):
    p = pm.Params()
    d = {"L": "14.0", "nsub": "4", "f_lip": "0.3", "c_N": "0.1", "f_Ka": "0", "seed": "3"}
    d.update({k: str(v) for k, v in TEMPLATES[name].items()})
    p.update(d)
    return p.finalize()


@pytest.mark.parametrize("name", list(TEMPLATES))
def test_template_soup_bit_exact(oracle, name):
    p = mk(name)
    s = oracle.soup(p.to_c(), p.N)
    for k in ("x", "y", "vx", "vy", "type", "mpos", "nxt"):
        assert np.array_equal(s[k], G[f"{name}_{k}"]), k


@pytest.mark.parametrize("name", list(TEMPLATES))
def test_library_host_soup_bit_exact(name):
    eng = importlib.import_module("welcome-youtube_amd.engine")
    p = mk(name)
    s = eng.soup_host(p)
    for k in ("x", "y", "vx", "vy", "type", "mpos", "nxt"):
        assert np.array_equal(s[k], G[f"{name}_{k}"]), k


@pytest.mark.parametrize("name", list(TEMPLATES))
def test_template_bonded_bit_exact(oracle, name):
    p = mk(name)
    st = {k: G[f"{name}_{k}"] for k in ("x", "y", "vx", "vy", "type", "mpos", "nxt")}
    st["x"], st["y"] = G[f"{name}_bx"], G[f"{name}_by"]
    fx, fy, mb, eb = oracle.bonded_gid(p.to_c(), st)
    assert np.array_equal(fx, G[f"{name}_bfx"]) and np.array_equal(fy, G[f"{name}_bfy"])
    assert mb == float(G[f"{name}_maxbond"])
    # internal forces of every molecule sum to zero
    assert abs(fx.sum()) < 1e-10 * np.abs(fx).sum()


@pytest.mark.parametrize("name", list(TEMPLATES))
def test_python_template_bonds_at_rest_in_the_soup(name):
    """The soup places every lipid as its template, so each template bond
    (Python mol_template, the snapshots' bond list) has its rest length there,
    and the bead count and heads match the reference soup."""
    io = importlib.import_module("welcome-youtube_amd.io")
    p = mk(name)
    M = pm.mol_template(p)
    s = {k: G[f"{name}_{k}"] for k in ("x", "y", "type", "mpos", "nxt")}
    assert M.mlen == int(s["mpos"].max()) + 1
    assert np.array_equal(s["type"][s["mpos"] >= 0][:M.mlen], np.array([1] * M.nH + [2] * (M.mlen - M.nH)))
    bonds = io.lipid_bonds(s, M.mlen, [(b[0], b[1]) for b in M.bonds])
    assert len(bonds) == int((s["mpos"] == 0).sum()) * len(M.bonds)
    L = p.L
    heads = np.flatnonzero(s["mpos"] == 0)
    for h in heads[:10]:
        chain = [int(h)]
        while len(chain) < M.mlen:
            chain.append(int(s["nxt"][chain[-1]]))
        for i, j, r0 in M.bonds:
            dx = (s["x"][chain[j]] - s["x"][chain[i]] + L / 2) % L - L / 2
            dy = (s["y"][chain[j]] - s["y"][chain[i]] + L / 2) % L - L / 2
            assert abs(np.hypot(dx, dy) - r0) < 1e-12, (i, j)


def test_template_trajectory_vs_reference_cli(oracle):
    """20 steps of the reference CLI with the tied pi H2T2 lipid of its
    membranes.ini (k_bond 300, a_W_T 200, warm-up ramp): the oracle's run stays
    within FP64 round-off of it, slot order identical."""
    ov = json.loads(str(G["traj_overrides"]))
    p = pm.load(overrides=ov)
    cp = p.to_c()
    sysm = oracle.System(cp, p.N)
    sysm.start(oracle.soup(cp, p.N), theta32=0)
    sysm.step(int(G["traj_step"]), theta32=0)
    s = sysm.get_state()
    L = p.L
    dx = np.abs((s["x"] - G["traj_x"] + L / 2) % L - L / 2).max()
    dv = max(np.abs(s["vx"] - G["traj_vx"]).max(), np.abs(s["vy"] - G["traj_vy"]).max())
    df = max(np.abs(s["fx"] - G["traj_fx"]).max(), np.abs(s["fy"] - G["traj_fy"]).max())
    assert dx < 1e-10 and dv < 1e-8 and df < 1e-9 * np.abs(G["traj_fx"]).max(), (dx, dv, df)
    assert np.array_equal(s["nxt"], G["traj_nxt"])
    _, gid = sysm.cells()
    assert np.array_equal(gid, G["traj_slot_gid"])


def test_template_parameters():
    with pytest.raises(pm.ParamError):
        pm.load(None, ["lip_arch=star"])
    with pytest.raises(pm.ParamError):
        pm.load(None, ["lip_arch=bola", "lip_nT=5"])  # 13 beads
    p = pm.load(None, ["lip_arch=y", "lip_nT=3"])
    assert p.mlen == 7 and p.nH == 1 and p.to_c().lip_arch == 2
