"""The oracle restatement against vectors written by the reference engine
itself (tests/golden/make_golden.py, from cel/engine/oercel.c built in
oracle/_ref).  CPU only."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, params_mod

pm = params_mod()

SOUPS = {"forcetest": (12.5, 4, dict(f_lip=0.15, f_Ka=0.02, f_Kb=0.02, c_N=0.1, seed=99)),
         "cfg1like": (20.0, 4, dict(f_lip=0.1, lip_nT=2, c_N=0.0, f_Ka=0.0, seed=1)),
         "precursor": (16.0, 2, dict(f_lip=0.3, lip_nT=2, c_N=0.2, f_Ka=0.0, seed=7))}


def mk(L, nsub, **kw):
    p = pm.Params()
    d = {"L": repr(L), "nsub": str(nsub)}
    d.update({k: str(v) for k, v in kw.items()})
    p.update(d)
    return p.finalize()


def load(name):
    return np.load(os.path.join(GOLDEN, name))


def test_hashes_bit_exact(oracle):
    g = load("hashes.npz")
    L = oracle.lib()
    assert all(L.or_mix13(int(z)) == int(m) for z, m in zip(g["z"], g["mix13"]))
    assert all(L.or_step_seed(int(s), int(t)) == int(v) for s, t, v in zip(g["seeds"], g["steps"], g["step_seed"]))
    th = [L.or_pair_theta(int(g["z"][k % 64]), int(g["gi"][k]), int(g["gj"][k])) for k in range(64)]
    assert np.array_equal(np.array(th), g["pair_theta"])
    hk = [L.or_hash_key(int(s), int(t), int(a)) for s, t, a in zip(g["seeds"], g["steps"], g["salts"])]
    assert np.array_equal(np.array(hk, np.uint64), g["hash_key"])
    hu = [L.or_hash_u01(int(g["hash_key"][k % 16]), int(g["gi"][k])) for k in range(64)]
    assert np.array_equal(np.array(hu), g["hash_u01"])


@pytest.mark.parametrize("name", list(SOUPS))
def test_soup_bit_exact(oracle, name):
    L, nsub, kw = SOUPS[name]
    p = mk(L, nsub, **kw)
    g = load("soups.npz")
    s = oracle.soup(p.to_c(), p.N)
    for k in ("x", "y", "vx", "vy", "type", "mpos", "nxt"):
        assert np.array_equal(s[k], g[f"{name}_{k}"]), k


@pytest.mark.parametrize("name", list(SOUPS))
@pytest.mark.parametrize("nsub", [1, 3, 4])
def test_sort_bit_exact(oracle, name, nsub):
    L, _, kw = SOUPS[name]
    cp = mk(L, nsub, **kw).to_c()
    g = load("sorts.npz")
    src, cs = oracle.sort(cp, g[f"{name}_x"], g[f"{name}_y"])
    assert np.array_equal(src, g[f"{name}_n{nsub}_src"])
    assert np.array_equal(cs, g[f"{name}_n{nsub}_cstart"])


@pytest.mark.parametrize("name", list(SOUPS))
def test_pair_forces_vs_reference(oracle, name):
    L, nsub, kw = SOUPS[name]
    p = mk(L, nsub, **kw)
    s = load("soups.npz")
    f = load("forces.npz")
    st = {k: s[f"{name}_{k}"] for k in ("x", "y", "type", "mpos", "nxt")}
    st["vtx"], st["vty"] = f[f"{name}_vtx"], f[f"{name}_vty"]
    out = oracle.pair_forces_gid(p.to_c(), st, 12345, theta32=0)
    ref = np.hypot(f[f"{name}_fx"], f[f"{name}_fy"]).max()
    d = max(np.abs(out["fx"] - f[f"{name}_fx"]).max(), np.abs(out["fy"] - f[f"{name}_fy"]).max())
    assert d < 1e-12 * ref
    # stirring + temperature cycle at t = 2.5 (kT = 1.3), cel/engine/tests/forcetest.c:80-84
    pe = mk(L, nsub, stir_A=0.7, stir_n=2, kT_amp=0.3, kT_period=10.0, **kw)
    pe.t_anchor = 2.5 - 12345 * pe.dt
    out = oracle.pair_forces_gid(pe.to_c(), st, 12345, theta32=0)
    d = max(np.abs(out["fx"] - f[f"{name}_env_fx"]).max(), np.abs(out["fy"] - f[f"{name}_env_fy"]).max())
    assert d < 1e-11 * ref


@pytest.mark.parametrize("name", list(SOUPS))
def test_bonded_forces_vs_reference(oracle, name):
    L, nsub, kw = SOUPS[name]
    p = mk(L, nsub, **kw)
    s = load("soups.npz")
    f = load("forces.npz")
    st = {"x": f[f"{name}_bx"], "y": f[f"{name}_by"], "mpos": s[f"{name}_mpos"], "nxt": s[f"{name}_nxt"]}
    fx, fy, mb, _ = oracle.bonded_gid(p.to_c(), st)
    assert np.array_equal(fx, f[f"{name}_bfx"]) and np.array_equal(fy, f[f"{name}_bfy"])
    assert mb == float(f[f"{name}_maxbond"])


@pytest.mark.parametrize("name", ["lipid", "water"])
def test_trajectory_vs_reference_cli(oracle, name):
    """20 steps of the reference CLI from a fresh soup (warm-up ramp, angle
    terms on): the oracle's run stays within FP64 round-off of it."""
    g = load("traj.npz")
    ov = json.loads(str(g[f"{name}_overrides"]))
    p = pm.load(overrides={k: v for k, v in ov.items()})
    cp = p.to_c()
    sysm = oracle.System(cp, p.N)
    sysm.start(oracle.soup(cp, p.N), theta32=0)
    sysm.step(int(g[f"{name}_step"]), theta32=0)
    s = sysm.get_state()
    L = p.L
    dx = np.abs((s["x"] - g[f"{name}_x"] + L / 2) % L - L / 2).max()
    dy = np.abs((s["y"] - g[f"{name}_y"] + L / 2) % L - L / 2).max()
    dv = max(np.abs(s["vx"] - g[f"{name}_vx"]).max(), np.abs(s["vy"] - g[f"{name}_vy"]).max())
    df = max(np.abs(s["fx"] - g[f"{name}_fx"]).max(), np.abs(s["fy"] - g[f"{name}_fy"]).max())
    fmax = np.abs(g[f"{name}_fx"]).max()
    assert dx < 1e-10 and dy < 1e-10 and dv < 1e-8 and df < 1e-9 * fmax, (dx, dy, dv, df)
    assert np.array_equal(s["type"], g[f"{name}_type"]) and np.array_equal(s["nxt"], g[f"{name}_nxt"])
    _, gid = sysm.cells()
    assert np.array_equal(gid, g[f"{name}_slot_gid"]), "slot order after 20 sorts"


def test_parameter_table_matches_reference():
    with open(os.path.join(GOLDEN, "params.json")) as f:
        t = json.load(f)
    assert [k for k, _, _ in pm.PDEFS] == t["order"]
    for k, _, d in pm.PDEFS:
        assert d == t["defaults"][k], k
    a = pm.default_amatrix()
    for i, n in enumerate(pm.TYPE_NAMES):
        assert a[i] == t["amatrix"][n]
