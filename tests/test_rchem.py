"""The reference chemistry without catalysts (chem = 1, no Ka / Kb): lipid and
nutrient decay and the food supply, restated in the oracle and pinned by the
reference itself -- its feed known answer (1138 conversions) through the
oracle's full step, and a 20-step reference CLI run with decay and vent feed
(tests/golden/rchem.npz): types, chains, positions, slot order and the events
of its events.csv.  CPU only."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, params_mod

pm = params_mod()

EVK = {"dlip": 3, "dN": 4, "feed": 5}


def test_feed_known_answer_through_the_oracle_step(oracle):
    """Uniform feed on water, L = 48 (N = 6,912), seed 7, p_feed 0.003, a
    check every 10 steps, 600 steps, no decay: 1138 conversions
    (cel/engine/tests/validate.sh:133-134, cel/engine/README.md:519-520)."""
    p = pm.load(None, ["L=48", "seed=7", "chem=1", "f_Ka=0", "f_Kb=0", "f_lip=0", "c_N=0", "feed=uniform",
                       "p_feed=0.003", "feed_cap=0", "n_chem=10", "p_nd=0", "p_ld=0", "warmup=0", "dt=0.02"])
    p.check_engine()
    assert p.N == 6912
    cp = p.to_c()
    s = oracle.System(cp, p.N)
    s.start(oracle.soup(cp, p.N))
    s.step(600)
    st = s.get_state()
    ev = s.events()
    assert int((st["type"] == 3).sum()) == 1138
    assert int((ev["kind"] == EVK["feed"]).sum()) == 1138


def test_rchem_run_vs_reference_cli(oracle):
    G = np.load(os.path.join(GOLDEN, "rchem.npz"))
    ov = json.loads(str(G["rc_overrides"]))
    p = pm.load(overrides=ov)
    p.check_engine()
    cp = p.to_c()
    s = oracle.System(cp, p.N)
    s.start(oracle.soup(cp, p.N), theta32=0)
    s.step(int(G["rc_step"]), theta32=0)
    st = s.get_state()
    for k in ("type", "mpos", "nxt"):
        assert np.array_equal(st[k], G[f"rc_{k}"]), k
    L = p.L
    dx = np.abs((st["x"] - G["rc_x"] + L / 2) % L - L / 2).max()
    dv = np.abs(st["vx"] - G["rc_vx"]).max()
    assert dx < 1e-10 and dv < 1e-8, (dx, dv)
    assert np.array_equal(s.cells()[1], G["rc_slot_gid"])
    # events: the reference lists dlip / dN per bead and one feed row per check
    rows = [ln.split(",") for ln in str(G["rc_events"]).strip().splitlines()[1:]]
    ev = s.events()
    assert len(rows) > 20
    for step in sorted({int(r[0]) for r in rows}):
        for kind in ("dlip", "dN"):
            ref = sorted((int(r[2]), float(r[3]), float(r[4])) for r in rows if int(r[0]) == step and r[1] == kind)
            got = sorted((int(e["gid_a"]), e["x"], e["y"]) for e in ev
                         if e["step"] == step and e["kind"] == EVK[kind])
            assert [g for g, _, _ in got] == [g for g, _, _ in ref], (step, kind)
            assert all(abs(a[1] - b[1]) < 6e-5 and abs(a[2] - b[2]) < 6e-5 for a, b in zip(got, ref))
        nfeed = [int(r[2]) for r in rows if int(r[0]) == step and r[1] == "feed"]
        got = int(((ev["step"] == step) & (ev["kind"] == EVK["feed"])).sum())
        assert got == (nfeed[0] if nfeed else 0), step


def test_rchem_parameters():
    p = pm.load(None, ["chem=1", "f_Ka=0", "feed_cap=0", "feed=vent", "vent_shape=hstrip", "vent_w=3"]).check_engine()
    c = p.to_c()
    assert c.rchem == 1 and c.feed_mode == 2 and c.vent_shape == 2 and c.vent_w == 3.0
