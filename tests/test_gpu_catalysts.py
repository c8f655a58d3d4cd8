"""The catalyst chemistry on the GPU (k_react: hash-priority reactions executed
in rounds of non-overlapping catalysts; catalyst decay; the feed cap by radix
select): one check from the reference GPU port's checkpoints
(tests/golden/catalysts.npz) against the port's own result, the same check
against the oracle, a graph run equal to eager launches, an ensemble equal to
single runs, and the invariants of a longer run (chains whole, momentum)."""
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, mkparams, params_mod

pytestmark = pytest.mark.gpu
pm = params_mod()

EVK = {"dlip": 3, "dN": 4, "feed": 5, "syn": 6, "rep": 7, "mut": 8, "dK": 9}
CONFIGS = ["linear", "capped", "pi", "synN2", "tries1", "bola", "yvent"]
STATE = ("x", "y", "vx", "vy", "fx", "fy", "type", "mpos", "nxt")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G():
    return np.load(os.path.join(GOLDEN, "catalysts.npz"))


def _load(G, name):
    ov = json.loads(str(G[f"{name}_overrides"]))
    p = pm.load(overrides=ov).check_engine()
    st = {k: G[f"{name}_in_{k}"].copy() for k in STATE}
    st["bond"] = np.full(p.N, -1, np.int32)
    return p, st


@pytest.mark.parametrize("name", CONFIGS)
def test_one_check_vs_the_gpu_port(eng, G, name):
    say(f"case: catalyst check '{name}' vs the GPU port's checkpoint")
    p, st_in = _load(G, name)
    e = eng.Engine(p)
    e.set_state(st_in, step=9)
    e.step(1, graph=False)
    st = e.get_state()
    for k in ("type", "mpos", "nxt"):
        assert np.array_equal(st[k], G[f"{name}_out_{k}"]), k
    L = p.L
    for k in ("x", "y"):
        d = np.abs((st[k] - G[f"{name}_out_{k}"] + L / 2) % L - L / 2)
        assert d.max() < 1e-13, (k, d.max())
    for k in ("vx", "vy"):
        assert np.array_equal(st[k], G[f"{name}_out_{k}"]), k
    rows = [ln.split(",") for ln in str(G[f"{name}_events"]).strip().splitlines()[1:]]
    ev = e.events()
    var = {"a": 0, "b": 1}
    for kind in ("syn", "rep", "mut", "dK", "dlip", "dN"):
        ref = sorted((int(r[2]), var.get(r[5], -1)) for r in rows if r[1] == kind)
        sel = ev[ev["kind"] == EVK[kind]]
        got = sorted((int(x["gid_a"]), int(x["gid_b"]) if kind in ("syn", "rep", "mut", "dK") else -1) for x in sel)
        assert got == ref, kind
    assert int((ev["kind"] == EVK["feed"]).sum()) == sum(int(r[2]) for r in rows if r[1] == "feed")
    e.close()


def _catalyst_params(L=24.0, seed=3, **kw):
    base = dict(f_lip=0.1, c_N=0.25, f_Ka=0.03, f_Kb=0.02, chem=1, n_chem=5, p_syn_a=0.2, p_syn_b=0.3,
                p_rep_a=0.1, p_rep_b=0.1, mu=0.3, p_kd=0.01, p_ld=0.005, p_nd=0.005, feed="uniform", p_feed=0.02,
                feed_cap=0.2, seed=seed)
    base.update(kw)
    return mkparams(L, **base).check_engine()


def test_check_vs_oracle_from_the_same_state(oracle, eng):
    """Both run 4 steps from the same soup, then the device takes the
    oracle's state and both run the step with the check: the same reactions,
    decay and capped feed, positions to the last bits."""
    p = _catalyst_params(feed_cap=0.26, p_feed=0.2)  # the cap binds (66 of about 190 candidates kept)
    say(f"case: catalyst check vs oracle, N={p.N}")
    cp = p.to_c()
    o = oracle.System(cp, p.N)
    o.start(oracle.soup(cp, p.N))
    o.step(4)
    o.events()
    st = o.get_state()
    e = eng.Engine(p)
    e.set_state(st, step=4)
    o = oracle.System(cp, p.N)
    o.set_state(st, 4)
    e.step(1, graph=False)
    o.step(1)
    a, b = e.get_state(), o.get_state()
    for k in ("type", "mpos", "nxt"):
        assert np.array_equal(a[k], b[k]), k
    L = p.L
    for k in ("x", "y"):
        d = np.abs((a[k] - b[k] + L / 2) % L - L / 2)
        assert d.max() < 1e-13, (k, d.max())
    for k in ("vx", "vy"):
        assert np.array_equal(a[k], b[k]), k
    ed, eo = e.events(), o.events()
    kinds = np.bincount(eo["kind"], minlength=10)
    say(f"events by kind (dlip, dN, feed, syn, rep, mut, dK): {kinds[3:].tolist()}")
    for k in ("step", "kind", "gid_a", "gid_b", "x", "y"):
        assert np.array_equal(ed[k], eo[k]), k
    assert kinds[6] > 0 and kinds[7] + kinds[8] > 0 and kinds[5] > 0
    assert int((b["type"] == 3).sum()) == int(np.floor(0.26 * p.N + 1e-9))
    e.close()


def test_graph_equals_eager_and_ensemble_equals_single(eng):
    """20 steps (four checks) as graph replays = eager launches, bitwise; two
    replicas in one engine = two single-replica engines, bitwise."""
    ps = [_catalyst_params(seed=3), _catalyst_params(L=20.0, seed=4, lip_arch="pi")]
    say("case: catalyst engine graph vs eager, ensemble vs single")
    outs = []
    for graph in (True, False):
        e = eng.Engine(ps[0])
        e.start()
        e.step(20, graph=graph)
        outs.append((e.get_state(), e.events()))
        e.close()
    for k in ("x", "y", "vx", "vy", "type", "mpos", "nxt"):
        assert np.array_equal(outs[0][0][k], outs[1][0][k]), k
    assert np.array_equal(outs[0][1], outs[1][1])
    assert int((outs[0][1]["kind"] == EVK["syn"]).sum()) > 0
    singles = []
    for p in ps:
        e = eng.Engine(p)
        e.start()
        e.step(20)
        singles.append((e.get_state(), e.events()))
        e.close()
    e = eng.Engine(ps)
    e.start()
    e.step(20)
    for r in range(2):
        s = e.get_state(r)
        for k in ("x", "y", "vx", "vy", "type", "mpos", "nxt"):
            assert np.array_equal(s[k], singles[r][0][k]), (r, k)
        assert np.array_equal(e.events(r), singles[r][1])
    e.close()


def test_long_run_invariants(eng):
    """600 steps with catalysts at N = 27,648: every lipid chain whole (one
    head, mlen beads in template order), momentum conserved (synthesis keeps
    the mean velocity of the beads it uses), the N count at or below the cap
    after the last check, and all reaction kinds seen."""
    p = _catalyst_params(L=96.0, seed=21, feed_cap=0.26, p_feed=0.05)
    say(f"case: catalyst long run, N={p.N}")
    e = eng.Engine(p)
    e.start()
    e.step(600)
    st = e.get_state()
    ev = e.events()
    ml = p.mlen
    heads = np.flatnonzero(st["mpos"] == 0)
    inlip = np.zeros(p.N, bool)
    for h in heads:
        g, n = h, 0
        while g >= 0 and n <= ml:
            assert st["mpos"][g] == n
            inlip[g] = True
            g = st["nxt"][g]
            n += 1
        assert n == ml
    assert np.array_equal(inlip, st["mpos"] >= 0)
    P = np.array([st["vx"].sum(), st["vy"].sum()])
    s = e.stats()
    say(f"lipids {len(heads)}, momentum {P}, T {s['T_kin']:.4f}, events {np.bincount(ev['kind'], minlength=10)[3:]}")
    assert np.abs(P).max() < 1e-6
    # the last check (step 600) ends with the capped feed
    assert int((st["type"] == 3).sum()) <= int(np.floor(0.26 * p.N + 1e-9))
    for k in ("syn", "rep", "mut", "dK", "dN", "dlip", "feed"):
        assert int((ev["kind"] == EVK[k]).sum()) > 0, k
    e.close()


def test_one_check_on_a_loaded_state_vs_oracle(oracle, eng):
    """oc_chem: one chemistry check (reactions, decay, capped feed) on a loaded
    state, without a step, against the oracle's; the device sorts again
    after the syntheses moved beads."""
    p = _catalyst_params(seed=5, feed_cap=0.26, p_feed=0.2)
    say(f"case: catalyst check on a loaded state, N={p.N}")
    cp = p.to_c()
    o = oracle.System(cp, p.N)
    o.start(oracle.soup(cp, p.N))
    o.step(3)
    o.events()
    st = o.get_state()
    e = eng.Engine(p)
    e.set_state(st, step=3)
    o = oracle.System(cp, p.N)
    o.set_state(st, 3)
    e.chem(5)
    o.chem(5)
    a, b = e.get_state(), o.get_state()
    for k in ("type", "mpos", "nxt"):
        assert np.array_equal(a[k], b[k]), k
    L = p.L
    for k in ("x", "y"):
        d = np.abs((a[k] - b[k] + L / 2) % L - L / 2)
        assert d.max() < 1e-13, (k, d.max())
    ed, eo = e.events(), o.events()
    assert int((eo["kind"] == EVK["syn"]).sum()) > 0
    for k in ("step", "kind", "gid_a", "gid_b", "x", "y"):
        assert np.array_equal(ed[k], eo[k]), k
    # the reloaded slot order is a valid sort: the next steps run
    e.step(4)
    assert np.isfinite(e.stats()["T_kin"])
    e.close()
