"""Chemistry statistics of the HIP step against the reference CPU engine and
against the stated probabilities.

Catalyst chemistry (the reference GPU port's hash-decided form, which the
reference compares with its CPU engine's serial chemistry by rates,
cel/engine_gpu/tests/chemstat.sh + compare.py:227-282): the two chemstat
modes (water L = 48 with catalysts and nutrients, 2000 checks; synthesis
only; replication + mutation + catalyst decay) against the CPU engine's own
run (tests/golden/chemrates.json, written by make_golden.py from
oracle/_ref): event totals and synthesis / replication per catalyst-check
with the reference's bar |z| < 4.5 (or within 15 %), the mutation fraction
against mu = 0.3.                                                    ~20 s

Precursor chemistry (own definition, no reference counterpart): per check,
a P next to a lipid tail converts with p_conv and a chemistry bond breaks
with p_break (binomial counts over the eligible beads / bonds), and two free
T that are each other's only candidate bond with p_form.          ~10 s
"""
import importlib
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, mkparams

pytestmark = pytest.mark.gpu

KINDS = {3: "dlip", 4: "dN", 5: "feed", 6: "syn", 7: "rep", 8: "mut", 9: "dK", 0: "conv", 1: "form", 2: "break"}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mode", ["syn", "rep"])
def test_catalyst_rates_vs_reference_cpu_engine(eng, mode):
    pm = importlib.import_module("welcome-youtube_amd.params")
    fx = json.load(open(os.path.join(GOLDEN, "chemrates.json")))
    ref = fx["modes"][mode]
    ov = dict(fx["base"], **ref["overrides"])
    p = pm.load(overrides=ov).check_engine()
    steps = int(ov["steps"])
    say(f"case: catalyst chemistry, chemstat mode {mode}, N={p.N}, {steps} steps")
    e = eng.Engine(p)
    e.start()
    cnt, nk = {}, []
    for _ in range(steps // 100):
        e.step(100)
        st = e.stats()
        nk.append(st["count"][4] + st["count"][5])
        for k in e.events()["kind"]:
            cnt[KINDS[int(k)]] = cnt.get(KINDS[int(k)], 0) + 1
    e.close()
    checks = steps // p.n_chem
    gk = float(np.mean(nk)) * checks
    rk, rc = ref["catalyst_checks"], ref["events"]
    say(f"  device: {checks} checks, {gk:.0f} catalyst-checks; events {sorted(cnt.items())}")
    say(f"  CPU engine: {ref['checks']} checks, {rk:.0f} catalyst-checks; events {sorted(rc.items())}")
    for k in ("syn", "rep", "mut", "dlip", "dK", "dN"):
        cr, cg = rc.get(k, 0), cnt.get(k, 0)
        if cr + cg == 0:
            continue
        z = (cg - cr) / np.sqrt(max(cr + cg, 1))
        ok = abs(z) < 4.5 or abs(cg - cr) <= 0.15 * max(cr, cg)
        msg = f"  {k:5s} device {cg} vs CPU engine {cr} (z = {z:+.2f})"
        if k in ("syn", "rep"):
            fr, fg = cr / rk, cg / gk
            zr = (fg - fr) / np.sqrt(max(cr, 1) / rk ** 2 + max(cg, 1) / gk ** 2)
            msg += f"; per catalyst-check {fg:.3e} vs {fr:.3e} (z = {zr:+.2f})"
            ok = ok and (abs(zr) < 4.5 or abs(fg - fr) <= 0.15 * max(fr, fg))
        say(msg)
        assert ok, k
    if mode == "rep":
        rep, mut = cnt.get("rep", 0), cnt.get("mut", 0)
        fr = mut / max(rep + mut, 1)
        se = np.sqrt(0.3 * 0.7 / max(rep + mut, 1))
        say(f"  mutation fraction {fr:.3f} +- {se:.3f} (mu = 0.3; CPU engine "
            f"{rc['mut'] / (rc['rep'] + rc['mut']):.3f})")
        assert abs(fr - 0.3) < 4 * se


PRECURSOR = dict(pchem=1, n_chem=1, p_conv=0.05, p_form=0.2, p_break=0.05)


def precursor_rate_counts(get_state, step1, events, p, nchecks):
    """(eligible, hits) per decision over nchecks checks of a system driven by
    get_state(), step1() (one step, returns the check's step) and events():
    P with a lipid tail within r_conv and their conversions, chemistry bonds
    and their breaks, free-T pairs that are each other's only candidate
    within r_form and their bonds."""
    cKDTree = pytest.importorskip("scipy.spatial").cKDTree
    L = p.L
    elig = {"conv": 0, "break": 0, "form": 0}
    hits = {"conv": 0, "break": 0, "form": 0}
    for _ in range(nchecks):
        s = get_state()  # types, chains and bonds at the start of the check
        step = step1()
        ev = events()
        assert np.all(ev["step"] == step)
        a = get_state()  # the check runs on the integrated positions (then the forces)
        pos = np.c_[a["x"], a["y"]] % L
        tree = cKDTree(pos, boxsize=L)
        P = np.flatnonzero(s["type"] == 3)
        tails = (s["type"] == 2) & (s["mpos"] >= 0)
        nb = tree.query_ball_point(pos[P], p.r_conv)
        has_tail = np.array([bool(np.any(tails[np.array(n, int)])) for n in nb], bool)
        conv = set(ev["gid_a"][ev["kind"] == 0].tolist())
        elig["conv"] += int(has_tail.sum())
        hits["conv"] += len(conv)
        assert conv <= set(P[has_tail].tolist())  # only eligible P convert
        bonded = np.flatnonzero((s["mpos"] < 0) & (s["bond"] >= 0) & (np.arange(p.N) < s["bond"]))
        elig["break"] += len(bonded)
        hits["break"] += int((ev["kind"] == 2).sum())
        fi = np.flatnonzero((s["type"] == 2) & (s["mpos"] < 0) & (s["bond"] < 0))
        pr = cKDTree(pos[fi], boxsize=L).query_pairs(p.r_form, output_type="ndarray")
        deg = np.bincount(pr.ravel(), minlength=len(fi))
        solo = pr[(deg[pr[:, 0]] == 1) & (deg[pr[:, 1]] == 1)]
        formed = {(int(x), int(y)) for x, y in zip(ev["gid_a"][ev["kind"] == 1], ev["gid_b"][ev["kind"] == 1])}
        for x, y in solo:
            gx, gy = int(fi[x]), int(fi[y])
            elig["form"] += 1
            hits["form"] += (min(gx, gy), max(gx, gy)) in formed
    return elig, hits


def check_binomial(elig, hits, nmin=200):
    for k, pk in (("conv", PRECURSOR["p_conv"]), ("break", PRECURSOR["p_break"]), ("form", PRECURSOR["p_form"])):
        n, x = elig[k], hits[k]
        z = (x - n * pk) / np.sqrt(n * pk * (1 - pk))
        say(f"  {k:5s}: {x} of {n} eligible, rate {x / max(n, 1):.4f} vs p = {pk} (z = {z:+.2f})")
        assert n > nmin and abs(z) < 4.0, k


def test_precursor_chemistry_rates(eng):
    """Per check from the state at its start: conversions of the P with a
    lipid tail within r_conv ~ Binomial(n, p_conv); breaks of the chemistry
    bonds ~ Binomial(n, p_break); bonds of free-T pairs that are each other's
    only candidate ~ Binomial(n, p_form) (the mutual-proposal rule lowers the
    rate of a pair whose members have other candidates: each proposes to its
    smallest draw below p_form)."""
    p = mkparams(40.0, nsub=4, f_lip=0.3, c_N=0.25, seed=33, **PRECURSOR)
    say(f"case: precursor chemistry rates, N={p.N}")
    e = eng.Engine(p)
    e.start()
    e.step(20)
    e.events()

    def step1():
        e.step(1, graph=False)
        return e.step_count()

    elig, hits = precursor_rate_counts(e.get_state, step1, e.events, p, 60)
    e.close()
    check_binomial(elig, hits)
