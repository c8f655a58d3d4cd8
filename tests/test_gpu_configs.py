"""GPU tests at the sizes of the benchmark configurations (the HIP step through
the C ABI against the oracle and against its own single-replica runs).

  config 2   N = 1,048,576 (L = 591.2067), 50% W / 20% P / 30% H-T-T lipid
             beads, precursor chemistry p_conv 0.005 / p_form 0.01 /
             p_break 0.001 every step: one step from a loaded state (cells,
             slot order, integration, types, bonds and events bit-exact,
             forces within 1e-5), pair forces per bead within 1e-5, then 500
             graph steps with the total momentum below 1e-6.       ~40 s
  config 3   256 replicas, N log-uniform in [2048, 65536], f_lip U[0.05, 0.4],
             a_WT U[60, 100], chemistry on, 20 steps: the smallest, the
             largest and two random replicas bitwise equal to single-replica
             runs.                                                  ~30 s
  config 4   N = 65,536, 30% lipids: the window gradient within 1e-3 of the
             FP64 oracle's over 12 steps and of central differences
             (eps = 1e-3) of the device's own forward over 20 steps,
             checkpoint every 20.                                   ~90 s (oracle)
  config 5a  N = 16,777,216 (L = 2364.8267), config-2 composition and
             chemistry, 20 graph steps: the slot order a permutation, every
             bead inside its cell, total momentum below 1e-6, no device
             error.                                                 ~40 s
  config 5b  32 x 262,144 (L = 295.6033), one step of the batched fit: its
             gradient the replica mean of the per-replica window gradients,
             two sampled replicas' gradients bitwise equal to single-replica
             engines', a rewound window repeating bit for bit.      ~40 s
"""
import importlib
import math
import sys

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
FORCE_ABS = 1e-5  # force units: FP32 resolution of pairs at the cutoff (tests/test_gpu_parity.py)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pm():
    return importlib.import_module("welcome-youtube_amd.params")


def abi():
    return importlib.import_module("welcome-youtube_amd.abi")


def test_config2_step_parity_and_momentum(oracle, eng):
    p = bench.config2_params(pm(), 1048576, 4)
    assert p.N == 1048576
    say("case: config 2, N=1,048,576: 300 steps to a state with free and bonded T")
    e = eng.Engine(p)
    e.start()
    e.step(300)
    st = e.get_state()
    s0 = e.step_count()
    e.close()
    free_t = int(((st["type"] == 2) & (st["mpos"] < 0) & (st["bond"] < 0)).sum())
    bonded_t = int(((st["type"] == 2) & (st["mpos"] < 0) & (st["bond"] >= 0)).sum())
    say(f"  free T {free_t}, bonded T {bonded_t}, P {int((st['type'] == 3).sum())}")
    assert free_t > 1000 and bonded_t > 1000

    say("case: config 2, one step from the loaded state, device vs oracle")
    d = eng.Engine(p)
    d.set_state(st, step=s0)
    o = oracle.System(p.to_c(), p.N)
    o.set_state(st, s0)
    d.step(1, graph=False)
    o.step(1)
    cs_d, g_d = d.cells()
    cs_o, g_o = o.cells()
    assert np.array_equal(cs_d, cs_o), "cstart differs"
    assert np.array_equal(g_d, g_o), "slot order differs"
    a, b = d.get_state(), o.get_state()
    for k in ("x", "y", "vx", "vy", "type", "mpos", "nxt", "bond"):
        assert np.array_equal(a[k], b[k]), f"{k} not bit-exact"
    ev_d, ev_o = d.events(), o.events()
    say(f"  events: device {len(ev_d)}, oracle {len(ev_o)}")
    assert len(ev_d) > 0 and len(ev_d) == len(ev_o)
    for k in ("step", "kind", "gid_a", "gid_b"):
        assert np.array_equal(ev_d[k], ev_o[k]), k
    dF = np.hypot(a["fx"] - b["fx"], a["fy"] - b["fy"])
    glob = dF.max() / np.hypot(b["fx"], b["fy"]).max()
    say(f"  step forces: max|dF|/max|F| = {glob:.3e}")
    assert glob < FORCE_TOL

    say("case: config 2, pair forces per bead on the stepped state (predicted velocity from the oracle's step)")
    sv = dict(b)
    rng = np.random.default_rng(2)
    sv["vtx"] = (b["vx"] + 0.3 * rng.standard_normal(p.N)).astype(np.float32).astype(np.float64)
    sv["vty"] = b["vy"].astype(np.float32).astype(np.float64)
    f = eng.Engine(p)
    ld = {k: sv[k] for k in ("x", "y", "type", "mpos", "nxt", "bond")}
    ld["vx"], ld["vy"] = sv["vtx"], sv["vty"]
    f.set_state(ld, step=0)
    f.forces(s0 + 1, bonded=False)
    out = f.get_state()
    f.close()
    ref = oracle.pair_forces_gid(p.to_c(), sv, s0 + 1, theta32=1, sabs=True)
    dd = np.hypot(out["fx"] - ref["fx"], out["fy"] - ref["fy"])
    rel = (dd / (ref["sabs"] + FORCE_ABS / FORCE_TOL)).max()
    say(f"  pair forces: max |dF| / (sum|f_ij| + {FORCE_ABS / FORCE_TOL:g}) = {rel:.3e}, "
        f"max|dF|/max|F| = {dd.max() / np.hypot(ref['fx'], ref['fy']).max():.3e}")
    assert rel < FORCE_TOL

    say("case: config 2, 500 graph steps, total momentum")
    d.step(500)
    s = d.stats()
    say(f"  step {s['step']}: T_kin {s['T_kin']:.5f}, momentum ({s['px']:.3e}, {s['py']:.3e})")
    assert abs(s["px"]) < 1e-6 and abs(s["py"]) < 1e-6
    assert abs(s["T_kin"] - 1.0) < 0.1
    d.close()


def test_config3_ensemble_256_equals_single_runs(eng):
    ps = bench.ensemble_params(pm(), 256)
    Ns = np.array([q.N for q in ps])
    say(f"case: config 3, 256 replicas, N {Ns.min()}..{Ns.max()}, total {Ns.sum()}")
    assert Ns.min() >= 2048 * 0.99 and Ns.max() <= 65536 * 1.01
    e = eng.Engine(ps)
    e.start()
    e.step(20)
    rng = np.random.default_rng(3)
    pick = [int(Ns.argmin()), int(Ns.argmax())] + [int(r) for r in rng.choice(256, 2, replace=False)]
    got = {r: (e.get_state(r), e.cells(r)[1], e.events(r)) for r in pick}
    e.close()
    for r in pick:
        say(f"  replica {r}: N={ps[r].N}")
        s = eng.Engine(ps[r])
        s.start()
        s.step(20)
        a, g, ev = got[r]
        b = s.get_state()
        for k in b:
            assert np.array_equal(a[k], b[k]), (r, k)
        assert np.array_equal(g, s.cells()[1]), r
        evs = s.events()
        assert len(ev) == len(evs) and all(np.array_equal(ev[k], evs[k]) for k in ("step", "kind", "gid_a", "gid_b"))
        s.close()


def test_config4_adjoint_window_65536(oracle, eng):
    p = bench.adjoint_params(pm())
    assert p.N == 65536
    say("case: config 4, N=65,536: 200 device steps to an equilibrated state")
    e = eng.Engine(p)
    e.start()
    e.step(200)
    st = e.get_state()
    s0 = e.step_count()
    spec = abi().rdf_spec(2, 2, 64, 3.0)
    ospec = oracle.rdf_spec(2, 2, 64, 3.0)

    say("case: config 4, 12-step window gradient vs the FP64 oracle adjoint")
    o = oracle.System(p.to_c(), p.N)
    o.set_state(st, s0)
    target = o.rdf(ospec) * 0.95
    lo, go, _ = o.window_grad(12, ospec, target, theta32=1)
    e.set_state(st, step=s0)
    ld, gd, _ = e.window_grad(12, spec, target, ckpt_every=20)
    rel = np.abs(gd[0] - go) / np.abs(go).max()
    say(f"  device {gd[0]}\n  oracle {go}\n  rel {rel} (max {rel.max():.2e}); loss {ld[0]:.8f} vs {lo:.8f}")
    assert rel.max() < 1e-3
    assert abs(ld[0] - lo) < 1e-5 * lo

    say("case: config 4, 20-step window gradient vs central differences (eps = 1e-3)")
    nsteps, eps = 20, 1e-3
    e.set_state(st, step=s0)
    tgt = e.window_forward(nsteps, spec)[0] * 1.05
    th0 = e.theta()

    def loss(th):
        e.set_state(st, step=s0)
        e.set_theta(th)
        g = e.window_forward(nsteps, spec)[0]
        return ((g - tgt) ** 2).sum()

    e.set_state(st, step=s0)
    e.set_theta(th0)
    _, gad, _ = e.window_grad(nsteps, spec, tgt, ckpt_every=20)
    fd = np.zeros(7)
    for q in range(7):
        tp, tm = th0.copy(), th0.copy()
        tp[q] += eps
        tm[q] -= eps
        fd[q] = (loss(tp) - loss(tm)) / (2 * eps)
    rel = np.abs(gad[0] - fd) / np.abs(fd).max()
    say(f"  adjoint {gad[0]}\n  fd      {fd}\n  rel {rel} (max {rel.max():.2e})")
    e.close()
    assert rel.max() < 1e-3


def test_config5a_16m_invariants(oracle, eng):
    p = bench.config2_params(pm(), 16777216, 4)
    assert p.N == 16777216
    say("case: config 5a, N=16,777,216: 20 graph steps")
    e = eng.Engine(p)
    e.start()
    e.step(20)  # raises on any device error bit
    s = e.stats()
    say(f"  step {s['step']}: T_kin {s['T_kin']:.4f}, momentum ({s['px']:.3e}, {s['py']:.3e})")
    assert abs(s["px"]) < 1e-6 and abs(s["py"]) < 1e-6
    cs, g = e.cells()
    st = e.get_state()
    e.close()
    assert np.array_equal(np.sort(g), np.arange(p.N, dtype=np.uint32)), "slot order is not a permutation"
    assert cs[0] == 0 and cs[-1] == p.N and np.all(np.diff(cs) >= 0)
    keys = oracle.cell_keys(p.to_c(), st["x"][g], st["y"][g])  # cell of every slot's bead
    cell_of_slot = np.repeat(np.arange(len(cs) - 1), np.diff(cs))
    assert np.array_equal(keys, cell_of_slot), "a bead lies outside the cell its slot belongs to"
    for k in ("x", "y"):
        assert np.all(st[k] >= 0) and np.all(st[k] < p.L)


def test_config5b_batched_fit_step(eng):
    fit = importlib.import_module("welcome-youtube_amd.fit")
    R, N, win = 32, 262144, 200
    ps = [bench.adjoint_params(pm(), N=N, seed=1000 + r) for r in range(R)]
    assert ps[0].N == N and abs(ps[0].L - 295.6033) < 1e-3
    say(f"case: config 5b, {R} x {N}: 200 equilibration steps")
    e = eng.Engine(ps)
    e.start()
    e.step(200)
    s0 = e.step_count()
    # every replica reloaded from its state in slot order (as a single-replica
    # engine is: the predicted velocity becomes vh), so both start alike
    st = {}
    for r in range(R):
        s, order = e.get_state(r), e.cells(r)[1]
        e.set_state(s, step=s0, r=r, slot_gid=order)
        if r in (0, 17):
            st[r] = (s, order)
    spec = abi().rdf_spec(2, 2, 64, 3.0)
    target = np.full(64, 0.9)
    target[:10] = 0.0

    say("case: config 5b, one Adam step on the replica-mean window gradient")
    hist = fit.fit_rdf(e, spec, target, steps=1, lr=0.5, window=win, ckpt_every=20)
    # the same window at the same theta (the one the Adam step ran at), per replica
    for r in range(R):
        e.set_theta(hist[0]["theta"], r)
    e.window_rewind()
    lb, gb, gbar1 = e.window_grad(win, spec, target, ckpt_every=20)
    free = list(fit.A_IJ)
    gm = gb.mean(0)
    dmax = np.abs(hist[0]["grad"][free] - gm[free]).max() / np.abs(gm[free]).max()
    say(f"  loss {hist[0]['loss']:.8f} (mean {lb.mean():.8f}); |g_fit - mean_r g_r| / |g| = {dmax:.2e}")
    assert abs(hist[0]["loss"] - lb.mean()) <= 1e-12 * lb.mean()
    assert dmax < 1e-9
    e.window_rewind()
    gbar2 = e.window_forward(win, spec)
    assert np.array_equal(gbar1, gbar2), "a rewound window does not repeat"
    e.close()
    for r, (s, order) in st.items():
        say(f"  replica {r} on a single-replica engine")
        q = eng.Engine(ps[r])
        q.set_state(s, step=s0, slot_gid=order)
        l1, g1, _ = q.window_grad(win, spec, target, ckpt_every=20)
        q.close()
        assert l1[0] == lb[r] and np.array_equal(g1[0], gb[r]), r
