"""Checks of the oracle that do not need the reference: brute force, finite
differences, closed forms, known-answer counts, and an independent numpy
statement of the precursor chemistry.  CPU only."""
import numpy as np
import pytest

from conftest import mkparams


def test_feed_known_answer(oracle):
    """Uniform food supply on water, L=48 (N=6912), seed 7, p_feed 0.003,
    n_chem 10, 600 steps: 1138 conversions (cel/engine/README.md:519-520,
    cel/engine/tests/validate.sh:133-134).  Depends only on hash_key/hash_u01."""
    assert oracle.lib().or_feed_count(7, 6912, 0.003, 10, 600) == 1138


def test_noise_moments_and_symmetry(oracle):
    L = oracle.lib()
    ss = L.or_step_seed(99, 12345)
    th = np.array([L.or_pair_theta(ss, i, j) for i in range(400) for j in range(i + 1, i + 300)])
    assert abs(th.mean()) < 0.01 and abs(th.var() - 1.0) < 0.01
    assert th.min() >= -np.sqrt(3) and th.max() < np.sqrt(3)
    assert L.or_pair_theta(ss, 5, 9) == L.or_pair_theta(ss, 9, 5)
    # the device's 32-bit cut of the same deviate
    d = max(abs(L.or_pair_theta(ss, i, j) - L.or_pair_theta32(ss, i, j)) for i in range(50) for j in range(50, 90))
    assert d <= np.sqrt(12) / 2**32 * 1.0001


@pytest.mark.parametrize("L,nsub", [(9.0, 1), (12.5, 4), (13.7, 3), (10.2, 8)])
def test_cell_list_forces_equal_brute_force(oracle, L, nsub):
    """Full-shell cell-list traversal finds every pair (O(N^2) minimum image)."""
    p = mkparams(L, nsub=nsub, f_lip=0.15, c_N=0.1, seed=3)
    rng = np.random.default_rng(1)
    s = oracle.soup(p.to_c(), p.N)
    s["vtx"] = s["vx"] + 0.3 * rng.standard_normal(p.N)
    s["vty"] = s["vy"]
    a = oracle.pair_forces_gid(p.to_c(), s, 77, theta32=0)
    bx, by = oracle.pair_brute(p.to_c(), s, 77, theta32=0)
    F = np.hypot(bx, by).max()
    assert np.abs(a["fx"] - bx).max() < 1e-12 * F and np.abs(a["fy"] - by).max() < 1e-12 * F
    assert abs(a["fx"].sum()) < 1e-9 and abs(a["fy"].sum()) < 1e-9


def bonded_energy(x, y, mpos, nxt, bond, L, kb, r0, ka):
    """Independent bonded energy: harmonic bonds and U = ka (1 - cos) angles."""
    def mi(d):
        return (d + L / 2) % L - L / 2
    U = 0.0
    for h in np.flatnonzero(mpos == 0):
        ch = [h]
        while nxt[ch[-1]] >= 0:
            ch.append(nxt[ch[-1]])
        v = [np.array([mi(x[ch[b + 1]] - x[ch[b]]), mi(y[ch[b + 1]] - y[ch[b]])]) for b in range(len(ch) - 1)]
        for d in v:
            U += 0.5 * kb * (np.linalg.norm(d) - r0) ** 2
        for b in range(len(v) - 1):
            U += ka * (1.0 - v[b] @ v[b + 1] / (np.linalg.norm(v[b]) * np.linalg.norm(v[b + 1])))
    for i in np.flatnonzero((mpos < 0) & (bond > np.arange(len(x)))):
        j = bond[i]
        d = np.array([mi(x[j] - x[i]), mi(y[j] - y[i])])
        U += 0.5 * kb * (np.linalg.norm(d) - r0) ** 2
    return U


def test_bonded_forces_finite_differences(oracle):
    p = mkparams(12.0, f_lip=0.3, k_angle=10.0, seed=4)
    rng = np.random.default_rng(2)
    s = oracle.soup(p.to_c(), p.N)
    lip = s["mpos"] >= 0
    s["x"][lip] = (s["x"][lip] + 0.15 * rng.standard_normal(lip.sum())) % p.L
    s["y"][lip] = (s["y"][lip] + 0.15 * rng.standard_normal(lip.sum())) % p.L
    # two chemistry bonds between non-lipid beads
    free = np.flatnonzero(~lip)[:4]
    s["bond"][free[0]], s["bond"][free[1]] = free[1], free[0]
    s["bond"][free[2]], s["bond"][free[3]] = free[3], free[2]
    s["x"][free[1]] = (s["x"][free[0]] + 0.4) % p.L
    s["y"][free[1]] = s["y"][free[0]]
    fx, fy, mb, eb = oracle.bonded_gid(p.to_c(), s)
    args = (s["mpos"], s["nxt"], s["bond"], p.L, p.k_bond, p.r0, p.k_angle)
    assert abs(eb - bonded_energy(s["x"], s["y"], *args)) < 1e-9 * max(1.0, eb)
    h = 1e-6
    worst = 0.0
    for i in list(np.flatnonzero(lip)[:60]) + list(free):
        for arr, fa in ((s["x"], fx), (s["y"], fy)):
            q0 = arr[i]
            arr[i] = q0 + h
            up = bonded_energy(s["x"], s["y"], *args)
            arr[i] = q0 - h
            um = bonded_energy(s["x"], s["y"], *args)
            arr[i] = q0
            worst = max(worst, abs(-(up - um) / (2 * h) - fa[i]))
    assert worst < 1e-5 * max(1.0, np.abs(fx).max())
    assert abs(fx.sum()) < 1e-10 and abs(fy.sum()) < 1e-10


def test_ballistic_motion_without_forces(oracle):
    """With every a_ij = 0, gamma = 0 and no bonds the force vanishes and the
    integrator must move every bead on a straight line: x_n = x_0 + n dt v."""
    p = mkparams(11.0, f_lip=0.0, gamma=0.0, a_W_W=0, seed=6)
    cp = p.to_c()
    s = oracle.soup(cp, p.N)
    sysm = oracle.System(cp, p.N)
    sysm.start(s)
    sysm.step(40)
    o = sysm.get_state()
    L = p.L
    exp_x = (s["x"] + 40 * p.dt * s["vx"]) % L
    d = np.abs((o["x"] - exp_x + L / 2) % L - L / 2).max()
    assert d < 1e-11 and np.array_equal(o["vx"], s["vx"]) and np.all(o["fx"] == 0)


def chem_numpy(p, st, step, oracle):
    """Independent statement of the precursor chemistry rules on gid-order
    arrays (all-pairs distances), returning sorted (kind, a, b) tuples."""
    L = oracle.lib()
    N = len(st["x"])
    x, y = st["x"], st["y"]
    dx = x[None, :] - x[:, None]
    dy = y[None, :] - y[:, None]
    dx = np.where(dx > p.L / 2, dx - p.L, np.where(dx < -p.L / 2, dx + p.L, dx))
    dy = np.where(dy > p.L / 2, dy - p.L, np.where(dy < -p.L / 2, dy + p.L, dy))
    r2 = dx * dx + dy * dy
    np.fill_diagonal(r2, np.inf)
    t, mp, bd = st["type"], st["mpos"], st["bond"]
    kc, kf, kb = (L.or_hash_key(p.seed, step, s) for s in (31, 32, 33))
    ev = []
    lipT = (t == 2) & (mp >= 0)
    for i in range(N):
        if mp[i] < 0 and bd[i] >= 0 and i < bd[i] and L.or_hash_u01(kb, i) < p.p_break:
            ev.append((2, i, bd[i]))
        if t[i] == 3 and L.or_hash_u01(kc, i) < p.p_conv and np.any(lipT & (r2[i] < p.r_conv * p.r_conv)):
            ev.append((0, i, i))
    free = (t == 2) & (mp < 0) & (bd < 0)
    best = np.full(N, -1)
    for i in np.flatnonzero(free):
        cand = [(L.or_pair_u01(kf, i, j), j) for j in np.flatnonzero(free & (r2[i] < p.r_form * p.r_form))]
        cand = [c for c in cand if c[0] < p.p_form]
        if cand:
            best[i] = min(cand)[1]
    for i in range(N):
        j = best[i]
        if j >= 0 and best[j] == i and i < j:
            ev.append((1, i, j))
    return sorted(ev)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_chemistry_oracle_vs_numpy(oracle, seed):
    """The oracle's cell-list chemistry equals an all-pairs numpy statement
    of the same rules (events and their order)."""
    p = mkparams(14.0, f_lip=0.3, c_N=0.2, pchem=1, n_chem=1, p_conv=0.6, p_form=0.6, p_break=0.4, seed=seed)
    cp = p.to_c()
    rng = np.random.default_rng(seed)
    sysm = oracle.System(cp, p.N)
    sysm.start(oracle.soup(cp, p.N))
    st = sysm.get_state()
    P = np.flatnonzero(st["type"] == 3)
    conv = rng.choice(P, size=len(P) // 2, replace=False)
    st["type"][conv] = 2
    q = len(conv) // 4
    a, b = conv[:q], conv[q:2 * q]
    st["bond"][:] = -1
    st["bond"][a], st["bond"][b] = b, a
    sysm.set_state(st, 0)
    sysm.chem(5)
    ev = sysm.events()
    got = sorted((int(e["kind"]), int(e["gid_a"]), int(e["gid_b"])) for e in ev)
    exp = chem_numpy(p, st, 5, oracle)
    assert len(exp) > 10 and got == exp


def test_stats_closed_forms(oracle):
    p = mkparams(12.0, f_lip=0.2, seed=2)
    cp = p.to_c()
    sysm = oracle.System(cp, p.N)
    sysm.start(oracle.soup(cp, p.N))
    sysm.step(3)
    s = sysm.get_state()
    o = sysm.stats()
    vx = s["vx"] + 0.5 * p.dt * s["fx"]
    vy = s["vy"] + 0.5 * p.dt * s["fy"]
    T = (vx @ vx + vy @ vy) / (2 * p.N - 2)
    assert abs(o.T_kin - T) < 1e-12 * T
    assert abs(o.px - vx.sum()) < 1e-10 and abs(o.py - vy.sum()) < 1e-10
    # pair energy: all pairs, a_ij/2 (1 - r)^2
    x, y = s["x"], s["y"]
    dx = (x[None, :] - x[:, None] + p.L / 2) % p.L - p.L / 2
    dy = (y[None, :] - y[:, None] + p.L / 2) % p.L - p.L / 2
    r = np.hypot(dx, dy)
    A = np.array(p.a)[s["type"][:, None], s["type"][None, :]]
    iu = np.triu_indices(p.N, 1)
    rr, aa = r[iu], A[iu]
    m = rr < 1
    E = (0.5 * aa[m] * (1 - rr[m]) ** 2).sum()
    assert abs(o.E_pair - E) < 1e-9 * E
    assert o.count[0] + o.count[1] + o.count[2] == p.N


@pytest.mark.parametrize("nsteps", [20, 200])
def test_window_adjoint_vs_central_differences(oracle, nsteps):
    """The FP64 oracle adjoint of the smooth tail-tail RDF loss against central
    differences with a small step (the loss is smooth at that scale; over 200
    chaotic steps eps = 1e-3 is no longer in its linear range)."""
    p = mkparams(12.0, f_lip=0.3, seed=3)
    cp = p.to_c()
    sysm = oracle.System(cp, p.N)
    sysm.start(oracle.soup(cp, p.N))
    sysm.step(150)
    st0 = sysm.get_state()
    spec = oracle.rdf_spec(2, 2, 32, 3.0)
    IJ = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]

    def run(theta, tgt):
        s = oracle.System(cp, p.N)
        s.set_state(st0, 150)
        a = np.array(p.a, dtype=float)
        for q, (i, j) in enumerate(IJ):
            a[i][j] = a[j][i] = theta[q]
        s.set_pair_params(a, theta[6])
        return s.window_grad(nsteps, spec, tgt)

    th0 = np.array([p.a[i][j] for i, j in IJ] + [p.gamma])
    tgt = run(th0, np.zeros(32))[2] * 1.05
    _, g, _ = run(th0, tgt)
    eps = 1e-6 if nsteps <= 20 else 1e-7
    for q in (2, 5, 6):
        e = np.eye(7)[q] * eps
        fd = (run(th0 + e, tgt)[0] - run(th0 - e, tgt)[0]) / (2 * eps)
        assert abs(fd - g[q]) < 2e-5 * np.abs(g).max(), (q, fd, g[q])


def test_rdf_of_ideal_gas_is_one(oracle):
    """The smooth RDF's normalisation, pinned by a closed form: for uniformly
    random positions E[g_b] = 1 - 1/n (n beads of the type, pairs counted once)
    in every bin whose kernel support stays at r > 0, since a symmetric kernel
    maps the linear pair density 2 pi r rho onto itself.  A factor 2 in the
    pair count, a wrong shell area or kernel constant fails this."""
    p = mkparams(30.0, f_lip=0.0, seed=1)
    N = p.N
    rng = np.random.default_rng(7)
    spec = oracle.rdf_spec(2, 2, 24, 3.0)
    acc = np.zeros(24)
    n_t = N // 2
    samples = 40
    for k in range(samples):
        s = {"x": rng.uniform(0, p.L, N), "y": rng.uniform(0, p.L, N), "vx": np.zeros(N), "vy": np.zeros(N),
             "type": np.where(np.arange(N) < n_t, 2, 0).astype(np.uint8), "mpos": np.full(N, -1, np.int8),
             "nxt": np.full(N, -1, np.int32), "bond": np.full(N, -1, np.int32)}
        o = oracle.System(p.to_c(), N)
        o.set_state(s, 0)
        acc += o.rdf(spec)
    g = acc / samples
    want = 1.0 - 1.0 / n_t
    # bins 1..23: kernel support [(b - 0.5) dr, (b + 1.5) dr] inside r > 0;
    # shot noise of bin b ~ 1/sqrt(pairs in its shell over all samples)
    dr = 3.0 / 24
    for b in range(1, 24):
        pairs = samples * 0.5 * n_t * (n_t / p.L ** 2) * np.pi * (((b + 1) * dr) ** 2 - (b * dr) ** 2)
        assert abs(g[b] - want) < 5.0 / np.sqrt(pairs), (b, g[b])
    # bin 0: the part of the kernel below r = 0 (where the linear density would
    # be negative) is cut off, so E[g_0] = want * int_0 K(r - c) r dr / c
    c, h = 0.5 * dr, dr
    r = np.linspace(0.0, c + h, 200001)
    t = (r - c) / h
    K = 35.0 / (32.0 * h) * (1 - t * t) ** 3
    f = K * r
    want0 = want * float(np.sum((f[1:] + f[:-1]) * 0.5 * np.diff(r))) / c
    pairs0 = samples * 0.5 * n_t * (n_t / p.L ** 2) * np.pi * dr ** 2
    assert want0 > want
    assert abs(g[0] - want0) < 5.0 / np.sqrt(pairs0), (g[0], want0)


def test_virial_pressure_vs_the_reference_formula(oracle):
    """The oracle's conservative virial equals the reference water check's own
    sum (1/(2V)) sum_{i<j} a (1 - r) r over pairs found by a periodic k-d tree
    (cel/engine_gpu/tests/compare.py:82-110), and the pressure is
    rho T_kin + that sum; for a mixture the pair's a_ij enters."""
    cKDTree = pytest.importorskip("scipy.spatial").cKDTree
    for kw in (dict(f_lip=0.0, c_N=0.0), dict(f_lip=0.2, c_N=0.1)):
        p = mkparams(17.3, **kw)
        cp = p.to_c()
        o = oracle.System(cp, p.N)
        o.start(oracle.soup(cp, p.N))
        o.step(20)
        st = o.stats()
        s = o.get_state()
        L = p.L
        pos = np.c_[s["x"], s["y"]] % L
        pr = cKDTree(pos, boxsize=L).query_pairs(1.0, output_type="ndarray")
        d = pos[pr[:, 0]] - pos[pr[:, 1]]
        d -= L * np.round(d / L)
        r = np.hypot(d[:, 0], d[:, 1])
        a = np.array(p.a)[s["type"][pr[:, 0]], s["type"][pr[:, 1]]]
        m = r < 1.0
        vir = (a[m] * (1 - r[m]) * r[m]).sum()
        assert abs(st.virial - vir) < 1e-10 * vir
        assert abs(st.pressure - (p.N / L**2 * st.T_kin + vir / (2 * L * L))) < 1e-10 * st.pressure


def pair_table(s, L, rmax, rng):
    """(lo, hi, value) of every gid pair closer than rmax (periodic), sorted by
    (lo, hi), with uniform values on [-sqrt3, sqrt3)"""
    cKDTree = pytest.importorskip("scipy.spatial").cKDTree
    pos = np.c_[s["x"], s["y"]] % L
    pr = cKDTree(pos, boxsize=L).query_pairs(rmax, output_type="ndarray")
    lo, hi = np.minimum(pr[:, 0], pr[:, 1]).astype(np.uint32), np.maximum(pr[:, 0], pr[:, 1]).astype(np.uint32)
    o = np.lexsort((hi, lo))
    return lo[o], hi[o], rng.uniform(-np.sqrt(3), np.sqrt(3), len(o))


def test_injected_noise_forces_vs_numpy_brute_force(oracle):
    """The oracle's pair forces with a caller's noise table equal an
    independent FP64 numpy statement of the DPD pair force with that noise
    (cel/engine/tests/forcetest.c:91-103: conservative + dissipative + random,
    F_ij = w (a_ij - gamma w r.v + sigma/sqrt(dt) theta_ij) r_hat)."""
    rng = np.random.default_rng(4)
    p = mkparams(12.5, nsub=4, f_lip=0.2, c_N=0.1, seed=8)
    cp = p.to_c()
    s = oracle.soup(cp, p.N)
    s["vtx"] = s["vx"] + 0.2 * rng.standard_normal(p.N)
    s["vty"] = s["vy"].copy()
    lo, hi, th = pair_table(s, p.L, 1.001, rng)
    oracle.inject(theta=(lo, hi, th))
    try:
        f = oracle.pair_forces_gid(cp, s, 77, theta32=1)
        assert oracle.inject_missing() == 0
    finally:
        oracle.inject()
    L = p.L
    d = np.c_[s["x"][lo] - s["x"][hi], s["y"][lo] - s["y"][hi]]
    d -= L * np.round(d / L)
    r = np.hypot(d[:, 0], d[:, 1])
    m = r < 1.0
    r, d, th_m, a = r[m], d[m], th[m], np.array(p.a)[s["type"][lo[m]], s["type"][hi[m]]]
    e = d / r[:, None]
    w = 1 - r
    rv = e[:, 0] * (s["vtx"][lo[m]] - s["vtx"][hi[m]]) + e[:, 1] * (s["vty"][lo[m]] - s["vty"][hi[m]])
    sdt = np.sqrt(2 * p.gamma * p.kT) / np.sqrt(p.dt)
    fm = a * w - p.gamma * w * w * rv + sdt * w * th_m
    bfx = np.zeros(p.N)
    bfy = np.zeros(p.N)
    np.add.at(bfx, lo[m], fm * e[:, 0])
    np.add.at(bfy, lo[m], fm * e[:, 1])
    np.add.at(bfx, hi[m], -fm * e[:, 0])
    np.add.at(bfy, hi[m], -fm * e[:, 1])
    assert np.abs(f["fx"] - bfx).max() < 1e-9 * max(1, np.abs(bfx).max())
    assert np.abs(f["fy"] - bfy).max() < 1e-9 * max(1, np.abs(bfy).max())
    # the noise enters: forces with the table differ from the hashed ones
    g = oracle.pair_forces_gid(cp, s, 77, theta32=1)
    assert np.abs(g["fx"] - bfx).max() > 1.0
