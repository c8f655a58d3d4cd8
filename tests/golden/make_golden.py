"""Write the golden fixtures of tests/golden from the reference engine itself.

Run in a container that holds the reference source (the fixtures are committed;
tests only read them):

    make -C oracle ref && python tests/golden/make_golden.py

Every vector comes from the reference compiled from cel/engine/oercel.c
(oracle/_ref: the static step functions behind ref_shim.c, and the reference
CLI).  The reference is built without FMA contraction (its own bitwise mode),
so the integer/ordering parts and the soup bits are exact references for the
restatement, and the FP64 forces are references to ~1e-13.

Fixtures (numpy .npz / JSON, small systems):
  hashes.npz        mix13, step_seed, pair_theta, hash_key, hash_u01 on inputs
  soups.npz         init_soup + set_velocities (gid order) of three soups
  sorts.npz         sort_cells permutation + cell starts (nsub 1, 3, 4)
  forces.npz        pair_forces (plain; stirring + kT cycle) and bonded_forces
  traj.npz          the CLI's checkpoint after 20 steps from a fresh soup
  params.json       the reference's --params table: keys, defaults, a_ij
  templates.npz     branched lipids (pi, y, bola): soups, bonded forces of
                    distorted lipids, and a 20-step CLI trajectory of pi lipids
  rchem.npz         a 20-step CLI run with the reference chemistry on and no
                    catalysts (decay, vent feed): checkpoint and events.csv
  catalysts.npz     the reference GPU port (cel/engine_gpu/oercel_gpu.cu, built
                    as its own CPU emulation, oracle/_ref/oercel_gpu_emu) over
                    one chemistry check with catalysts: its checkpoint before
                    (step 9) and after (step 10) and its events.csv, for seven
                    configurations (linear / pi / y / bola lipids, syn_nN,
                    syn_tries 1, feed cap, vent, nsub 2 / 3 / 4)

  chemrates.json    the reference CPU engine's catalyst chemistry statistics
                    in the two modes of cel/engine_gpu/tests/chemstat.sh (event
                    totals, catalyst-checks, checks)

`python tests/golden/make_golden.py templates` (or `rchem`, `catalysts`,
`chemrates`) writes that file only.
"""
import ctypes as C
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))
import oracle  # noqa: E402
import importlib  # noqa: E402

pm = importlib.import_module("welcome-youtube_amd.params")

# soups: (name, L, nsub, overrides)
SOUPS = [
    ("forcetest", 12.5, 4, dict(f_lip=0.15, f_Ka=0.02, f_Kb=0.02, c_N=0.1, seed=99)),
    ("cfg1like", 20.0, 4, dict(f_lip=0.1, lip_nT=2, c_N=0.0, f_Ka=0.0, seed=1)),
    ("precursor", 16.0, 2, dict(f_lip=0.3, lip_nT=2, c_N=0.2, f_Ka=0.0, seed=7)),
]


def mk(L, nsub, **kw):
    p = pm.Params()
    d = {"L": repr(L), "nsub": str(nsub)}
    d.update({k: str(v) for k, v in kw.items()})
    p.update(d)
    return p.finalize()


def hashes(R):
    rng = np.random.default_rng(2024)
    z = rng.integers(0, 2**63, 64, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, 64, dtype=np.uint64)
    seeds = z[:16]
    steps = rng.integers(0, 10**7, 16)
    gi = rng.integers(0, 2**29, 64).astype(np.uint32)
    gj = rng.integers(0, 2**29, 64).astype(np.uint32)
    salts = rng.integers(0, 64, 16).astype(np.uint64)
    out = {
        "z": z, "mix13": np.array([R.ref_mix13(int(v)) for v in z], np.uint64),
        "seeds": seeds, "steps": steps,
        "step_seed": np.array([R.ref_step_seed(int(s), int(t)) for s, t in zip(seeds, steps)], np.uint64),
        "gi": gi, "gj": gj,
        "pair_theta": np.array([R.ref_pair_theta(int(z[k % 64]), int(gi[k]), int(gj[k])) for k in range(64)]),
        "salts": salts,
        "hash_key": np.array([R.ref_hash_key(int(s), int(t), int(a)) for s, t, a in zip(seeds, steps, salts)],
                             np.uint64),
    }
    keys = out["hash_key"]
    out["hash_u01"] = np.array([R.ref_hash_u01(int(keys[k % 16]), int(gi[k])) for k in range(64)])
    return out


# branched lipid templates: (name, L, overrides)
TEMPLATES = [
    ("pi", 14.0, dict(lip_arch="pi", lip_nT=2, lip_tie=1, r0_head=0.7, k_head=5, cos_head=0.3, k_angle=10)),
    ("pi_membranes", 14.0, dict(lip_arch="pi", lip_nT=1, lip_tie=1, r0_head=0.7, k_bond=300)),
    ("y", 14.0, dict(lip_arch="y", lip_nT=2, lip_tie=1, k_head=4, cos_head=0.2, k_angle=8)),
    ("bola", 14.0, dict(lip_arch="bola", lip_nT=3, lip_tie=2, r0_head=0.6, k_head=3, cos_head=-0.1, k_angle=6)),
]


def templates():
    out = {}
    rng = np.random.default_rng(11)
    for name, L, kw in TEMPLATES:
        p = mk(L, 4, f_lip=0.3, c_N=0.1, f_Ka=0.0, seed=3, **kw)
        cp = p.to_c()
        s = oracle.ref_soup(cp, p.N)
        for k, v in s.items():
            out[f"{name}_{k}"] = v
        st = dict(s, vtx=s["vx"], vty=s["vy"])
        xb, yb = s["x"].copy(), s["y"].copy()
        lip = s["mpos"] >= 0
        xb[lip] = (xb[lip] + 0.15 * rng.standard_normal(lip.sum())) % L
        yb[lip] = (yb[lip] + 0.15 * rng.standard_normal(lip.sum())) % L
        fx, fy, mb = oracle.ref_forces_gid(cp, dict(st, x=xb, y=yb), 0, 0.0, with_pair=False, with_bonded=True)
        out[f"{name}_bx"], out[f"{name}_by"] = xb, yb
        out[f"{name}_bfx"], out[f"{name}_bfy"], out[f"{name}_maxbond"] = fx, fy, np.array(mb)
    # a 20-step CLI run of the membranes lipid (tied pi H2T2) with its a_ij
    with tempfile.TemporaryDirectory() as tmp:
        ov = dict(L="16", dt="0.02", f_lip="0.2", c_N="0.1", f_Ka="0", lip_arch="pi", lip_nT="1", lip_tie="1",
                  r0_head="0.7", k_bond="300", a_W_T="200", a_H_N="22", a_W_N="45", a_T_N="120", warmup="10",
                  seed="4", nsub="4", chem="0", steps="20", snap_every="0", stats_every="0", ckpt_every="0",
                  log_every="0", events="0", threads="2")
        oracle.run_ref_cli(os.path.join(tmp, "pi"), ov)
        ck = oracle.read_ckpt(os.path.join(tmp, "pi", "ckpt.bin"))
        g = oracle.ckpt_gid_order(ck)
        for k in ("x", "y", "vx", "vy", "fx", "fy", "type", "mpos", "nxt"):
            out[f"traj_{k}"] = g[k]
        out["traj_slot_gid"] = ck["gid"]
        out["traj_step"] = np.array(ck["step"])
        out["traj_maxbond"] = np.array(ck["maxbond"])
        out["traj_overrides"] = np.array(json.dumps(ov))
    np.savez_compressed(os.path.join(HERE, "templates.npz"), **out)


def rchem():
    """The reference CLI with its chemistry on and no catalysts: lipid and N
    decay, vent feed (no cap), checks every 5 steps, 20 steps; its checkpoint
    and events.csv."""
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        ov = dict(L="16", dt="0.01", f_lip="0.3", lip_nT="2", c_N="0.2", f_Ka="0", f_Kb="0", k_angle="5",
                  warmup="0", seed="6", nsub="4", chem="1", n_chem="5", p_ld="0.05", p_nd="0.05", feed="vent",
                  vent_shape="vstrip", vent_x="4", vent_w="6", p_feed="0.2", feed_cap="0", steps="20",
                  snap_every="0", stats_every="0", ckpt_every="0", log_every="0", events="1", threads="2")
        oracle.run_ref_cli(os.path.join(tmp, "rc"), ov)
        ck = oracle.read_ckpt(os.path.join(tmp, "rc", "ckpt.bin"))
        g = oracle.ckpt_gid_order(ck)
        for k in ("x", "y", "vx", "vy", "fx", "fy", "type", "mpos", "nxt"):
            out[f"rc_{k}"] = g[k]
        out["rc_slot_gid"] = ck["gid"]
        out["rc_step"] = np.array(ck["step"])
        out["rc_overrides"] = np.array(json.dumps(ov))
        with open(os.path.join(tmp, "rc", "events.csv")) as f:
            out["rc_events"] = np.array(f.read())
    np.savez_compressed(os.path.join(HERE, "rchem.npz"), **out)


# catalyst chemistry configurations: (name, overrides on top of CAT_BASE)
CAT_BASE = dict(L="16", dt="0.01", f_lip="0.1", lip_nT="2", c_N="0.3", f_Ka="0.03", f_Kb="0.02", warmup="0",
                nsub="4", chem="1", n_chem="10", p_syn_a="0.3", p_syn_b="0.5", p_rep_a="0.2", p_rep_b="0.2",
                mu="0.3", p_kd="0.02", p_ld="0.01", p_nd="0.01", feed="uniform", p_feed="0.05", feed_cap="0.25",
                snap_every="0", stats_every="0", ckpt_every="0", log_every="0", events="1", threads="1")
CATS = [
    ("linear", dict(seed="5")),
    ("capped", dict(seed="8", c_N="0.2", feed_cap="0.22", p_feed="0.2")),
    ("pi", dict(seed="9", lip_arch="pi", c_N="0.35")),
    ("synN2", dict(seed="10", syn_nN="2")),
    ("tries1", dict(seed="11", syn_tries="1", nsub="2", c_N="0.15")),
    ("bola", dict(seed="12", lip_arch="bola", c_N="0.4", L="20")),
    ("yvent", dict(seed="13", lip_arch="y", nsub="3", feed="vent", vent_r="5", p_feed="0.3", c_N="0.15")),
]


def catalysts():
    """The GPU port's CPU emulation: 9 steps from its soup (no check yet:
    n_chem = 10), then a restart from that checkpoint for the step with the
    check; both checkpoints (gid order) and the events.csv of the check."""
    emu = os.path.join(REPO, "oracle", "_ref", "oercel_gpu_emu")
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, kw in CATS:
            ov = dict(CAT_BASE, **kw)
            a, b = os.path.join(tmp, name + "_a"), os.path.join(tmp, name + "_b")
            for d, extra in ((a, dict(steps="9")), (b, dict(steps="10", restart=os.path.join(a, "ckpt.bin")))):
                cmd = [emu, "/dev/null", d] + [f"{k}={v}" for k, v in dict(ov, **extra).items()]
                subprocess.run(cmd, capture_output=True, text=True, check=True)
            for tag, d in (("in", a), ("out", b)):
                g = oracle.ckpt_gid_order(oracle.read_ckpt(os.path.join(d, "ckpt.bin")))
                for k in ("x", "y", "vx", "vy", "fx", "fy", "type", "mpos", "nxt"):
                    if tag == "in" or k not in ("fx", "fy"):
                        out[f"{name}_{tag}_{k}"] = g[k]
            with open(os.path.join(b, "events.csv")) as f:
                out[f"{name}_events"] = np.array(f.read())
            out[f"{name}_overrides"] = np.array(json.dumps(ov))
        # a 30-step CLI run from the port's soup with the reference's default
        # physics (warm-up ramp, dt, a_ij, feed cap) and busier catalysts
        ov = dict(L="24", f_Ka="0.03", f_Kb="0.02", c_N="0.2", p_syn_a="0.2", p_syn_b="0.3", p_rep_a="0.05",
                  p_rep_b="0.05", mu="0.3", p_kd="0.01", p_ld="0.01", p_feed="0.05", seed="17", steps="30",
                  snap_every="0", stats_every="0", ckpt_every="0", log_every="0", events="1", threads="1")
        d = os.path.join(tmp, "cli")
        cmd = [emu, "/dev/null", d] + [f"{k}={v}" for k, v in ov.items()]
        subprocess.run(cmd, capture_output=True, text=True, check=True)
        g = oracle.ckpt_gid_order(oracle.read_ckpt(os.path.join(d, "ckpt.bin")))
        for k in ("x", "y", "vx", "vy", "type", "mpos", "nxt"):
            out[f"cli_{k}"] = g[k]
        with open(os.path.join(d, "events.csv")) as f:
            out["cli_events"] = np.array(f.read())
        out["cli_overrides"] = np.array(json.dumps(ov))
    np.savez_compressed(os.path.join(HERE, "catalysts.npz"), **out)


# the reference's own chemistry statistics setup (cel/engine_gpu/tests/chemstat.sh):
# well-mixed water with catalysts and nutrients, checks every 10 steps, two modes
CHEMSTAT_BASE = dict(dt="0.03", L="48", f_lip="0", f_Ka="0.02", f_Kb="0.01", c_N="0.12", chem="1", n_chem="10",
                     p_nd="0.001", feed="uniform", p_feed="0.01", feed_cap="0.12", steps="20000", snap_every="0",
                     stats_every="100", log_every="0", ckpt_every="0", events="1", seed="11")
CHEMSTAT_MODES = {
    "syn": dict(p_syn_a="0.02", p_syn_b="0.06", p_rep_a="0", p_rep_b="0", p_kd="0", p_ld="0.02"),
    "rep": dict(p_syn_a="0", p_syn_b="0", p_rep_a="0.005", p_rep_b="0.005", mu="0.3", p_kd="0.0025", p_ld="0"),
}


def chem_counts(events_csv, stats_csv, n_chem):
    """Event totals per kind (feed: conversions), catalyst-checks (the mean
    catalyst count of the stats rows times the checks) and checks, as
    cel/engine_gpu/tests/compare.py:227-253 sums them."""
    cnt = {}
    for ln in events_csv.splitlines()[1:]:
        f = ln.split(",")
        cnt[f[1]] = cnt.get(f[1], 0) + (int(f[2]) if f[1] == "feed" else 1)
    rows = [ln.split(",") for ln in stats_csv.splitlines()[1:]]
    steps = max(int(r[0]) for r in rows)
    nk = np.mean([int(r[7]) + int(r[8]) for r in rows])  # nKa + nKb
    nc = steps // n_chem
    return {"events": cnt, "catalyst_checks": float(nk * nc), "checks": nc}


def chemrates():
    """The reference CPU engine (its serial chemistry) in both chemstat modes:
    event totals and catalyst-checks."""
    out = {"base": CHEMSTAT_BASE, "modes": {}}
    with tempfile.TemporaryDirectory() as tmp:
        for mode, kw in CHEMSTAT_MODES.items():
            ov = dict(CHEMSTAT_BASE, threads="4", **kw)
            d = os.path.join(tmp, mode)
            oracle.run_ref_cli(d, ov, fast=True, timeout=3000)
            ev = open(os.path.join(d, "events.csv")).read()
            st = open(os.path.join(d, "stats.csv")).read()
            out["modes"][mode] = dict(chem_counts(ev, st, 10), overrides=kw)
    with open(os.path.join(HERE, "chemrates.json"), "w") as f:
        json.dump(out, f, indent=1)


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref is not built: make -C oracle ref")
    R = oracle.ref()
    np.savez_compressed(os.path.join(HERE, "hashes.npz"), **hashes(R))

    soups, sorts, forces = {}, {}, {}
    rng = np.random.default_rng(7)
    for name, L, nsub, kw in SOUPS:
        p = mk(L, nsub, **kw)
        cp = p.to_c()
        s = oracle.ref_soup(cp, p.N)
        for k, v in s.items():
            soups[f"{name}_{k}"] = v
        soups[f"{name}_L"] = np.array(L)
        # sort: moved positions, several nsub
        x = (s["x"] + 0.4 * rng.standard_normal(p.N)) % L
        y = (s["y"] + 0.4 * rng.standard_normal(p.N)) % L
        for ns in (1, 3, 4):
            cp.nsub = ns
            src, cs = oracle.ref_sort(cp, x, y)
            sorts[f"{name}_n{ns}_src"] = src
            sorts[f"{name}_n{ns}_cstart"] = cs
        sorts[f"{name}_x"], sorts[f"{name}_y"] = x, y
        cp.nsub = nsub
        # pair forces at step 12345, plain and with stirring + temperature cycle
        vtx = s["vx"] + 0.3 * rng.standard_normal(p.N)
        vty = s["vy"].copy()
        st = dict(s, vtx=vtx, vty=vty)
        forces[f"{name}_vtx"], forces[f"{name}_vty"] = vtx, vty
        fx, fy, _ = oracle.ref_forces_gid(cp, st, 12345, cp.t_anchor + 12345 * cp.dt)
        forces[f"{name}_fx"], forces[f"{name}_fy"] = fx, fy
        cpe = mk(L, nsub, stir_A=0.7, stir_n=2, kT_amp=0.3, kT_period=10.0, **kw).to_c()
        fx, fy, _ = oracle.ref_forces_gid(cpe, st, 12345, 2.5)
        forces[f"{name}_env_fx"], forces[f"{name}_env_fy"] = fx, fy
        # bonded forces of distorted lipids, with angle terms
        xb, yb = s["x"].copy(), s["y"].copy()
        lip = s["mpos"] >= 0
        xb[lip] = (xb[lip] + 0.15 * rng.standard_normal(lip.sum())) % L
        yb[lip] = (yb[lip] + 0.15 * rng.standard_normal(lip.sum())) % L
        sb = dict(st, x=xb, y=yb)
        fx, fy, mb = oracle.ref_forces_gid(cp, sb, 0, 0.0, with_pair=False, with_bonded=True)
        forces[f"{name}_bx"], forces[f"{name}_by"] = xb, yb
        forces[f"{name}_bfx"], forces[f"{name}_bfy"], forces[f"{name}_maxbond"] = fx, fy, np.array(mb)
    np.savez_compressed(os.path.join(HERE, "soups.npz"), **soups)
    np.savez_compressed(os.path.join(HERE, "sorts.npz"), **sorts)
    np.savez_compressed(os.path.join(HERE, "forces.npz"), **forces)

    # trajectory: the reference CLI, fresh soup, 20 steps, warm-up ramp active
    traj = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, ov in (("lipid", dict(L="16", dt="0.01", f_lip="0.2", lip_nT="2", c_N="0", f_Ka="0", k_angle="10",
                                        warmup="50", seed="5", nsub="4")),
                         ("water", dict(L="14.5", dt="0.03", f_lip="0", c_N="0", f_Ka="0", warmup="0", seed="9",
                                        nsub="2"))):
            ov = dict(ov, chem="0", steps="20", snap_every="0", stats_every="0", ckpt_every="0", log_every="0",
                      events="0", threads="2")
            oracle.run_ref_cli(os.path.join(tmp, name), ov)
            ck = oracle.read_ckpt(os.path.join(tmp, name, "ckpt.bin"))
            g = oracle.ckpt_gid_order(ck)
            for k in ("x", "y", "vx", "vy", "fx", "fy", "type", "mpos", "nxt"):
                traj[f"{name}_{k}"] = g[k]
            traj[f"{name}_slot_gid"] = ck["gid"]
            traj[f"{name}_step"] = np.array(ck["step"])
            traj[f"{name}_maxbond"] = np.array(ck["maxbond"])
            traj[f"{name}_overrides"] = np.array(json.dumps(ov))
    np.savez_compressed(os.path.join(HERE, "traj.npz"), **traj)

    # parameter table: keys and defaults only (the reference's --params)
    out = subprocess.run([oracle.ref_binary(), "--params"], capture_output=True, text=True, check=True).stdout
    keys = re.findall(r"^\| `([^`]+)` \| ([^|]*) \|", out, re.M)
    amat = re.findall(r"^\| \*\*(\w+)\*\* \|(.*)\|$", out, re.M)
    tab = {"defaults": {k: (v.strip() if v.strip() != "(empty)" else "") for k, v in keys},
           "order": [k for k, _ in keys],
           "amatrix": {n: [float(x) for x in row.split("|") if x.strip()] for n, row in amat}}
    with open(os.path.join(HERE, "params.json"), "w") as f:
        json.dump(tab, f, indent=1)
    templates()
    rchem()
    catalysts()
    chemrates()
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    if sys.argv[1:] in (["templates"], ["rchem"], ["catalysts"], ["chemrates"]):
        if not oracle.ref_available():
            sys.exit("oracle/_ref is not built: make -C oracle ref")
        {"templates": templates, "rchem": rchem, "catalysts": catalysts, "chemrates": chemrates}[sys.argv[1]]()
    else:
        main()
