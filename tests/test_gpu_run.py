"""The run driver (python -m welcome-youtube_amd.run) against the reference's
own restart checks (cel/engine/tests/validate.sh:80-105): a restart from a
checkpoint continues the uninterrupted run bit for bit (snapshots, stats rows,
events, final checkpoint), and a run stopped by SIGTERM, then restarted into
the same directory, leaves the same files as the uninterrupted run."""
import importlib
import os
import signal
import subprocess
import sys
import time

import pytest

from conftest import REPO

pytestmark = pytest.mark.gpu

OV = dict(L="20", chem="0", pchem="1", f_lip="0.2", c_N="0.2", lip_nH="1", lip_nT="2", f_Ka="0", f_Kb="0",
          k_angle="0", warmup="20", dt="0.01", p_conv="0.05", p_form="0.05", p_break="0.01", a_N_T="50", seed="5",
          events="1", log_every="0")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def args(**kw):
    d = dict(OV, **kw)
    return [f"{k}={v}" for k, v in d.items()]


def rows_after(path, step):
    with open(path) as f:
        lines = f.readlines()
    return [ln for ln in lines[1:] if int(ln.split(",", 1)[0]) > step]


def same_bytes(a, b):
    with open(a, "rb") as fa, open(b, "rb") as fb:
        return fa.read() == fb.read()


def test_restart_continues_bitwise(tmp_path):
    run = importlib.import_module("welcome-youtube_amd.run")
    ini = tmp_path / "empty.ini"
    ini.write_text("# all parameters on the command line\n")
    a, b = str(tmp_path / "A"), str(tmp_path / "B")
    say("run A: 300 steps, checkpoint at 150")
    run.run(str(ini), a, args(steps=300, snap_every=100, stats_every=10, ckpt_every=150, ckpt_keep=1))
    ck = os.path.join(a, "ckpt_0000000150.bin")
    assert os.path.exists(ck)
    say("run B: restart from step 150")
    run.run(str(ini), b, args(steps=300, snap_every=100, stats_every=10, ckpt_every=150, restart=ck))
    for st in ("0000000200", "0000000300"):
        assert same_bytes(os.path.join(a, f"snap_{st}.bin"), os.path.join(b, f"snap_{st}.bin")), st
    assert same_bytes(os.path.join(a, "ckpt.bin"), os.path.join(b, "ckpt.bin"))
    ra, rb = rows_after(os.path.join(a, "stats.csv"), 150), rows_after(os.path.join(b, "stats.csv"), 150)
    assert len(ra) == 15 and ra == rb
    ea, eb = rows_after(os.path.join(a, "events.csv"), 150), rows_after(os.path.join(b, "events.csv"), 150)
    say(f"events after step 150: {len(ea)}")
    assert len(ea) > 0 and ea == eb


def test_sigterm_stop_and_restart_same_files(tmp_path):
    run = importlib.import_module("welcome-youtube_amd.run")
    ini = tmp_path / "empty.ini"
    ini.write_text("")
    cad = dict(steps=30000, snap_every=10000, stats_every=100, ckpt_every=10000)
    c, d = str(tmp_path / "C"), str(tmp_path / "D")
    say("run C: uninterrupted")
    run.run(str(ini), c, args(**cad))
    say("run D: subprocess, SIGTERM once stepping")
    env = dict(os.environ, PYTHONPATH=REPO + os.pathsep + os.environ.get("PYTHONPATH", ""))
    proc = subprocess.Popen([sys.executable, "-m", "welcome-youtube_amd.run", str(ini), d] + args(**cad), cwd=REPO,
                            env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
    log = os.path.join(d, "log.txt")
    t0 = time.time()
    while time.time() - t0 < 60:
        if os.path.exists(log) and "setup" in open(log).read():
            break
        if proc.poll() is not None:
            break
        time.sleep(0.02)
    time.sleep(0.2)
    proc.send_signal(signal.SIGTERM)
    _, err = proc.communicate(timeout=60)
    assert proc.returncode == 0, err.decode()[-2000:]
    text = open(log).read()
    stopped = "stopping at step" in text
    say("D stopped early:", stopped)
    assert stopped, "the run finished before the signal arrived"
    say("run D: restart into the same directory")
    run.run(str(ini), d, args(restart=os.path.join(d, "ckpt.bin"), **cad))
    for f in ("stats.csv", "events.csv", "snap_0000030000.bin", "ckpt.bin", "snap_0000010000.bin"):
        assert same_bytes(os.path.join(c, f), os.path.join(d, f)), f


def test_cli_with_catalysts_vs_the_gpu_port(tmp_path):
    """The reference's default physics with busier catalysts, 30 steps from a
    fresh soup, three chemistry checks: the reference GPU port's own CPU build
    (tests/golden/catalysts.npz) and this CLI write the same events (as
    multisets per step; the port lists reactions in priority order) and end
    with the same types and lipid chains; positions agree to 1e-6 typically
    and 1e-4 at worst (the two FP32 pair kernels round differently and 30
    steps at dt 0.03 amplify that)."""
    import json

    import numpy as np

    from conftest import GOLDEN

    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    run = importlib.import_module("welcome-youtube_amd.run")
    G = np.load(os.path.join(GOLDEN, "catalysts.npz"))
    ov = json.loads(str(G["cli_overrides"]))
    ini = tmp_path / "empty.ini"
    ini.write_text("# all parameters on the command line\n")
    d = str(tmp_path / "cat")
    say("CLI run with catalysts, 30 steps")
    run.run(str(ini), d, [f"{k}={v}" for k, v in ov.items()])
    with open(os.path.join(d, "events.csv")) as f:
        ours = sorted(ln.strip().split(",") for ln in f.readlines()[1:])
    ref = sorted(ln.split(",") for ln in str(G["cli_events"]).strip().splitlines()[1:])
    say(f"events: {len(ours)} (port {len(ref)})")
    assert [(r[0], r[1], r[2], r[5]) for r in ours] == [(r[0], r[1], r[2], r[5]) for r in ref]
    assert all(abs(float(a[3]) - float(b[3])) < 2e-4 and abs(float(a[4]) - float(b[4])) < 2e-4
               for a, b in zip(ours, ref) if a[3] != "nan")
    g = oracle.ckpt_gid_order(oracle.read_ckpt(os.path.join(d, "ckpt.bin")))
    for k in ("type", "mpos", "nxt"):
        assert np.array_equal(g[k], G[f"cli_{k}"]), k
    L = float(ov["L"])
    for k in ("x", "y"):
        dd = np.abs((g[k] - G[f"cli_{k}"] + L / 2) % L - L / 2)
        say(f"{k}: median {np.median(dd):.2e}, max {dd.max():.2e}")
        assert np.median(dd) < 1e-6 and dd.max() < 1e-4, (k, np.median(dd), dd.max())


def test_event_overflow_does_not_stop_the_run(tmp_path, monkeypatch):
    """With an event buffer far smaller than a chunk's events (OC_EVENT_CAP),
    events are drained at every output point, an overflow is logged as a
    warning, and the run still ends with its checkpoint (ckpt_every = 0)."""
    monkeypatch.setenv("OC_EVENT_CAP", "8")
    run = importlib.import_module("welcome-youtube_amd.run")
    io = importlib.import_module("welcome-youtube_amd.io")
    ini = tmp_path / "empty.ini"
    ini.write_text("# all parameters on the command line\n")
    d = str(tmp_path / "ovf")
    say("run: 300 steps, 8-event buffer, no periodic checkpoint")
    assert run.run(str(ini), d, args(steps=300, stats_every=50, ckpt_every=0, p_conv="0.3", p_form="0.3")) == 300
    log = open(os.path.join(d, "log.txt")).read()
    assert "event buffer overflowed" in log and "done at step 300" in log
    s, hdr, _ = io.read_checkpoint(os.path.join(d, "ckpt.bin"))
    assert hdr["step"] == 300
    assert not os.path.exists(os.path.join(d, "ckpt.bin.tmp"))
    rows = open(os.path.join(d, "events.csv")).read().splitlines()[1:]
    assert len(rows) > 0 and all(0 < int(r.split(",", 1)[0]) <= 300 for r in rows)
