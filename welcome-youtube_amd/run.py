"""Command-line run driver: the reference engine's process-level interface
around the HIP step.

    python -m welcome-youtube_amd.run <params.ini> <outdir> [key=value ...]

The loop of the reference's main (cel/engine/oercel.c:2130-2414), same
arguments and files:
  - parameters: defaults, the ini, then key=value overrides (two-pass order);
    params.ini in outdir holds every value used, and a restart whose
    parameters differ keeps the previous file as params_before_<step>.ini;
  - start: the random soup (init = soup) or restart = <checkpoint> (its slot
    order, step, time anchors, warm-up end, seed and running max bond), so a
    restart continues the uninterrupted run bit for bit;
  - output cadences: stats.csv every stats_every steps (and step 0 of a fresh
    run), snap_<step>.bin every snap_every, events.csv (events = 1) with the
    chemistry's events (reference syn / rep / mut / dK / dlip / dN / feed,
    precursor conv / form / break), ckpt.bin every ckpt_every and at the end
    (ckpt_<step>.bin copies with ckpt_keep = 1), a progress line in log.txt
    every log_every steps;
  - a restart into an existing directory truncates stats.csv and events.csv
    after the checkpoint's step and appends;
  - SIGTERM / SIGINT or walltime seconds stop the run cleanly with a
    checkpoint ("stopping at step ..."), to be continued by a restart.

Steps run as hipGraph replays between output points; the stop flag and the
walltime are checked at least every `chunk` steps (100).
"""
import math
import os
import signal
import sys
import time

from . import io
from .abi import OcError
from .engine import Engine
from .params import ParamError, load, mol_template

USAGE = "usage: python -m welcome-youtube_amd.run <params.ini> <outdir> [key=value ...]"
CHUNK = 100

_stop = [False]


def _on_signal(sig, frame):
    _stop[0] = True


def truncate_csv_after(path, step):
    """Keep the header and the rows with step <= `step`
    (truncate_csv_after, cel/engine/oercel.c:2013-2051)."""
    with open(path) as f:
        lines = f.readlines()
    keep = lines[:1]
    for ln in lines[1:]:
        try:
            s = int(ln.split(",", 1)[0])
        except ValueError:
            continue
        if s <= step:
            keep.append(ln)
    with open(path, "w") as f:
        f.writelines(keep)


def _next_output(step, steps, p):
    nxt = min(steps, step + CHUNK)
    for every in (p.stats_every, p.snap_every, p.ckpt_every, p.log_every):
        if every > 0:
            nxt = min(nxt, (step // every + 1) * every)
    return nxt


def run(ini, outdir, overrides=()):
    """Run the engine as the reference's CLI does; returns the final step."""
    p = load(ini, list(overrides))
    os.makedirs(outdir, exist_ok=True)
    log = open(os.path.join(outdir, "log.txt"), "a")

    def logmsg(msg):
        log.write(msg)
        log.flush()
        sys.stdout.write(msg)
        sys.stdout.flush()

    _stop[0] = False
    old = {sg: signal.signal(sg, _on_signal) for sg in (signal.SIGTERM, signal.SIGINT)}
    try:
        return _run(p, outdir, logmsg)
    finally:
        for sg, h in old.items():
            signal.signal(sg, h)
        log.close()


def _run(p, outdir, logmsg):
    t_start = time.perf_counter()
    restarting = bool(p.restart)
    if restarting:
        s0, hdr, _ = io.read_checkpoint(p.restart)
        q = io.restart_params(p, hdr)
        if len(s0["x"]) != q.N:
            raise ParamError(f"checkpoint holds N={len(s0['x'])} beads, the parameters give N={q.N}")
        eng = Engine(q)
        io.load_checkpoint(eng, p.restart)
        step = int(hdr["step"])
        logmsg(f"restart from {p.restart} at step {step} (t={hdr['time']:.4f}), N={q.N}\n")
        p = q
    else:
        eng = Engine(p)
        eng.start()
        step = 0
        if p.warmup > 0:
            logmsg(f"warm-up: repulsion above 25 ramped in over steps 0..{p.warmup}\n")

    # parameters used
    ppath = os.path.join(outdir, "params.ini")
    txt = p.ini_text()
    if restarting and os.path.exists(ppath):
        with open(ppath) as f:
            if f.read() != txt:
                os.replace(ppath, os.path.join(outdir, f"params_before_{step:010d}.ini"))
                logmsg("parameters changed at restart; previous ones kept\n")
    with open(ppath + ".new", "w") as f:
        f.write(txt)
    os.replace(ppath + ".new", ppath)

    spath, epath = os.path.join(outdir, "stats.csv"), os.path.join(outdir, "events.csv")
    if restarting and os.path.exists(spath):
        truncate_csv_after(spath, step)
        sf = open(spath, "a")
    else:
        sf = open(spath, "w")
        sf.write(io.STATS_HEADER + "\n")
    ef = None
    if p.events:
        if restarting and os.path.exists(epath):
            truncate_csv_after(epath, step)
            ef = open(epath, "a")
        else:
            ef = open(epath, "w")
            ef.write(io.EVENTS_HEADER + "\n")

    tbonds = [(b[0], b[1]) for b in mol_template(p).bonds]

    def snapshot(st):
        s = eng.get_state()
        io.write_snapshot(os.path.join(outdir, f"snap_{st:010d}.bin"), s, st, p.t_anchor + (st - p.step_anchor) * p.dt,
                          p.L, p.mlen, tbonds)

    def flush_events():
        # drained at every output point (at most CHUNK steps apart), so the
        # device buffer holds one chunk's events; an overflow is reported in
        # the log and the run goes on
        try:
            ev = eng.events()
        except OcError as err:
            if "overflowed" not in str(err):
                raise
            logmsg(f"warning: {err} (before step {step})\n")
            ev = eng.events()
        if ef is None:
            return
        for row in io.events_rows(ev):
            ef.write(row + "\n")
        ef.flush()

    logmsg(f"oercel_mi: N={p.N} L={p.L:g} dt={p.dt:g} gamma={p.gamma:g} sigma={p.sigma:g} steps->{p.steps}\n")
    if not restarting:
        st = eng.stats()
        if p.stats_every > 0:
            row = io.stats_row(st)
            sf.write("0,0," + row.split(",", 2)[2] + "\n")
        if p.snap_every > 0:
            snapshot(0)
    elif p.snap_every > 0 and step % p.snap_every == 0 and not os.path.exists(
            os.path.join(outdir, f"snap_{step:010d}.bin")):
        snapshot(step)

    step0 = step
    t_loop0 = time.perf_counter()
    t_last, s_last = t_loop0, step
    stopped = None
    logmsg(f"setup {t_loop0 - t_start:.3f} s\n")
    while step < p.steps:
        nxt = _next_output(step, p.steps, p)
        eng.step(nxt - step)
        step = nxt
        if p.stats_every > 0 and step % p.stats_every == 0:
            st = eng.stats()
            if not math.isfinite(st["T_kin"]):
                raise RuntimeError(f"non-finite temperature at step {step}")
            sf.write(io.stats_row(st) + "\n")
        if p.snap_every > 0 and step % p.snap_every == 0:
            snapshot(step)
        now = time.perf_counter()
        if _stop[0]:
            stopped = "signal"
        elif p.walltime > 0 and now - t_start > p.walltime:
            stopped = "walltime"
        flush_events()
        if (p.ckpt_every > 0 and step % p.ckpt_every == 0) or stopped or step >= p.steps:
            sf.flush()
            io.save_checkpoint(eng, os.path.join(outdir, "ckpt.bin"),
                               keep_path=os.path.join(outdir, f"ckpt_{step:010d}.bin") if p.ckpt_keep else None)
        if p.log_every > 0 and step % p.log_every == 0:
            logmsg(f"step {step} t={p.t_anchor + (step - p.step_anchor) * p.dt:.2f} "
                   f"{1e3 * (now - t_last) / max(step - s_last, 1):.3f} ms/step\n")
            t_last, s_last = now, step
        if stopped:
            logmsg(f"stopping at step {step} ({stopped}), checkpoint written\n")
            break
    flush_events()
    sf.close()
    if ef is not None:
        ef.close()
    n = step - step0
    if n > 0:
        logmsg(f"timing: {n} steps, {1e3 * (time.perf_counter() - t_loop0) / n:.4f} ms/step\n")
    logmsg(f"done at step {step}{' (stopped early; restart to continue)' if stopped else ''}\n")
    eng.close()
    return step


def main(argv=None):
    argv = sys.argv[1:] if argv is None else list(argv)
    if len(argv) < 2 or any(a in ("-h", "--help") for a in argv[:2]):
        print(USAGE, file=sys.stderr)
        return 2
    try:
        run(argv[0], argv[1], argv[2:])
    except ParamError as e:
        print(f"error: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
