"""Host logic of the run driver (no GPU): usage, the output cadence and the
CSV truncation a restart applies (truncate_csv_after,
cel/engine/oercel.c:2013-2051)."""
import importlib

from conftest import params_mod

run = importlib.import_module("welcome-youtube_amd.run")


def test_usage_and_bad_override(capsys):
    assert run.main([]) == 2
    assert "usage" in capsys.readouterr().err


def test_truncate_csv_after(tmp_path):
    p = tmp_path / "stats.csv"
    p.write_text("step,time\n0,0\n10,0.1\n20,0.2\n30,0.3\n")
    run.truncate_csv_after(str(p), 20)
    assert p.read_text() == "step,time\n0,0\n10,0.1\n20,0.2\n"
    run.truncate_csv_after(str(p), 5)
    assert p.read_text() == "step,time\n0,0\n"


def test_next_output_stops_at_every_cadence():
    p = params_mod().load(None, ["L=16", "chem=0", "stats_every=30", "snap_every=70", "ckpt_every=0", "log_every=0",
                                 "steps=250"])
    seq, s = [], 0
    while s < p.steps:
        s = run._next_output(s, p.steps, p)
        seq.append(s)
    assert seq == [30, 60, 70, 90, 120, 140, 150, 180, 210, 240, 250]
    assert run._next_output(0, 10 ** 6, params_mod().load(None, ["L=16", "chem=0", "stats_every=0", "snap_every=0",
                                                                 "ckpt_every=0", "log_every=0"])) == run.CHUNK


def test_events_rows_follow_the_reference_format():
    """events.csv rows as the reference's flush_events writes them
    (cel/engine/oercel.c:2036-2047): variant a / b for catalyst events, "-"
    for decay, one feed row per check with the count, no position for the
    precursor chemistry's rows."""
    io = importlib.import_module("welcome-youtube_amd.io")
    abi = importlib.import_module("welcome-youtube_amd.abi")
    import numpy as np
    ev = np.zeros(8, dtype=abi.event_dtype())
    rows = [(10, 6, 5, 1, 1.23456, 2.5), (10, 7, 9, 0, 1.23456, 2.5), (10, 8, 11, 1, 3.0, 4.0),
            (10, 9, 12, 0, 0.5, 0.25), (10, 3, 20, 20, 7.0, 8.0), (10, 5, 30, 30, 1.0, 1.0),
            (10, 5, 31, 31, 2.0, 2.0), (20, 1, 40, 41, 0.0, 0.0)]
    for q, (st, k, a, b, x, y) in enumerate(rows):
        ev[q]["step"], ev[q]["kind"], ev[q]["gid_a"], ev[q]["gid_b"], ev[q]["x"], ev[q]["y"] = st, k, a, b, x, y
    assert io.events_rows(ev) == [
        "10,syn,5,1.2346,2.5000,b", "10,rep,9,1.2346,2.5000,a", "10,mut,11,3.0000,4.0000,b",
        "10,dK,12,0.5000,0.2500,a", "10,dlip,20,7.0000,8.0000,-", "10,feed,2,nan,nan,-", "20,form,40,nan,nan,41"]
