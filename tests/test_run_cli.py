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
