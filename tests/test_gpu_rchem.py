"""The reference chemistry without catalysts on the GPU: its feed known answer
(1138 conversions over 600 steps) and, from identical states, decay and feed
decisions (with the precursor chemistry in the same check) bit-exact against
the oracle."""
import sys

import numpy as np
import pytest

from conftest import mkparams

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_feed_known_answer_on_the_gpu(eng):
    """cel/engine/tests/validate.sh:133-134: uniform feed on water, L = 48,
    seed 7, p_feed 0.003, a check every 10 steps, 600 steps -> 1138."""
    p = mkparams(48.0, f_lip=0.0, seed=7, chem=1, feed="uniform", p_feed=0.003, feed_cap=0, n_chem=10, p_nd=0,
                 p_ld=0, dt=0.02)
    p.check_engine()
    e = eng.Engine(p)
    e.start()
    e.step(600)
    st = e.get_state()
    ev = e.events()
    n = int((st["type"] == 3).sum())
    say(f"feed conversions after 600 steps: {n}")
    assert n == 1138 and int((ev["kind"] == 5).sum()) == 1138
    e.close()


@pytest.mark.parametrize("feed", ["uniform", "vent"])
def test_decay_and_feed_bit_exact(oracle, eng, feed):
    say(f"case: decay + feed ({feed}) vs oracle")
    p = mkparams(20.0, f_lip=0.3, c_N=0.2, seed=9, chem=1, n_chem=1, p_ld=0.05, p_nd=0.05, feed=feed,
                 vent_shape="disk", vent_r=6.0, p_feed=0.05, feed_cap=0, pchem=1, p_conv=0.2, p_form=0.2,
                 p_break=0.05)
    p.check_engine()
    o = oracle.System(p.to_c(), p.N)
    o.start(oracle.soup(p.to_c(), p.N))
    o.step(8)
    o.events()
    st = o.get_state()
    e = eng.Engine(p)
    e.set_state(st, step=8)
    o = oracle.System(p.to_c(), p.N)
    o.set_state(st, 8)
    e.step(1, graph=False)  # one step: the integration of the loaded state is bit-exact
    o.step(1)
    a, b = e.get_state(), o.get_state()
    for k in ("x", "y", "vx", "vy", "type", "mpos", "nxt", "bond"):
        assert np.array_equal(a[k], b[k]), k
    ed, eo = e.events(), o.events()
    kinds = np.bincount(eo["kind"], minlength=6)
    say(f"events by kind (conv, form, break, dlip, dN, feed): {kinds.tolist()}")
    assert kinds[3] > 0 and kinds[4] > 0 and kinds[5] > 0
    for k in ("step", "kind", "gid_a", "gid_b", "x", "y"):
        assert np.array_equal(ed[k], eo[k]), k
    e.close()
