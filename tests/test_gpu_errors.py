"""Error paths of the C ABI on the GPU: each returns an error code with a
message (raised as OcError) and leaves the engine usable."""
import importlib

import numpy as np
import pytest

from conftest import mkparams

pytestmark = pytest.mark.gpu


def test_error_paths_leave_the_engine_usable(eng):
    abi = importlib.import_module("welcome-youtube_amd.abi")
    p = mkparams(12.0, f_lip=0.3, seed=2)
    e = eng.Engine(p)
    with pytest.raises(abi.OcError, match="no state"):
        e.step(1)
    e.start()
    with pytest.raises(abi.OcError, match="no forward window"):
        e.window_rewind()
    s = e.get_state()
    bad = np.arange(p.N, dtype=np.uint32)
    bad[1] = bad[0]
    with pytest.raises(abi.OcError, match="permutation"):
        e.set_state(s, step=0, slot_gid=bad)
    with pytest.raises(ValueError):
        e.set_state(s, step=0, slot_gid=bad[:-1])
    with pytest.raises(abi.OcError, match="max bond"):
        e.max_bond(0, set_to=-1.0)
    with pytest.raises(abi.OcError, match="pair variant"):
        e._chk(e.lib.oc_set_pair_variant(e.h, 4))
    with pytest.raises(abi.OcError):
        e.set_tuning(pair_block=96)
    e.set_state(s, step=0, slot_gid=e.cells()[1])
    e.step(5)
    st = e.stats()
    assert np.isfinite(st["T_kin"]) and e.step_count() == 5
    e.close()


def test_reference_chemistry_limits(eng):
    """The window gradient refuses the reference chemistry (not
    differentiated); a step that moves a bead by the chemistry margin 0.5
    before a check is reported (the reactions' windows rely on the last
    sort's cell index), and the engine keeps working after a state reload."""
    abi = importlib.import_module("welcome-youtube_amd.abi")
    oc = importlib.import_module("welcome-youtube_amd")
    p = mkparams(16.0, f_lip=0.1, c_N=0.3, f_Ka=0.03, chem=1, n_chem=1, p_syn_a=0.5, p_feed=0.5, p_nd=0.5, p_ld=0.5,
                 seed=4)
    e = eng.Engine(p)
    e.start()
    with pytest.raises(abi.OcError, match="reference chemistry"):
        e.window_forward(10, oc.abi.rdf_spec(2, 2, 16, 3.0))
    s = e.get_state()
    fast = dict(s, vx=s["vx"].copy())
    fast["vx"][5] = 60.0  # dt 0.01: a 0.6 move in one step
    e.set_state(fast, step=0)
    with pytest.raises(abi.OcError, match="moved 0.5 or more"):
        e.step(1, graph=False)
    # the refused check changed no type, chain or bond (no reaction, decay or feed)
    after = e.get_state()
    for k in ("type", "mpos", "nxt", "bond"):
        assert np.array_equal(after[k], fast[k]), k
    e.set_state(s, step=0)
    e.step(3)
    assert np.isfinite(e.stats()["T_kin"])
    e.close()
