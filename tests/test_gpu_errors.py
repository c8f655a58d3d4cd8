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
        e._chk(e.lib.oc_set_pair_variant(e.h, 3))
    with pytest.raises(abi.OcError):
        e.set_tuning(pair_block=96)
    e.set_state(s, step=0, slot_gid=e.cells()[1])
    e.step(5)
    st = e.stats()
    assert np.isfinite(st["T_kin"]) and e.step_count() == 5
    e.close()
