"""kinetic_T on caller-held device arrays (oc_kinetic_dev) through the C ABI:
against an FP64 numpy restatement of kinetic_T (cel/engine/oercel.c:2058-2071)
and against an engine's own reduction of the same state."""
import ctypes as C
import importlib
import sys

import numpy as np
import pytest

from conftest import mkparams

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def kinetic_np(vx, vy, fx, fy, dt):
    ux, uy = vx + 0.5 * dt * fx, vy + 0.5 * dt * fy
    return (ux * ux + uy * uy).sum() / (2.0 * len(vx) - 2.0), ux.sum(), uy.sum()


def dev_kinetic(lib, arrs, dt, torch):
    n = arrs[0].numel()
    out = torch.zeros(3, dtype=torch.float64, device="cuda")
    nw = lib.oc_kinetic_dev_work(n)
    work = torch.empty(max(nw // 8, 1), dtype=torch.float64, device="cuda")
    rc = lib.oc_kinetic_dev(*[C.c_void_p(a.data_ptr()) for a in arrs], n, dt, C.c_void_p(out.data_ptr()),
                            C.c_void_p(work.data_ptr()), nw, None)
    assert rc == 0, lib.oc_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("n", [2, 4095, 4096, 4097, 1000003])
def test_kinetic_dev_vs_numpy(n):
    torch = pytest.importorskip("torch")
    abi = importlib.import_module("welcome-youtube_amd.abi")
    lib = abi.load_library()
    say(f"case: oc_kinetic_dev n={n}")
    rng = np.random.default_rng(n)
    h = [rng.standard_normal(n), rng.standard_normal(n), 30 * rng.standard_normal(n), 30 * rng.standard_normal(n)]
    d = [torch.from_numpy(a).cuda() for a in h]
    got = dev_kinetic(lib, d, 0.01, torch)
    ref = kinetic_np(*h, 0.01)
    assert abs(got[0] - ref[0]) <= 1e-13 * ref[0]
    assert abs(got[1] - ref[1]) <= 1e-12 * np.abs(h[0]).sum() and abs(got[2] - ref[2]) <= 1e-12 * np.abs(h[1]).sum()
    again = dev_kinetic(lib, d, 0.01, torch)
    assert np.array_equal(got, again)  # fixed partition: the same bits


def test_kinetic_dev_vs_engine_stats(eng):
    torch = pytest.importorskip("torch")
    abi = importlib.import_module("welcome-youtube_amd.abi")
    lib = abi.load_library()
    say("case: oc_kinetic_dev vs engine stats")
    p = mkparams(24.0, f_lip=0.2, seed=5)
    e = eng.Engine(p)
    e.start()
    e.step(50)
    st = e.stats()
    s = e.get_state()
    e.close()
    d = [torch.from_numpy(np.ascontiguousarray(s[k])).cuda() for k in ("vx", "vy", "fx", "fy")]
    got = dev_kinetic(lib, d, p.dt, torch)
    assert abs(got[0] - st["T_kin"]) < 1e-12 * st["T_kin"]
    assert abs(got[1] - st["px"]) < 1e-10 and abs(got[2] - st["py"]) < 1e-10
