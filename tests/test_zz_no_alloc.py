"""The allocation-free kinetic_T entry (oc_kinetic_dev_noalloc, partials in the
library's fixed device scratch) through the C ABI: the same bits as
oc_kinetic_dev with a caller workspace, on two calls in a row, the second
with fewer beads; sizes past its limit are refused."""
import ctypes as C
import importlib
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_noalloc_equals_workspace_entry():
    torch = pytest.importorskip("torch")
    abi = importlib.import_module("welcome-youtube_amd.abi")
    lib = abi.load_library()
    rng = np.random.default_rng(7)
    for n in (3_000_017, 70_001):  # the second call smaller: stale partials of the first must not leak in
        say(f"case: oc_kinetic_dev_noalloc n={n}")
        h = [rng.standard_normal(n), rng.standard_normal(n), 25 * rng.standard_normal(n), 25 * rng.standard_normal(n)]
        d = [torch.from_numpy(a).cuda() for a in h]
        ptrs = [C.c_void_p(a.data_ptr()) for a in d]
        out_a = torch.zeros(3, dtype=torch.float64, device="cuda")
        out_b = torch.zeros(3, dtype=torch.float64, device="cuda")
        rc = lib.oc_kinetic_dev_noalloc(*ptrs, n, 0.01, C.c_void_p(out_a.data_ptr()), None)
        assert rc == 0, lib.oc_last_error()
        nw = lib.oc_kinetic_dev_work(n)
        work = torch.empty(nw // 8, dtype=torch.float64, device="cuda")
        rc = lib.oc_kinetic_dev(*ptrs, n, 0.01, C.c_void_p(out_b.data_ptr()), C.c_void_p(work.data_ptr()), nw, None)
        assert rc == 0, lib.oc_last_error()
        torch.cuda.synchronize()
        a, b = out_a.cpu().numpy(), out_b.cpu().numpy()
        assert np.array_equal(a, b), (a, b)
        ux, uy = h[0] + 0.005 * h[2], h[1] + 0.005 * h[3]
        assert abs(a[0] - (ux * ux + uy * uy).sum() / (2.0 * n - 2.0)) <= 1e-13 * a[0]
    # beyond the fixed scratch: refused before any launch
    one = torch.zeros(1, dtype=torch.float64, device="cuda")
    p = C.c_void_p(one.data_ptr())
    assert lib.oc_kinetic_dev_noalloc(p, p, p, p, 16384 * 4096 + 1, 0.01, p, None) == abi.OC_EINVAL
    assert b"67,108,864" in lib.oc_last_error()
