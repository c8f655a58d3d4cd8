import importlib
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def pkg():
    return importlib.import_module("welcome-youtube_amd")


def params_mod():
    return importlib.import_module("welcome-youtube_amd.params")


def engine_mod():
    return importlib.import_module("welcome-youtube_amd.engine")


def oracle_mod():
    import oracle
    return oracle


def mkparams(L, nsub=4, **kw):
    """Parameters for a test system: the BASELINE physics defaults (dt 0.01,
    H-T-T lipids, harmonic bonds only, no warm-up), overridden by kw."""
    P = params_mod().Params
    base = dict(L=L, nsub=nsub, dt=0.01, chem=0, warmup=0, k_angle=0, lip_nH=1, lip_nT=2, f_lip=0.1, f_Ka=0, f_Kb=0,
                c_N=0, a_W_H=25, a_H_T=80, a_W_T=80, a_N_T=50, a_N_W=25, a_N_H=25)
    base.update(kw)
    p = P()
    p.update({k: str(v) for k, v in base.items()})
    return p.finalize()


@pytest.fixture(scope="session")
def oracle():
    return oracle_mod()


@pytest.fixture(scope="session")
def eng():
    return engine_mod()


@pytest.fixture(scope="session")
def rng():
    return np.random.default_rng(12345)
