"""The precursor chemistry's rates on the oracle (CPU): per check,
conversions ~ Binomial(eligible P, p_conv), breaks ~ Binomial(bonds, p_break),
bonds of free-T pairs that are each other's only candidate ~ Binomial(pairs,
p_form).  The device runs the same check in tests/test_gpu_chemrates.py."""
from conftest import mkparams
from test_gpu_chemrates import PRECURSOR, check_binomial, precursor_rate_counts


def test_precursor_chemistry_rates_oracle(oracle):
    p = mkparams(34.0, nsub=4, f_lip=0.3, c_N=0.25, seed=35, **PRECURSOR)
    cp = p.to_c()
    o = oracle.System(cp, p.N)
    o.start(oracle.soup(cp, p.N))
    o.step(20)
    o.events()

    def step1():
        o.step(1)
        return o.step_count

    elig, hits = precursor_rate_counts(o.get_state, step1, o.events, p, 60)
    check_binomial(elig, hits, nmin=150)
