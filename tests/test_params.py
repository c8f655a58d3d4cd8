"""The parameter API: the reference's key table and set_param / read_ini /
finalize_params rules (cel/engine/oercel.c:237-619, 2141-2159).  CPU only."""
import pytest

from conftest import params_mod

pm = params_mod()


def test_defaults_and_derived():
    p = pm.Params().finalize()
    assert p.L == 64.0 and p.dt == 0.03 and p.lambda_ == 0.65 and p.k_angle == 10.0 and p.nsub == 4
    assert p.seed == 1 and p.steps == 10000 and p.init == "soup" and p.restart == ""
    assert abs(p.sigma - 3.0) < 1e-15
    assert p.N == 12288  # lround(3 * 64^2)


def test_repulsion_keys_symmetric_and_generic():
    p = pm.Params()
    p.set("a_W_T", "70")
    assert p.a[0][2] == p.a[2][0] == 70.0
    # a_K_X sets Ka and Kb; the generic key is applied first whatever its position
    p.update([("a_Ka_T", "99"), ("a_K_T", "11")])
    assert p.a[4][2] == 99.0 and p.a[5][2] == 11.0 and p.a[2][5] == 11.0
    with pytest.raises(pm.ParamError):
        p.set("a_W", "1")
    with pytest.raises(pm.ParamError):
        p.set("a_W_Q", "1")
    with pytest.raises(pm.ParamError):
        p.set("a_W_T", "abc")
    p.set("a_W_T", "12.5xyz")  # strtod prefix, as the reference accepts
    assert p.a[0][2] == 12.5


def test_typed_values():
    p = pm.Params()
    p.set("steps", "1e6")
    assert p.steps == 1000000
    p.set("steps", "2.5")
    assert p.steps == 3  # llround
    p.set("seed", "0x10")
    assert p.seed == 16
    p.set("threads", " 7 ")
    assert p.threads == 7
    for k, v in (("threads", "1.5"), ("steps", "abc"), ("dt", "1e-3x"), ("seed", "")):
        with pytest.raises(pm.ParamError):
            p.set(k, v)
    with pytest.raises(pm.ParamError):
        p.set("lip_arch", "x" * 16)
    with pytest.raises(pm.ParamError, match="unknown"):
        p.set("nonsense", "1")
    with pytest.raises(pm.ParamError, match="removed"):
        p.set("feed_every", "1")


@pytest.mark.parametrize("kv,msg", [
    ({"L": "7.9"}, "L must be"), ({"nsub": "9"}, "nsub"), ({"dt": "0"}, "dt"), ({"kT_amp": "1"}, "kT_amp"),
    ({"init": "seeded"}, "seeded"), ({"feed": "1"}, "flood"), ({"feed_cap": "1"}, "feed_cap"),
    ({"lip_arch": "star"}, "lip_arch must be"), ({"lip_nT": "0"}, "lipid length"), ({"init": "restart"}, "restart"),
    ({"p_conv": "2"}, "p_conv"), ({"r_form": "1.5"}, "r_conv and r_form"),
])
def test_finalize_errors(kv, msg):
    with pytest.raises(pm.ParamError, match=msg):
        pm.load(overrides=kv)


def test_finalize_clamps():
    p = pm.load(overrides={"threads": "0", "n_chem": "0", "warmup": "-5", "vent_x": "-1"})
    assert p.threads == 1 and p.n_chem == 1 and p.warmup == 0 and p.vent_x == 32.0


def test_ini_and_overrides(tmp_path):
    ini = tmp_path / "a.ini"
    ini.write_text("# comment\nL = 20   # trailing\n\n a_W_T=60\ndt=0.01\n")
    p = pm.load(str(ini), ["dt=0.02", "a_K_W=5"])
    assert p.L == 20.0 and p.dt == 0.02 and p.a[0][2] == 60.0 and p.a[4][0] == 5.0
    bad = tmp_path / "b.ini"
    bad.write_text("L 20\n")
    with pytest.raises(pm.ParamError, match="expected key = value"):
        pm.load(str(bad))
    with pytest.raises(pm.ParamError):
        pm.load(overrides=["Lx20"])


def test_engine_takes_the_reference_defaults():
    """The reference's default parameters (chem = 1 with catalysts, feed cap)
    map onto the C parameters without a refusal."""
    p = pm.load().check_engine()
    c = p.to_c()
    assert c.rchem == 1 and c.f_Ka == 0.005 and c.p_syn_b == 0.06 and c.feed_cap == 0.12
    assert c.r_cat == 1.2 and c.syn_tries == 8 and c.syn_relax == 20 and c.syn_nN == 0 and c.mu == 0.02
    pm.load(overrides={"chem": "0", "pchem": "1"}).check_engine()


@pytest.mark.parametrize("N", [12000, 65536, 262144, 1048576, 16777216])
def test_box_side_for_exact_N(N):
    L = pm.L_for_N(N)
    assert pm._llround(3.0 * L * L) == N
    assert abs(L - (N / 3.0) ** 0.5) < 1e-9


def test_to_c_roundtrip():
    p = pm.load(overrides={"L": "33.5", "nsub": "3", "a_W_T": "77", "pchem": "1", "p_form": "0.2", "seed": "42"})
    c = p.to_c()
    assert c.L == 33.5 and c.nsub == 3 and c.a[0][2] == 77.0 and c.a[2][0] == 77.0 and c.chem == 1
    assert c.p_form == 0.2 and c.seed == 42 and c.a[7][7] == 0.0


def test_ini_text_reloads():
    p = pm.load(overrides={"L": "21.25", "a_H_T": "66", "pchem": "1"})
    kv = [ln.split("=", 1) for ln in p.ini_text().splitlines() if ln and not ln.startswith("#")]
    q = pm.Params()
    q.update([(k.strip(), v.strip()) for k, v in kv])
    q.finalize()
    assert q.L == p.L and q.a == p.a and q.pchem == 1
