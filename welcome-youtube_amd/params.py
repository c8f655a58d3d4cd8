"""Parameter API of the engine: the reference's typed key table, its ini and
key=value rules, and its checks, restated in Python.

Reference: the Params struct and PDEFS table (cel/engine/oercel.c:178-299), the
default repulsion matrix (:305-322), set_param (:338-395), is_generic_akey
(:397-406), read_ini (:408-434), the two-pass key order of main (:2156-2158),
finalize_params (:578-619) and N = lround(rho L^2) (:2188).

Same key names, types, defaults and error behaviour: an unknown key, a bad
value or a removed key raises ParamError (the reference prints the message and
exits).  Every key of the reference is accepted and used: the chemistry
(chem = 1; catalysts, decay, feed and its cap), IO cadences, lipid
architectures.  Lipid templates (lip_arch linear, pi, y, bola and
their keys) are the reference's mol_build.

Precursor-chemistry keys (no counterpart in the reference): pchem, p_conv,
r_conv, p_form, r_form, p_break.
"""
import math
import re

from . import abi

TYPE_NAMES = abi.TYPE_NAMES


class ParamError(ValueError):
    pass


# (key, type, default) in the reference's table order; types: int, long, dbl, str, u64.
# str entries carry their C buffer size (value length must stay below it).
PDEFS = [
    ("L", "dbl", "64"), ("rho", "dbl", "3"), ("dt", "dbl", "0.03"), ("lambda", "dbl", "0.65"),
    ("gamma", "dbl", "4.5"), ("kT", "dbl", "1"), ("k_bond", "dbl", "100"), ("r0", "dbl", "0.5"),
    ("k_angle", "dbl", "10"), ("lip_nH", "int", "1"), ("lip_nT", "int", "3"), ("lip_arch", ("str", 16), "linear"),
    ("r0_head", "dbl", "0.5"), ("k_head", "dbl", "0"), ("cos_head", "dbl", "0"), ("lip_tie", "int", "0"),
    ("syn_nN", "int", "0"), ("steps", "long", "10000"), ("snap_every", "long", "1000"),
    ("stats_every", "long", "100"), ("ckpt_every", "long", "10000"), ("ckpt_keep", "int", "0"),
    ("log_every", "long", "1000"), ("warmup", "long", "1000"), ("threads", "int", "2"), ("nsub", "int", "4"),
    ("seed", "u64", "1"), ("init", ("str", 64), "soup"), ("restart", ("str", 1024), ""),
    ("restart_reseed", "int", "0"), ("walltime", "dbl", "0"), ("chem", "int", "1"), ("n_chem", "long", "10"),
    ("r_cat", "dbl", "1.2"), ("syn_tries", "int", "8"), ("syn_relax", "int", "20"), ("p_syn_a", "dbl", "0.02"),
    ("p_syn_b", "dbl", "0.06"), ("p_rep_a", "dbl", "0.001"), ("p_rep_b", "dbl", "0.001"), ("mu", "dbl", "0.02"),
    ("p_ld", "dbl", "2e-4"), ("p_kd", "dbl", "5e-4"), ("p_nd", "dbl", "1e-4"), ("feed", ("str", 16), "uniform"),
    ("p_feed", "dbl", "1e-3"), ("feed_cap", "dbl", "0.12"), ("vent_shape", ("str", 16), "disk"),
    ("vent_x", "dbl", "-1"), ("vent_y", "dbl", "-1"), ("vent_r", "dbl", "5"), ("vent_w", "dbl", "4"),
    ("events", "int", "1"), ("stir_A", "dbl", "0"), ("stir_n", "int", "1"), ("kT_amp", "dbl", "0"),
    ("kT_period", "dbl", "1000"), ("f_lip", "dbl", "0.10"), ("f_Ka", "dbl", "0.005"), ("f_Kb", "dbl", "0"),
    ("c_N", "dbl", "0.12"),
]
NREF_KEYS = len(PDEFS)

# precursor chemistry (own keys, appended after the reference's table)
PCHEM_DEFS = [
    ("pchem", "int", "0"), ("p_conv", "dbl", "0.005"), ("r_conv", "dbl", "0.7"), ("p_form", "dbl", "0.01"),
    ("r_form", "dbl", "0.6"), ("p_break", "dbl", "0.001"),
]
ALL_DEFS = PDEFS + PCHEM_DEFS
_DEF = {k: (t, d) for k, t, d in ALL_DEFS}

REMOVED_KEYS = ("seed_R", "seed_spacing", "seed_cN_in", "seed_nKa", "feed_every", "feed_grid", "feed_tmin")

_FLOAT_RE = re.compile(r"\s*[+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan)", re.I)
_HEX_RE = re.compile(r"\s*[+-]?0[xX](?:[0-9a-fA-F]+\.?[0-9a-fA-F]*|\.[0-9a-fA-F]+)(?:[pP][+-]?\d+)?")


def _strtod(s):
    """C strtod: (value, number of characters consumed); consumed 0 = no number."""
    m = _HEX_RE.match(s)
    if m:
        txt = m.group(0).strip()
        if "p" not in txt.lower():
            txt += "p0"
        return float.fromhex(txt), m.end()
    m = _FLOAT_RE.match(s)
    if not m:
        return 0.0, 0
    return float(m.group(0)), m.end()


def _strtol(s, base):
    m = re.match(r"\s*[+-]?(0[xX][0-9a-fA-F]+|0[0-7]*|[1-9]\d*)" if base == 0 else r"\s*[+-]?\d+", s)
    if not m:
        return 0, 0
    txt = m.group(0).strip()
    if base == 0:
        sign = -1 if txt.startswith("-") else 1
        body = txt.lstrip("+-")
        if body.lower().startswith("0x"):
            v = int(body, 16)
        elif len(body) > 1 and body.startswith("0"):
            v = int(body, 8)
        else:
            v = int(body)
        return sign * v, m.end()
    return int(txt), m.end()


def _llround(v):
    """C llround: halfway cases away from zero."""
    return int(math.floor(v + 0.5)) if v >= 0 else -int(math.floor(-v + 0.5))


def default_amatrix():
    """cel/engine/oercel.c:305-322."""
    a = [[25.0] * 7 for _ in range(7)]

    def seta(i, j, v):
        a[i][j] = a[j][i] = v

    W, H, T, N, KA, KB = 0, 1, 2, 3, 4, 5
    seta(W, H, 20.0)
    seta(W, T, 80.0)
    seta(H, T, 50.0)
    seta(N, W, 25.0)
    seta(N, H, 25.0)
    seta(N, T, 40.0)
    for k in (KA, KB):
        seta(k, W, 25.0)
        seta(k, H, 30.0)
        seta(k, T, 120.0)
        seta(k, N, 25.0)
    return a


def _type_from_name(s):
    if s in TYPE_NAMES:
        return TYPE_NAMES.index(s)
    if s == "K":
        return -2
    return -1


def is_generic_akey(key):
    """cel/engine/oercel.c:397-406: a_K_X or a_X_K."""
    if not key.startswith("a_"):
        return False
    rest = key[2:]
    if "_" not in rest:
        return False
    t1, t2 = rest.split("_", 1)
    return t1 == "K" or t2 == "K"


class Params:
    """The engine's parameters: attribute access by reference key name
    (``p.L``, ``p.lambda_`` for the Python keyword), the 7x7 matrix ``p.a``.
    """

    def __init__(self, **overrides):
        self.__dict__["_v"] = {}
        self.__dict__["a"] = default_amatrix()
        for k, t, d in ALL_DEFS:
            self.set(k, d)
        self.__dict__["sigma"] = None
        self.__dict__["warm_end"] = 0
        self.__dict__["t_anchor"] = 0.0
        self.__dict__["step_anchor"] = 0
        if overrides:
            self.update(overrides)

    # attribute access
    def __getattr__(self, k):
        if k == "lambda_":
            k = "lambda"
        v = self.__dict__["_v"]
        if k in v:
            return v[k]
        raise AttributeError(k)

    def __setattr__(self, k, val):
        if k == "lambda_":
            k = "lambda"
        if k in self.__dict__["_v"]:
            self.__dict__["_v"][k] = val
        else:
            self.__dict__[k] = val

    def copy(self):
        q = Params.__new__(Params)
        q.__dict__.update({k: (v.copy() if isinstance(v, dict) else v) for k, v in self.__dict__.items()})
        q.__dict__["a"] = [row[:] for row in self.a]
        return q

    def set(self, key, val):
        """set_param, cel/engine/oercel.c:338-395."""
        val = str(val)
        if key.startswith("a_"):
            rest = key[2:]
            if "_" not in rest:
                raise ParamError(f"bad repulsion key '{key}' (use a_X_Y)")
            n1, n2 = rest.split("_", 1)
            t1, t2 = _type_from_name(n1), _type_from_name(n2)
            if t1 == -1 or t2 == -1:
                raise ParamError(f"bad repulsion key '{key}'")
            v, n = _strtod(val)
            if n == 0:
                raise ParamError(f"bad value for {key}: '{val}'")
            l1 = [4, 5] if t1 == -2 else [t1]
            l2 = [4, 5] if t2 == -2 else [t2]
            for i in l1:
                for j in l2:
                    self.a[i][j] = self.a[j][i] = v
            return
        if key in _DEF:
            t, _ = _DEF[key]
            vals = self.__dict__["_v"]
            if t == "int":
                v, n = _strtol(val, 10)
                if n == 0 or val[n:].strip():
                    raise ParamError(f"bad integer for {key}: '{val}'")
                vals[key] = int(v)
            elif t == "long":
                v, n = _strtod(val)
                if n == 0 or val[n:].strip():
                    raise ParamError(f"bad integer for {key}: '{val}'")
                vals[key] = _llround(v)
            elif t == "u64":
                v, n = _strtol(val, 0)
                if n == 0 or val[n:].strip():
                    raise ParamError(f"bad integer for {key}: '{val}'")
                vals[key] = v % (1 << 64)
            elif t == "dbl":
                v, n = _strtod(val)
                if n == 0 or val[n:].strip():
                    raise ParamError(f"bad number for {key}: '{val}'")
                vals[key] = float(v)
            else:
                if len(val) >= t[1]:
                    raise ParamError(f"value too long for {key}")
                vals[key] = val
            return
        if key in REMOVED_KEYS:
            raise ParamError(f"parameter '{key}' was removed from the engine")
        raise ParamError(f"unknown parameter '{key}'")

    def update(self, kv):
        """Apply key/value pairs in the reference's two-pass order: generic
        a_K_X keys first, then every other key in order (cel/engine/oercel.c:2156-2158)."""
        items = list(kv.items()) if isinstance(kv, dict) else list(kv)
        for pass_ in (0, 1):
            for k, v in items:
                if is_generic_akey(k) == (pass_ == 0):
                    self.set(k, v)
        return self

    # ------------------------------------------------------------------ finalize
    def finalize(self):
        """finalize_params, cel/engine/oercel.c:578-619 (checks and clamps), for
        the linear lipid architecture this engine implements."""
        v = self.__dict__["_v"]
        self.__dict__["sigma"] = math.sqrt(2.0 * v["gamma"] * v["kT"])
        if v["lip_arch"] not in ARCHS:
            raise ParamError("lip_arch must be linear, pi, y or bola")
        if v["lip_arch"] == "linear":
            mlen = v["lip_nH"] + v["lip_nT"]
            if v["lip_nH"] < 1 or v["lip_nT"] < 1 or mlen > abi.MAXMLEN:
                raise ParamError(f"lipid length must be 2..{abi.MAXMLEN} with >=1 head and tail")
        else:
            if v["lip_nT"] < 1:
                raise ParamError("lip_nT must be >= 1")
            mlen = mol_template(self).mlen
            if mlen > abi.MAXMLEN:
                raise ParamError(f"{v['lip_arch']} lipid with lip_nT={v['lip_nT']} has {mlen} beads "
                                 f"(max {abi.MAXMLEN})")
        if v["syn_nN"] < 0 or v["syn_nN"] > mlen:
            raise ParamError(f"syn_nN must be 0..{mlen}")
        if v["r0_head"] <= 0:
            raise ParamError("r0_head must be > 0")
        if v["L"] < 8:
            raise ParamError("L must be >= 8")
        if v["dt"] <= 0:
            raise ParamError("dt must be > 0")
        if v["threads"] < 1:
            v["threads"] = 1
        if v["nsub"] < 1 or v["nsub"] > 8:
            raise ParamError("nsub must be 1..8")
        if v["n_chem"] < 1:
            v["n_chem"] = 1
        if v["syn_tries"] < 1:
            v["syn_tries"] = 1
        if v["warmup"] < 0:
            v["warmup"] = 0
        if v["init"] == "seeded":
            raise ParamError("init = seeded was removed: the only start state is a random soup or a checkpoint of one")
        if v["init"] not in ("soup", "restart"):
            raise ParamError("init must be soup or restart")
        if v["init"] == "restart" and not v["restart"]:
            raise ParamError("init = restart needs restart = <checkpoint>")
        if v["feed"] in ("none", "0"):
            v["feed"] = "none"
        elif v["feed"] == "1":
            raise ParamError("feed = 1 (flood fill) was removed; use feed = uniform, vent or none")
        elif v["feed"] not in ("uniform", "vent"):
            raise ParamError("feed must be none, uniform or vent")
        if v["p_feed"] < 0 or v["p_feed"] > 1:
            raise ParamError("p_feed must be in [0, 1]")
        if v["feed_cap"] < 0 or v["feed_cap"] >= 1:
            raise ParamError("feed_cap must be in [0, 1) (0 = no cap)")
        if v["vent_shape"] not in ("disk", "vstrip", "hstrip"):
            raise ParamError("vent_shape must be disk, vstrip or hstrip")
        if v["vent_x"] < 0:
            v["vent_x"] = 0.5 * v["L"]
        if v["vent_y"] < 0:
            v["vent_y"] = 0.5 * v["L"]
        if v["feed"] == "vent":
            if v["vent_x"] >= v["L"] or v["vent_y"] >= v["L"]:
                raise ParamError("vent_x and vent_y must lie in [0, L)")
            if v["vent_shape"] == "disk" and (v["vent_r"] <= 0 or v["vent_r"] >= 0.5 * v["L"]):
                raise ParamError("vent_r must be in (0, L/2)")
            if v["vent_shape"] != "disk" and (v["vent_w"] <= 0 or v["vent_w"] > v["L"]):
                raise ParamError("vent_w must be in (0, L]")
        if v["stir_n"] < 1:
            raise ParamError("stir_n must be >= 1")
        if v["kT"] <= 0:
            raise ParamError("kT must be > 0")
        if abs(v["kT_amp"]) >= v["kT"]:
            raise ParamError("|kT_amp| must be < kT (the temperature must stay positive)")
        if v["kT_amp"] != 0 and v["kT_period"] <= 0:
            raise ParamError("kT_period must be > 0")
        for k in ("p_conv", "p_form", "p_break"):
            if not 0.0 <= v[k] <= 1.0:
                raise ParamError(f"{k} must be in [0, 1]")
        if not 0.0 < v["r_conv"] <= 1.0 or not 0.0 < v["r_form"] <= 1.0:
            raise ParamError("r_conv and r_form must be in (0, 1] (the cell list covers rc = 1)")
        return self

    def check_engine(self):
        """Features the HIP engine implements (raises for the others; every
        feature of the parameter table is implemented).  The reference
        chemistry (chem = 1) runs in the reference GPU port's form:
        hash-decided catalyst reactions in hash priority order, decay, and the
        food supply with its cap as the smallest hash keys."""
        return self

    @property
    def N(self):
        """cel/engine/oercel.c:2188."""
        return _llround(self.rho * self.L * self.L)

    @property
    def mlen(self):
        return mol_template(self).mlen

    @property
    def nH(self):
        """heads per lipid (lip_nH for linear lipids, 2 / 1 / 3 for pi / y / bola)"""
        return mol_template(self).nH

    def to_c(self):
        """oc_params for the C ABI (call finalize() first)."""
        c = abi.oc_params()
        v = self.__dict__["_v"]
        c.L, c.rho, c.dt, c.lambda_ = v["L"], v["rho"], v["dt"], v["lambda"]
        c.gamma, c.kT = v["gamma"], v["kT"]
        c.k_bond, c.r0, c.k_angle = v["k_bond"], v["r0"], v["k_angle"]
        for i in range(8):
            for j in range(8):
                c.a[i][j] = self.a[i][j] if i < 7 and j < 7 else 0.0
        c.stir_A, c.kT_amp, c.kT_period = v["stir_A"], v["kT_amp"], v["kT_period"]
        c.stir_n, c.nsub, c.lip_nH, c.lip_nT = v["stir_n"], v["nsub"], v["lip_nH"], v["lip_nT"]
        c.warmup, c.warm_end, c.seed = v["warmup"], self.warm_end, v["seed"]
        c.f_lip, c.f_Ka, c.f_Kb, c.c_N = v["f_lip"], v["f_Ka"], v["f_Kb"], v["c_N"]
        c.chem, c.n_chem = v["pchem"], v["n_chem"]
        c.p_conv, c.r_conv, c.p_form, c.r_form, c.p_break = v["p_conv"], v["r_conv"], v["p_form"], v["r_form"], v[
            "p_break"]
        c.t_anchor, c.step_anchor = self.t_anchor, self.step_anchor
        c.lip_arch = ARCHS.index(v["lip_arch"])
        c.lip_tie, c.r0_head, c.k_head, c.cos_head = v["lip_tie"], v["r0_head"], v["k_head"], v["cos_head"]
        c.rchem = 1 if v["chem"] else 0
        c.feed_mode = ("none", "uniform", "vent").index(v["feed"])
        c.p_ld, c.p_nd, c.p_feed = v["p_ld"], v["p_nd"], v["p_feed"]
        c.vent_shape = ("disk", "vstrip", "hstrip").index(v["vent_shape"])
        c.vent_x, c.vent_y, c.vent_r, c.vent_w = v["vent_x"], v["vent_y"], v["vent_r"], v["vent_w"]
        c.p_syn_a, c.p_syn_b, c.p_rep_a, c.p_rep_b = v["p_syn_a"], v["p_syn_b"], v["p_rep_a"], v["p_rep_b"]
        c.mu, c.p_kd, c.r_cat, c.feed_cap = v["mu"], v["p_kd"], v["r_cat"], v["feed_cap"]
        c.syn_tries, c.syn_relax, c.syn_nN = v["syn_tries"], v["syn_relax"], v["syn_nN"]
        return c

    def items(self):
        for k, _, _ in ALL_DEFS:
            yield k, self.__dict__["_v"][k]

    def ini_text(self):
        """params.ini text in the reference's style (cel/engine/oercel.c:626-653): every key, full precision."""
        out = ["# parameters actually used (all keys, defaults filled in)"]
        for k, t, _ in ALL_DEFS:
            val = self.__dict__["_v"][k]
            out.append(f"{k} = {val!r}" if t == "dbl" else f"{k} = {val}")
        out.append("# repulsion matrix a_ij (symmetric)")
        for i in range(7):
            for j in range(i, 7):
                out.append(f"a_{TYPE_NAMES[i]}_{TYPE_NAMES[j]} = {self.a[i][j]!r}")
        return "\n".join(out) + "\n"


ARCHS = ("linear", "pi", "y", "bola")


class MolTemplate:
    """The lipid template of mol_build (cel/engine/oercel.c:479-576): bead
    count, heads, bonds (i, j, rest length) and angle terms (i, j, k with j the
    middle bead, kind 0: U = K (1 - cos), kind 1: U = K (cos - c0)^2), in the
    reference's order, and the soup-init offsets (ix, iy) of each bead from
    bead 0."""

    def __init__(self, arch):
        self.arch, self.mlen, self.nH = arch, 0, 0
        self.bonds, self.angles = [], []
        self.px, self.py = [], []

    def bond(self, i, j, r0):
        self.bonds.append((i, j, r0))

    def angle(self, i, j, k, kind, K, c0):
        if K != 0.0:  # zero terms are skipped
            self.angles.append((i, j, k, kind, K, c0))

    def chain(self, c, r0, ka):
        for b in range(len(c) - 1):
            self.bond(c[b], c[b + 1], r0)
        for b in range(len(c) - 2):
            self.angle(c[b], c[b + 1], c[b + 2], 0, ka, 0.0)


def mol_template(p):
    """mol_build, cel/engine/oercel.c:479-576, for a Params p."""
    v = p.__dict__["_v"]
    arch, nT, r0 = v["lip_arch"], v["lip_nT"], v["r0"]
    M = MolTemplate(arch)
    if arch == "linear":
        M.nH, M.mlen = v["lip_nH"], v["lip_nH"] + nT
        for b in range(M.mlen - 1):
            M.bond(b, b + 1, r0)
        for b in range(M.mlen - 2):
            M.angle(b, b + 1, b + 2, 0, v["k_angle"], 0.0)
        M.px = [(b - 0.5 * (M.mlen - 1)) * r0 for b in range(M.mlen)]
        M.py = [0.0] * M.mlen
    elif arch == "pi":
        h = v["r0_head"]
        M.nH, M.mlen = 2, 2 + 2 * nT
        M.bond(0, 1, h)
        ca = [0] + [1 + k for k in range(1, nT + 1)]
        cb = [1] + [1 + nT + k for k in range(1, nT + 1)]
        M.chain(ca, r0, v["k_angle"])
        M.chain(cb, r0, v["k_angle"])
        for m in range(1, min(v["lip_tie"], nT) + 1):
            M.bond(ca[m], cb[m], h)
        M.angle(1, 0, ca[1], 1, v["k_head"], v["cos_head"])
        M.angle(0, 1, cb[1], 1, v["k_head"], v["cos_head"])
        M.px, M.py = [0.0] * M.mlen, [0.0] * M.mlen
        M.py[0], M.py[1] = -0.5 * h, 0.5 * h
        for k in range(1, nT + 1):
            M.px[ca[k]], M.py[ca[k]] = k * r0, -0.5 * h
            M.px[cb[k]], M.py[cb[k]] = k * r0, 0.5 * h
    elif arch == "y":
        M.nH, M.mlen = 1, 1 + 2 * nT
        ca = [0] + list(range(1, nT + 1))
        cb = [0] + [nT + k for k in range(1, nT + 1)]
        M.chain(ca, r0, v["k_angle"])
        M.chain(cb, r0, v["k_angle"])
        for m in range(1, min(v["lip_tie"], nT) + 1):
            M.bond(ca[m], cb[m], 1.2 * r0)
        M.angle(ca[1], 0, cb[1], 1, v["k_head"], v["cos_head"])
        M.px, M.py = [0.0] * M.mlen, [0.0] * M.mlen
        for k in range(1, nT + 1):
            M.px[ca[k]], M.py[ca[k]] = 0.8 * r0 + (k - 1) * r0, -0.6 * r0
            M.px[cb[k]], M.py[cb[k]] = 0.8 * r0 + (k - 1) * r0, 0.6 * r0
    elif arch == "bola":
        h = v["r0_head"]
        M.nH, M.mlen = 3, 3 + 2 * nT
        M.bond(0, 1, h)
        ca = [0] + [2 + k for k in range(1, nT + 1)] + [2]
        cb = [1] + [2 + nT + k for k in range(1, nT + 1)] + [2]
        M.chain(ca, r0, v["k_angle"])
        M.chain(cb, r0, v["k_angle"])
        for m in range(1, min(v["lip_tie"], nT) + 1):
            M.bond(ca[m], cb[m], h * (1.0 - m / (nT + 1)))
        M.angle(1, 0, ca[1], 1, v["k_head"], v["cos_head"])
        M.angle(0, 1, cb[1], 1, v["k_head"], v["cos_head"])
        X = math.sqrt((nT + 1) * r0 * (nT + 1) * r0 - 0.25 * h * h)
        M.px, M.py = [0.0] * M.mlen, [0.0] * M.mlen
        M.py[0], M.py[1], M.px[2] = -0.5 * h, 0.5 * h, X
        for k in range(1, nT + 1):
            f = k / (nT + 1)
            M.px[ca[k]], M.py[ca[k]] = f * X, -0.5 * h * (1.0 - f)
            M.px[cb[k]], M.py[cb[k]] = f * X, 0.5 * h * (1.0 - f)
    else:
        raise ParamError("lip_arch must be linear, pi, y or bola")
    return M


def read_ini(path):
    """read_ini, cel/engine/oercel.c:408-434: '#' comments, key = value lines."""
    kv = []
    with open(path) as f:
        for ln, line in enumerate(f, 1):
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            if "=" not in line:
                raise ParamError(f"{path}:{ln}: expected key = value")
            k, v = line.split("=", 1)
            kv.append((k.strip(), v.strip()))
    return kv


def load(ini=None, overrides=(), finalize=True):
    """Defaults, then the ini file, then key=value overrides (strings or a dict),
    applied in the reference's two-pass order; then finalize_params."""
    kv = read_ini(ini) if ini else []
    if isinstance(overrides, dict):
        kv += list(overrides.items())
    else:
        for o in overrides:
            if "=" not in o:
                raise ParamError(f"override '{o}' is not key=value")
            k, v = o.split("=", 1)
            kv.append((k.strip(), v.strip()))
    p = Params()
    p.update(kv)
    if finalize:
        p.finalize()
    return p


def L_for_N(N, rho=3.0):
    """The box side L (nearest double to sqrt(N/rho)) for which lround(rho L^2) == N."""
    L = math.sqrt(N / rho)
    for k in range(64):
        if _llround(rho * L * L) == N:
            return L
        L = math.nextafter(L, math.inf if rho * L * L < N else -math.inf)
    raise ParamError(f"no L gives N = {N} at rho = {rho}")
