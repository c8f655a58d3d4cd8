"""The reference engine's file formats, for interchange with its tools and its
CPU engine:

  snapshot    snap_<step>.bin, "OCEL" v1: positions (float32) and types in gid
              order plus the bond list (cel/engine/oercel.c:1830-1884)
  checkpoint  ckpt.bin, "OCKP" v2 (v3 with the architecture code for branched
              lipids): header (step, time, L, seed, rng state, max bond, warm-up
              end, time anchors, soup seed) and the slot-ordered state
              gid/type/mpos/nxt/x/y/vh/f (cel/engine/oercel.c:1886-1936), read
              back with the reference's checks (:1942-2010)
  stats.csv   step,time,T_kin,nW..nZ,n_lipids,max_bond (:2264-2267, :2349)
  events.csv  step,kind,i,x,y,variant (:2273)

States are dicts of numpy arrays in gid order, as Engine.get_state returns them.
Chemistry bonds of the precursor chemistry are appended to the snapshot's bond
list after the lipid bonds; a checkpoint holding them carries them in an
extension block after the reference's END! marker ("OCMB" + partner gid per
bead), which the reference's reader never reaches.  The reference's reader
accepts T beads only inside lipids, so checkpoints with free T are for this
engine only.
"""
import os
import struct

import numpy as np

from .abi import W, H, T, N, KA, KB, Z  # noqa: F401

STATS_HEADER = "step,time,T_kin,nW,nH,nT,nN,nKa,nKb,nZ,n_lipids,max_bond"
EVENTS_HEADER = "step,kind,i,x,y,variant"


class FormatError(ValueError):
    pass


def _mix13(z):
    m = (1 << 64) - 1
    z = ((z ^ (z >> 30)) * 0xbf58476d1ce4e5b9) & m
    z = ((z ^ (z >> 27)) * 0x94d049bb133111eb) & m
    return z ^ (z >> 31)


def rng_seed_state(seed):
    """xoshiro256** state after rng_seed(seed) (cel/engine/oercel.c:156-162)."""
    z = seed & ((1 << 64) - 1)
    out = []
    for _ in range(4):
        z = (z + 0x9e3779b97f4a7c15) & ((1 << 64) - 1)
        out.append(_mix13(z))
    return out


def lipid_bonds(s, mlen, tbonds=None):
    """Template bonds of the lipids, ordered by the gid of the bond's first
    template bead then template order, each as (min gid, max gid).  tbonds:
    the template's (bi, bj) pairs (default: a linear chain, (b, b+1))."""
    if tbonds is None:
        tbonds = [(b, b + 1) for b in range(mlen - 1)]
    mpos, nxt = s["mpos"], s["nxt"]
    heads = np.flatnonzero(mpos == 0)
    out = []
    for h in heads:
        chain = [int(h)]
        while len(chain) < mlen and nxt[chain[-1]] >= 0:
            chain.append(int(nxt[chain[-1]]))
        if len(chain) != mlen:
            raise FormatError(f"broken lipid chain at gid {int(h)}")
        for q, (bi, bj) in enumerate(tbonds):
            i, j = chain[bi], chain[bj]
            out.append((i, q, min(i, j), max(i, j)))
    out.sort(key=lambda t: (t[0], t[1]))
    return [(a, b) for _, _, a, b in out]


def chem_bonds(s):
    bd = s.get("bond")
    if bd is None:
        return []
    g = np.flatnonzero((s["mpos"] < 0) & (bd > np.arange(len(bd))))
    return [(int(i), int(bd[i])) for i in g]


def write_snapshot(path, s, step, time, L, mlen, tbonds=None):
    """OCEL v1 (cel/engine/oercel.c:1830-1884); tbonds: the lipid template's
    (bi, bj) bond pairs (default: a linear chain)."""
    Nb = len(s["x"])
    bonds = lipid_bonds(s, mlen, tbonds) + chem_bonds(s)
    with open(path, "wb") as f:
        f.write(b"OCEL")
        f.write(struct.pack("<iii", 1, Nb, len(bonds)))
        f.write(struct.pack("<qdd", int(step), float(time), float(L)))
        f.write(np.asarray(s["x"], dtype=np.float32).tobytes())
        f.write(np.asarray(s["y"], dtype=np.float32).tobytes())
        f.write(np.asarray(s["type"], dtype=np.uint8).tobytes())
        if bonds:
            f.write(np.asarray(bonds, dtype=np.int32).ravel().tobytes())


def read_snapshot(path):
    with open(path, "rb") as f:
        b = f.read()
    if b[:4] != b"OCEL":
        raise FormatError(f"{path} is not an Oercel snapshot")
    ver, n, nb = struct.unpack_from("<iii", b, 4)
    if ver != 1:
        raise FormatError(f"unsupported snapshot version {ver}")
    step, time, L = struct.unpack_from("<qdd", b, 16)
    o = 40
    x = np.frombuffer(b, "<f4", n, o)
    y = np.frombuffer(b, "<f4", n, o + 4 * n)
    t = np.frombuffer(b, "u1", n, o + 8 * n)
    bonds = np.frombuffer(b, "<i4", 2 * nb, o + 9 * n).reshape(nb, 2)
    return {"step": step, "time": time, "L": L, "x": x.copy(), "y": y.copy(), "type": t.copy(), "bonds": bonds.copy()}


def write_checkpoint(path, s, slot_gid, hdr, keep_path=None):
    """OCKP checkpoint (cel/engine/oercel.c:1886-1936): version 2, or version 3
    (version 2 + the lipid architecture code) when hdr["arch"] is not linear,
    as the reference writes them.  s: gid-order state (x, y, vx, vy = vh, fx,
    fy, type, mpos, nxt[, bond]); slot_gid: gid of each slot (the cell order,
    e.g. Engine.cells()[1]); hdr: step, time, L, seed, mlen, nH, and
    optionally arch (0 linear, 1 pi, 2 y, 3 bola), rng (4 ints), maxbond,
    warm_end, t_anchor, step_anchor, dt_anchor, soup_seed.  Written to
    path + ".tmp", flushed to disk and renamed onto path, so a kill during the
    write leaves the previous checkpoint; keep_path (the reference's ckpt_keep
    copy) is a hard link to the same file."""
    slot_gid = np.asarray(slot_gid, dtype=np.int64)
    Nb = len(slot_gid)
    slot = np.empty(Nb, np.int64)
    slot[slot_gid] = np.arange(Nb)
    nxt_g = np.asarray(s["nxt"])
    nxt_slot = np.where(nxt_g >= 0, slot[np.maximum(nxt_g, 0)], -1).astype(np.int32)[slot_gid]
    rng = hdr.get("rng") or rng_seed_state(int(hdr["seed"]))
    arch = int(hdr.get("arch", 0))
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(b"OCKP")
        f.write(struct.pack("<iiii", 2 if arch == 0 else 3, Nb, int(hdr["mlen"]), int(hdr["nH"])))
        if arch != 0:
            f.write(struct.pack("<i", arch))
        f.write(struct.pack("<qddQ", int(hdr["step"]), float(hdr["time"]), float(hdr["L"]), int(hdr["seed"])))
        f.write(struct.pack("<4Q", *[int(v) & ((1 << 64) - 1) for v in rng]))
        f.write(struct.pack("<dqdqdQ", float(hdr.get("maxbond", 0.0)), int(hdr.get("warm_end", 0)),
                            float(hdr.get("t_anchor", 0.0)), int(hdr.get("step_anchor", 0)),
                            float(hdr.get("dt_anchor", hdr.get("dt", 0.0))), int(hdr.get("soup_seed", hdr["seed"]))))
        f.write(slot_gid.astype(np.uint32).tobytes())
        f.write(np.asarray(s["type"], np.uint8)[slot_gid].tobytes())
        f.write(np.asarray(s["mpos"], np.int8)[slot_gid].tobytes())
        f.write(nxt_slot.tobytes())
        for k in ("x", "y", "vx", "vy", "fx", "fy"):
            f.write(np.asarray(s[k], np.float64)[slot_gid].tobytes())
        f.write(b"END!")
        bd = s.get("bond")
        if bd is not None and np.any(bd >= 0):
            f.write(b"OCMB")
            f.write(np.asarray(bd, np.int32).tobytes())
        f.flush()
        os.fsync(f.fileno())
    if keep_path:
        try:
            os.link(tmp, keep_path)
        except FileExistsError:
            pass
    os.replace(tmp, path)


def read_checkpoint(path):
    """(state in gid order, header, slot_gid) of an OCKP v2/v3 file, with the
    reference's consistency checks (cel/engine/oercel.c:1942-2010, minus the
    rule that T beads belong to lipids when an OCMB block is present)."""
    with open(path, "rb") as f:
        b = f.read()
    try:
        return _read_checkpoint(b, path)
    except struct.error:
        raise FormatError("checkpoint truncated") from None


def _read_checkpoint(b, path):
    if b[:4] != b"OCKP":
        raise FormatError(f"{path} is not an Oercel checkpoint")
    ver, Nb, ml, nH = struct.unpack_from("<iiii", b, 4)
    if ver == 1:
        raise FormatError(f"{path} is a version 1 checkpoint; its history cannot be verified")
    if ver not in (2, 3):
        raise FormatError(f"unsupported checkpoint version {ver}")
    o = 20
    arch = 0
    if ver == 3:
        (arch,) = struct.unpack_from("<i", b, o)
        o += 4
        if arch not in (1, 2, 3):
            raise FormatError(f"checkpoint: unknown lipid architecture code {arch}")
    step, time, L, seed = struct.unpack_from("<qddQ", b, o)
    o += 32
    rng = list(struct.unpack_from("<4Q", b, o))
    o += 32
    maxbond, warm_end, t_anchor, step_anchor, dt_anchor, soup_seed = struct.unpack_from("<dqdqdQ", b, o)
    o += 48
    need = o + Nb * (4 + 1 + 1 + 4 + 6 * 8) + 4
    if len(b) < need:
        raise FormatError("checkpoint truncated")
    gid = np.frombuffer(b, "<u4", Nb, o).astype(np.int64)
    o += 4 * Nb
    typ = np.frombuffer(b, "u1", Nb, o)
    o += Nb
    mpos = np.frombuffer(b, "i1", Nb, o)
    o += Nb
    nxt = np.frombuffer(b, "<i4", Nb, o)
    o += 4 * Nb
    arr = {}
    for k in ("x", "y", "vx", "vy", "fx", "fy"):
        arr[k] = np.frombuffer(b, "<f8", Nb, o)
        o += 8 * Nb
    if b[o:o + 4] != b"END!":
        raise FormatError("checkpoint end marker missing")
    o += 4
    bond = None
    if b[o:o + 4] == b"OCMB":
        bond = np.frombuffer(b, "<i4", Nb, o + 4).copy()
    if np.any(gid >= Nb) or len(np.unique(gid)) != Nb:
        raise FormatError("checkpoint: gid table is not a permutation")
    if np.any(typ >= Z):
        raise FormatError("checkpoint: bead type Z is unused")
    lipid_type = (typ == H) | (typ == T)
    bad = (nxt >= Nb) | ((mpos < 0) & (typ == H))
    if bond is None:
        bad |= (mpos < 0) == lipid_type
    if np.any(bad):
        raise FormatError("checkpoint: inconsistent lipid links")
    s = {}
    for k, v in arr.items():
        a = np.empty(Nb)
        a[gid] = v
        s[k] = a
    s["type"] = np.empty(Nb, np.uint8)
    s["type"][gid] = typ
    s["mpos"] = np.empty(Nb, np.int8)
    s["mpos"][gid] = mpos
    nx = np.full(Nb, -1, np.int32)
    has = nxt >= 0
    nx[gid[has]] = gid[nxt[has]]
    s["nxt"] = nx
    s["bond"] = bond if bond is not None else np.full(Nb, -1, np.int32)
    hdr = {"version": ver, "N": Nb, "mlen": ml, "nH": nH, "arch": arch, "step": step, "time": time, "L": L, "seed": seed,
           "rng": rng, "maxbond": maxbond, "warm_end": warm_end, "t_anchor": t_anchor, "step_anchor": step_anchor,
           "dt_anchor": dt_anchor, "soup_seed": soup_seed}
    return s, hdr, gid


def stats_row(st):
    """One stats.csv row from Engine.stats() (cel/engine/oercel.c:2349-2351)."""
    c = st["count"]
    return (f"{st['step']},{st['time']:.6f},{st['T_kin']:.6f},{c[0]},{c[1]},{c[2]},{c[3]},{c[4]},{c[5]},{c[6]},"
            f"{st['n_lipids']},{st['max_bond']:.6f}")


def events_rows(ev):
    """events.csv rows (step,kind,i,x,y,variant), in the order of `ev`:
      - the precursor chemistry: conv / form / break, i = first gid,
        variant = second gid, no position;
      - the reference chemistry (flush_events, cel/engine/oercel.c:2036-2047):
        syn (i = the catalyst), rep / mut (i = the new catalyst) at the
        catalyst's position and dK at the bead's, variant a / b (the kind of
        catalyst); dlip (i = the lipid's first bead) and dN at the bead's
        position, variant "-"; feed as one row per check with i = the number
        of conversions."""
    names = {0: "conv", 1: "form", 2: "break", 3: "dlip", 4: "dN", 6: "syn", 7: "rep", 8: "mut", 9: "dK"}
    rows, feed = [], {}
    for e in ev:
        k, st = int(e["kind"]), int(e["step"])
        if k <= 2:
            rows.append(f"{st},{names[k]},{int(e['gid_a'])},nan,nan,{int(e['gid_b'])}")
        elif k == 5:
            if st not in feed:
                feed[st] = 0
                rows.append(st)  # placeholder, filled below
            feed[st] += 1
        else:
            v = "ab"[int(e["gid_b"])] if k >= 6 else "-"
            rows.append(f"{st},{names[k]},{int(e['gid_a'])},{float(e['x']):.4f},{float(e['y']):.4f},{v}")
    return [f"{r},feed,{feed[r]},nan,nan,-" if isinstance(r, int) else r for r in rows]


def save_checkpoint(engine, path, r=0, keep_path=None):
    """Checkpoint of replica r of an Engine in the reference's format (version
    3 with the architecture code for branched lipids); keep_path: an extra
    name for the same file (ckpt_keep)."""
    from .params import ARCHS
    p = engine.params[r]
    s = engine.get_state(r)
    _, slot_gid = engine.cells(r)
    step = engine.step_count(r)
    hdr = {"step": step, "time": p.t_anchor + (step - p.step_anchor) * p.dt, "L": p.L, "seed": p.seed,
           "mlen": p.mlen, "nH": p.nH, "arch": ARCHS.index(p.lip_arch),
           "warm_end": max(p.warm_end, p.step_anchor + p.warmup),
           "t_anchor": p.t_anchor, "step_anchor": p.step_anchor, "dt_anchor": p.dt, "soup_seed": p.seed,
           "maxbond": engine.max_bond(r)}
    write_checkpoint(path, s, slot_gid, hdr, keep_path=keep_path)
    return hdr


def check_architecture(params, hdr):
    """The reference's restart check (cel/engine/oercel.c:1976): the lipid
    length, heads and architecture of the checkpoint are the parameters'."""
    from .params import ARCHS
    want = (params.mlen, params.nH, ARCHS.index(params.lip_arch))
    got = (int(hdr["mlen"]), int(hdr["nH"]), int(hdr.get("arch", 0)))
    if got != want:
        raise FormatError(f"checkpoint lipid architecture (mlen, nH, arch) = {got} differs from the parameters' {want}")


def restart_params(params, hdr):
    """Parameters for continuing a checkpoint, with the reference's restart
    rules (cel/engine/oercel.c:1985-1989, 2174-2186): the checkpoint's seed
    unless restart_reseed, its warm-up end, and its time anchor while dt is
    unchanged (a new anchor at the checkpoint's time otherwise)."""
    q = params.copy()
    q.warm_end = int(hdr["warm_end"])
    if hdr["dt_anchor"] == q.dt:
        q.t_anchor, q.step_anchor = float(hdr["t_anchor"]), int(hdr["step_anchor"])
    else:
        q.t_anchor, q.step_anchor = float(hdr["time"]), int(hdr["step"])
    if not q.restart_reseed:
        q.seed = int(hdr["seed"])
    return q


def load_checkpoint(engine, path, r=0):
    """Load a checkpoint into replica r of an Engine created with
    restart_params(params, header) (same L), in the checkpoint's slot order:
    the continuation is bitwise the uninterrupted run's, as the reference's
    restart is (cel/engine/tests/validate.sh:80-105)."""
    s, hdr, slot_gid = read_checkpoint(path)
    p = engine.params[r]
    check_architecture(p, hdr)
    if abs(hdr["L"] - p.L) > 1e-12:
        raise FormatError(f"checkpoint L={hdr['L']} differs from parameter L={p.L}")
    engine.set_state(s, step=hdr["step"], r=r, slot_gid=slot_gid)
    engine.max_bond(r, set_to=hdr["maxbond"])
    return hdr
