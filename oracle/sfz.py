"""TEST INFRASTRUCTURE ONLY: a pure-Python restatement of the reference's SFZ
loader (src/sfz.rs) and of the WAV reader it uses (src/wav.rs:166-365), for
checking the library's loader region by region on small synthetic
instruments. Never imported by the product.

f32 values and operations go through numpy.float32; log2f through glibc via
ctypes; the pool converter is oracle.sf2's.
"""
from __future__ import annotations

import ctypes
import math
import os
import re
import struct

import numpy as np

from .sf2 import POOL_GUARD, resample_i16, round_half_away

F = np.float32
_libm = ctypes.CDLL("libm.so.6")
_libm.log2f.restype, _libm.log2f.argtypes = ctypes.c_float, [ctypes.c_float]
_libc = ctypes.CDLL("libc.so.6")
_libc.strtof.restype, _libc.strtof.argtypes = ctypes.c_float, [ctypes.c_char_p, ctypes.c_void_p]


def log2f(x):
    return F(_libm.log2f(float(x)))


# ---- WAV ----
def read_wav(path: str) -> dict:
    with open(path, "rb") as f:
        b = f.read()
    if len(b) < 12 or b[:4] != b"RIFF":
        raise ValueError(f"{path}: not a RIFF file")
    if b[8:12] != b"WAVE":
        raise ValueError(f"{path}: not a WAVE file")
    pos, n = 12, len(b)
    w = dict(rate=0, channels=0, loop=None, root=None, fine=F(0))
    tag = bits = 0
    data = None
    while pos + 4 <= n:
        cid = b[pos:pos + 4]
        if pos + 8 > n:
            raise ValueError(f"{path}: truncated")
        size = struct.unpack_from("<I", b, pos + 4)[0]
        pos += 8
        padded = size + (size & 1)
        if size > n:
            raise ValueError(f"{path}: chunk claims {size} bytes")
        if cid == b"fmt ":
            if pos + 16 > n:
                raise ValueError(f"{path}: truncated")
            tag, w["channels"], w["rate"] = struct.unpack_from("<HHI", b, pos)
            bits = struct.unpack_from("<H", b, pos + 14)[0]
            if tag == 0xFFFE and size >= 40:
                tag = struct.unpack_from("<H", b, pos + 24)[0]
        elif cid == b"data":
            if pos + size > n:
                raise ValueError(f"{path}: truncated")
            data = b[pos:pos + size]
        elif cid == b"smpl":
            if pos + padded > n:
                raise ValueError(f"{path}: truncated")
            c = b[pos:pos + padded]
            if len(c) >= 36:
                note, frac, loops = struct.unpack_from("<III", c, 20)
                if note < 128:
                    w["root"] = note
                w["fine"] = F(frac / 4294967296.0 * 100.0)
                if loops > 0 and len(c) >= 60:
                    w["loop"] = struct.unpack_from("<II", c, 44)
        pos += padded
    if data is None:
        raise ValueError(f"{path}: no data chunk")
    if w["channels"] == 0:
        raise ValueError(f"{path}: no fmt chunk")
    d = np.frombuffer(data, np.uint8)
    if (tag, bits) == (1, 8):
        x = (d.astype(np.float32) - F(128)) / F(128)
    elif (tag, bits) == (1, 16):
        x = np.frombuffer(data[:len(data) // 2 * 2], "<i2").astype(np.float32) / F(32768)
    elif (tag, bits) == (1, 24):
        m = len(data) // 3
        t = d[:m * 3].reshape(m, 3).astype(np.uint32)
        v = ((t[:, 0] << 8) | (t[:, 1] << 16) | (t[:, 2] << 24)).view(np.int32) >> 8
        x = v.astype(np.float32) / F(8388608)
    elif (tag, bits) == (1, 32):
        x = np.frombuffer(data[:len(data) // 4 * 4], "<i4").astype(np.float32) / F(2147483648)
    elif (tag, bits) == (3, 32):
        x = np.frombuffer(data[:len(data) // 4 * 4], "<f4").astype(np.float32)
    elif (tag, bits) == (3, 64):
        x = np.frombuffer(data[:len(data) // 8 * 8], "<f8").astype(np.float32)
    else:
        raise ValueError(f"{path}: unsupported format tag {tag} / {bits} bits")
    w["x"] = x
    return w


def channel_i16(w, ch):
    nch = w["channels"]
    ch = min(ch, max(nch - 1, 0))
    s = np.clip(w["x"][ch::max(nch, 1)], F(-1), F(1)) * F(32767)
    r = np.where(np.isnan(s), F(0), s)
    return np.array([int(round_half_away(float(v))) for v in r], np.int16)


# ---- opcode values ----
_INT = re.compile(r"[+-]?[0-9]+\Z")
_FLT = re.compile(r"[+-]?(([0-9]+\.?[0-9]*|\.[0-9]+)([eE][+-]?[0-9]+)?|inf|infinity|nan)\Z",
                  re.IGNORECASE)
WS = " \t\n\r\f\v"


def p_i32(s):
    if _INT.match(s):
        v = int(s)
        return v if -2 ** 31 <= v < 2 ** 31 else None
    return None


def p_f32(s):
    """Decimal straight to the nearest f32, as Rust's str::parse::<f32> reads it."""
    return F(_libc.strtof(s.encode(), None)) if _FLT.match(s) else None


def as_i32(f):
    if f != f:
        return 0
    return max(min(int(f) if math.isfinite(f) else (2 ** 31 - 1 if f > 0 else -2 ** 31),
                   2 ** 31 - 1), -2 ** 31)


def note_name(s):
    """src/sfz.rs:88-116 (c4 = 60)."""
    if not s:
        return None
    step = {"c": 0, "d": 2, "e": 4, "f": 5, "g": 7, "a": 9, "b": 11}.get(s[0].lower())
    if step is None:
        return None
    i, acc = 1, 0
    while i < len(s) and (s[i] in "#b" or (s[i] == "-" and i == 1)):
        if s[i] == "#":
            acc += 1
        elif s[i] == "b":
            acc -= 1
        else:
            break
        i += 1
    o = p_i32(s[i:].strip(WS))
    return None if o is None else (o + 1) * 12 + step + acc


class Ops(dict):
    def f32(self, k):
        return p_f32(self[k].strip(WS)) if k in self else None

    def i32(self, k):
        if k not in self:
            return None
        t = self[k].strip(WS)
        v = p_i32(t)
        if v is None:
            f = p_f32(t)
            v = None if f is None else as_i32(float(f))
        return v

    def key(self, k):
        if k not in self:
            return None
        t = self[k].strip(WS)
        v = p_i32(t)
        return v if v is not None else note_name(t)

    def canon(self):
        if "key" in self:
            k = self.pop("key")
            for a in ("lokey", "hikey", "pitch_keycenter"):
                self.setdefault(a, k)
        for a, b in (("loopmode", "loop_mode"), ("loopstart", "loop_start"),
                     ("loopend", "loop_end"), ("pitch", "tune")):
            if a in self:
                self.setdefault(b, self.pop(a))


def value_end(s):
    """src/sfz.rs:226-252."""
    i, last = 0, None
    while i < len(s):
        if s[i] in "=<":
            return i if last is None else last
        if s[i] in " \t\n\f\r":
            j = i
            while j < len(s) and s[j] in " \t\n\f\r":
                j += 1
            last, i = i, j
            continue
        i += 1
    return len(s)


def resolve(root, default_path, rel):
    cand = os.path.join(root, default_path, rel.replace("\\", "/"))
    if os.path.exists(cand):
        return cand
    parent, name = os.path.dirname(cand) or root, os.path.basename(cand).lower()
    try:
        for e in os.listdir(parent):  # readdir order
            if e.lower() == name:
                return os.path.join(parent, e)
    except OSError:
        pass
    return cand


def parse_file(root, path, out, depth=0):
    if depth > 16:
        raise ValueError("#include nesting is too deep")
    with open(path, "rb") as f:
        text = f.read().decode("utf-8")
    header, cur, started = "global", Ops(), False
    for raw in text.split("\n"):
        raw = raw[:-1] if raw.endswith("\r") else raw
        i = raw.find("//")
        line = (raw if i < 0 else raw[:i]).strip(WS)
        if not line:
            continue
        if line.startswith("#include"):
            inc = line[8:].strip(WS).strip('"').strip(WS)
            p = resolve(root, "", inc)
            if started:
                out.append((header, cur))
                cur, started = Ops(), False
            try:
                parse_file(root, p, out, depth + 1)
            except (OSError, ValueError, UnicodeDecodeError):
                pass
            continue
        if line.startswith("#define"):
            continue
        rest = line
        while rest:
            rest = rest.lstrip(WS)
            if rest.startswith("<"):
                end = rest.find(">", 1)
                if end < 0:
                    break
                if started:
                    out.append((header, cur))
                header = rest[1:end].strip(WS).lower()
                cur, started = Ops(), True
                rest = rest[end + 1:]
                continue
            eq = rest.find("=")
            if eq < 0:
                break
            key = rest[:eq].strip(WS).lower()
            after = rest[eq + 1:]
            ve = value_end(after)
            value = after[:ve].strip(WS)
            rest = after[ve:]
            if not key:
                continue
            cur[key] = value
            started = True
    if started:
        out.append((header, cur))


def region_from(ops, sample, info):
    """src/sfz.rs:547-676."""
    r = dict(sample=sample, key_lo=0, key_hi=127, vel_lo=0, vel_hi=127, root_key_override=-1,
             fixed_key=-1, fixed_vel=-1, coarse_tune=0, fine_tune=0, scale_tuning=100,
             attenuation_cb=F(0), amp_veltrack=F(100), pan=F(0), loop_mode=0, addr_start=0,
             addr_end=0, addr_loop_start=0, addr_loop_end=0, exclusive_class=0, delay=F(0),
             attack=F(0.001), hold=F(0), decay=F(0.001), sustain=F(1), release=F(0.001),
             keynum_to_hold=0, keynum_to_decay=0, filter_fc_cents=F(13500), filter_q_cb=F(0),
             filter_veltrack_cents=F(0))
    cl = lambda v, lo, hi: lo if v < lo else (hi if v > hi else v)
    if (k := ops.key("lokey")) is not None:
        r["key_lo"] = cl(k, 0, 127)
    if (k := ops.key("hikey")) is not None:
        r["key_hi"] = cl(k, 0, 127)
    if (k := ops.key("pitch_keycenter")) is not None:
        r["root_key_override"] = cl(k, 0, 127)
    if r["root_key_override"] < 0:
        r["root_key_override"] = info["root_key"]
    if (k := ops.i32("lovel")) is not None:
        r["vel_lo"] = cl(k, 0, 127)
    if (k := ops.i32("hivel")) is not None:
        r["vel_hi"] = cl(k, 0, 127)
    tune = ops.f32("tune")
    t = cl(F(round_half_away(float(tune if tune is not None else F(0)))), F(-32768), F(32767))
    r["fine_tune"] = 0 if t != t else int(t)
    if (k := ops.i32("transpose")) is not None:
        r["coarse_tune"] = cl(k, -127, 127)
    if (f := ops.f32("pitch_keytrack")) is not None:
        t = cl(F(round_half_away(float(f))), F(-1200), F(1200))
        r["scale_tuning"] = 0 if t != t else int(t)
    if (f := ops.f32("volume")) is not None:
        r["attenuation_cb"] = -f * F(10)
    if (f := ops.f32("pan")) is not None:
        r["pan"] = cl(f / F(100), F(-1), F(1))
    if (f := ops.f32("amp_veltrack")) is not None:
        r["amp_veltrack"] = cl(f, F(-100), F(100))
    mode = ops.get("loop_mode", "")
    r["loop_mode"] = {"loop_continuous": 1, "loop_sustain": 2, "one_shot": 0, "no_loop": 0}.get(
        mode, 1 if info["loop_end"] > info["loop_start"] + 1 else 0)
    if (k := ops.i32("offset")) is not None:
        r["addr_start"] = max(k, 0)

    def to_source(v):
        if info["resample_ratio"] > 0:
            return as_i32(float(F(round_half_away(float(F(v) / info["resample_ratio"])))))
        return v
    wrap = lambda v: (v + 2 ** 31) % 2 ** 32 - 2 ** 31
    if (k := ops.i32("end")) is not None and k > 0:
        r["addr_end"] = wrap(k - to_source(info["len"]))
    if (k := ops.i32("loop_start")) is not None:
        r["addr_loop_start"] = wrap(k - to_source(info["loop_start"]))
    if (k := ops.i32("loop_end")) is not None:
        r["addr_loop_end"] = wrap(k - to_source(info["loop_end"]))

    def env(name, d):
        v = ops.f32(name)
        v = F(d) if v is None else v
        return F(0) if v != v or v < 0 else v
    r["delay"] = env("ampeg_delay", 0.0)
    r["attack"] = env("ampeg_attack", 0.001)
    r["hold"] = env("ampeg_hold", 0.0)
    r["decay"] = env("ampeg_decay", 100.0)
    s = ops.f32("ampeg_sustain")
    r["sustain"] = cl((F(100) if s is None else s) / F(100), F(0), F(1))
    r["release"] = env("ampeg_release", 0.05)
    if (f := ops.f32("cutoff")) is not None and f > 0:
        r["filter_fc_cents"] = F(1200) * log2f(f / F(8.176))
    if (f := ops.f32("resonance")) is not None:
        r["filter_q_cb"] = f * F(10)
    if (f := ops.f32("fil_veltrack")) is not None:
        r["filter_veltrack_cents"] = cl(f, F(-9600), F(9600))
    if "fil_type" in ops and not ops["fil_type"].startswith("lpf"):
        r["filter_fc_cents"], r["filter_veltrack_cents"] = F(13500), F(0)
    g = cl(ops.i32("group") or 0, 0, 255)
    off = ops.i32("off_by")
    if off is not None and cl(off, 0, 255) == g and g > 0:
        r["exclusive_class"] = g
    return r


def load(path: str, sample_rate: int, resample_pool: bool = True):
    """sfz::load -> (pool, samples, regions, presets, preset_regions, pool_rate)."""
    root = os.path.dirname(path) or "."
    sections = []
    parse_file(root, path, sections)
    for _, o in sections:
        o.canon()
    dp, glob, master, group, sets = "", Ops(), Ops(), Ops(), []
    for h, o in sections:
        if h == "control":
            if "default_path" in o:
                dp = o["default_path"].replace("\\", "/")
        elif h == "global":
            glob, master, group = o, Ops(), Ops()
        elif h == "master":
            master, group = o, Ops()
        elif h == "group":
            group = o
        elif h == "region":
            m = Ops(glob)
            m.update(master)
            m.update(group)
            m.update(o)
            sets.append((dp, m))
    if not sets:
        raise ValueError(f"{path}: no <region> sections")
    pool_rate = sample_rate if resample_pool else 0
    pool, samples, cache, regions, ids, at = [], [], {}, [], [], 0
    for dpath, ops in sets:
        if "sample" not in ops or ops.i32("end") == -1:
            continue
        sp = resolve(root, dpath, ops["sample"])
        if (sp, 0) in cache:
            chans = [cache[(sp, 0)]] + ([cache[(sp, 1)]] if (sp, 1) in cache else [])
        else:
            try:
                w = read_wav(sp)
            except (OSError, ValueError):
                continue
            chans = []
            for ch in range(min(max(w["channels"], 1), 2)):
                raw = channel_i16(w, ch)
                src = max(w["rate"], 1)
                ratio = pool_rate / src if pool_rate and pool_rate != src else 1.0
                data = resample_i16(raw, ratio) if abs(ratio - 1.0) > 1e-12 else raw
                n = len(data)
                ls, le = w["loop"] if w["loop"] is not None else (0, max(n - 1, 0))
                ls, le = int(round_half_away(ls * ratio)), int(round_half_away(le * ratio))
                cache[(sp, ch)] = len(samples)
                chans.append(len(samples))
                samples.append(dict(start=at, len=n, loop_start=min(ls, max(n - 1, 0)),
                                    loop_end=min(le, n), rate=pool_rate if pool_rate else src,
                                    root_key=w["root"] if w["root"] is not None else 60,
                                    correction_cents=w["fine"], resample_ratio=F(ratio)))
                pool += [np.asarray(data, np.int16), np.zeros(POOL_GUARD, np.int16)]
                at += n + POOL_GUARD
        for ch, s in enumerate(chans):
            r = region_from(ops, s, samples[s])
            if len(chans) == 2:
                r["pan"] = F(-1) if ch == 0 else F(1)
            ids.append(len(regions))
            regions.append(r)
    if not regions:
        raise ValueError(f"{path}: every region was unusable")
    return (np.concatenate(pool), samples, regions, [(0, 0, 0, len(ids))], ids, pool_rate)
