"""TEST INFRASTRUCTURE ONLY: a pure-Python restatement of the reference's
SoundFont 2 loader (src/sf2.rs) and of the pool's sample-rate converter
(src/resample.rs), for checking the library's loader field by field and sample
by sample on small synthetic soundfonts. Never imported by the product.

f64 work is Python floats; every f32 value and f32 operation goes through
numpy.float32 (the converter's accumulation runs tap by tap over all outputs at
once, each output's sum in the same tap order as the reference's loop). The
timecent and centibel curves use oracle.driver's glibc bindings.
"""
from __future__ import annotations

import math
import struct

import numpy as np

from .driver import exp2f, powf

F = np.float32
GEN_COUNT = 60
POOL_GUARD = 8
ABSOLUTE_ONLY = {0, 1, 2, 3, 4, 12, 45, 50, 53, 54, 57, 58, 46, 47, 41, 43, 44}


def default_gens():
    """The SoundFont 2.01 defaults, src/sf2.rs:79-101."""
    g = [0] * GEN_COUNT
    g[8] = 13500
    for op in (21, 23, 25, 26, 27, 28, 30, 33, 34, 35, 36, 38):
        g[op] = -12000
    g[43] = g[44] = 0x7F00  # key and velocity ranges (0, 127)
    g[46] = g[47] = -1
    g[56] = 100
    g[58] = -1
    return g


def timecents(tc: int):
    return F(0) if F(tc) <= F(-12000) else exp2f(F(tc) / F(1200))


def cb_to_gain(cb):
    return powf(F(10), -F(cb) / F(200))


def range_of(g, is_set, op):
    if not is_set:
        return (0, 127)
    raw = g[op] & 0xFFFF
    lo, hi = raw & 0xFF, raw >> 8
    return (hi, lo) if lo > hi else (lo, hi)


def sat16(v):
    return max(-32768, min(32767, v))


def region_of(g, sample, key, vel) -> dict:
    """build_region, src/sf2.rs:532-581."""
    mode = g[54] & 3
    return dict(
        sample=sample, key_lo=key[0], key_hi=key[1], vel_lo=vel[0], vel_hi=vel[1],
        root_key_override=g[58], fixed_key=g[46], fixed_vel=g[47], coarse_tune=g[51],
        fine_tune=g[52], scale_tuning=g[56], keynum_to_hold=g[39], keynum_to_decay=g[40],
        attenuation_cb=F(g[48]), amp_veltrack=F(100), pan=min(max(F(g[17]) / F(500), F(-1)), F(1)),
        loop_mode={1: 1, 3: 2}.get(mode, 0), addr_start=g[0] + g[4] * 32768,
        addr_end=g[1] + g[12] * 32768, addr_loop_start=g[2] + g[45] * 32768,
        addr_loop_end=g[3] + g[50] * 32768, exclusive_class=max(0, min(255, g[57])),
        delay=timecents(g[33]), attack=timecents(g[34]), hold=timecents(g[35]),
        decay=timecents(g[36]), sustain=cb_to_gain(max(F(g[37]), F(0))), release=timecents(g[38]),
        filter_fc_cents=F(g[8]), filter_q_cb=F(g[9]), filter_veltrack_cents=F(0))


def scan(raw: bytes):
    """The sub-chunks of the sdta and pdta lists, src/sf2.rs:153-207."""
    if raw[:4] != b"RIFF" or len(raw) < 8:
        raise ValueError("not a RIFF file")
    if raw[8:12] != b"sfbk":
        raise ValueError("RIFF form is not sfbk")
    lists = {b"sdta": {}, b"pdta": {}}
    pos, n = 12, len(raw)
    while pos + 8 <= n:
        cid, size = raw[pos:pos + 4], struct.unpack_from("<I", raw, pos + 4)[0]
        size = min(size, n - pos - 8)
        if cid == b"LIST":
            lid = raw[pos + 8:pos + 12]
            if lid in lists:
                end, inner = pos + 8 + size, pos + 12
                while inner + 8 <= end:
                    sid, ssz = raw[inner:inner + 4], struct.unpack_from("<I", raw, inner + 4)[0]
                    ssz = min(ssz, end - inner - 8)
                    lists[lid].setdefault(sid, (inner + 8, ssz))
                    inner += 8 + ssz + (ssz & 1)
        pos += 8 + size + (size & 1)
    return lists[b"sdta"], lists[b"pdta"]


def load(path: str, sample_rate: int, resample_pool: bool = True, budget: int = 2 << 30):
    """sf2::load -> (pool, samples, regions, presets, preset_regions, pool_rate) as
    lists of dicts (presets as (bank, program, first, count))."""
    with open(path, "rb") as f:
        raw = f.read()
    sdta, pdta = scan(raw)
    if b"smpl" not in sdta:
        raise ValueError("soundfont has no smpl chunk")
    smpl_off, smpl_size = sdta[b"smpl"]
    smpl_frames = smpl_size // 2

    def rec(name, size):
        if name not in pdta:
            raise ValueError(f"pdta is missing {name.decode()}")
        off, n = pdta[name]
        return [raw[off + i * size: off + (i + 1) * size] for i in range(n // size)]
    phdr, pbag, pgen = rec(b"phdr", 38), rec(b"pbag", 4), rec(b"pgen", 4)
    inst, ibag, igen, shdr = rec(b"inst", 22), rec(b"ibag", 4), rec(b"igen", 4), rec(b"shdr", 46)
    u16 = lambda b, o: struct.unpack_from("<H", b, o)[0]
    i16 = lambda b, o: struct.unpack_from("<h", b, o)[0]
    shdrs = [dict(zip(("start", "end", "sloop", "eloop", "rate"), struct.unpack_from("<5I", c, 20)),
                  pitch=c[40], corr=struct.unpack_from("<b", c, 41)[0], type=u16(c, 44))
             for c in shdr]
    if len(phdr) < 2:
        raise ValueError("soundfont has no presets")
    regions, presets, pre_regions, used = [], [], [], set()
    for pi in range(len(phdr) - 1):
        blo, bhi = u16(phdr[pi], 24), u16(phdr[pi + 1], 24)
        if blo >= bhi or bhi > len(pbag):
            continue
        pglob, pglob_set, first = [0] * GEN_COUNT, [False] * GEN_COUNT, True
        start = len(pre_regions)
        for bi in range(blo, bhi):
            glo = u16(pbag[bi], 0)
            ghi = u16(pbag[bi + 1], 0) if bi + 1 < len(pbag) else len(pgen)
            if glo > ghi or ghi > len(pgen):
                continue
            zone, zset, instrument = [0] * GEN_COUNT, [False] * GEN_COUNT, None
            for gi in range(glo, ghi):
                op, amt = u16(pgen[gi], 0), i16(pgen[gi], 2)
                if op == 41:
                    instrument = amt & 0xFFFF
                elif op < GEN_COUNT:
                    zone[op], zset[op] = amt, True
            if instrument is None:
                if first:
                    pglob, pglob_set = zone, zset
                first = False
                continue
            first = False
            pg = [zone[i] if zset[i] else pglob[i] for i in range(GEN_COUNT)]
            ps = [zset[i] or pglob_set[i] for i in range(GEN_COUNT)]
            pkey, pvel = range_of(pg, ps[43], 43), range_of(pg, ps[44], 44)
            if instrument >= len(inst):
                continue
            iblo = u16(inst[instrument], 20)
            ibhi = u16(inst[instrument + 1], 20) if instrument + 1 < len(inst) else len(ibag)
            if iblo >= ibhi or ibhi > len(ibag):
                continue
            iglob, ifirst = default_gens(), True
            for ibi in range(iblo, ibhi):
                glo = u16(ibag[ibi], 0)
                ghi = u16(ibag[ibi + 1], 0) if ibi + 1 < len(ibag) else len(igen)
                if glo > ghi or ghi > len(igen):
                    continue
                ig, sid = list(iglob), None
                for gi in range(glo, ghi):
                    op, amt = u16(igen[gi], 0), i16(igen[gi], 2)
                    if op == 53:
                        sid = amt & 0xFFFF
                    elif op < GEN_COUNT:
                        ig[op] = amt
                if sid is None:
                    if ifirst:
                        iglob = ig
                    ifirst = False
                    continue
                ifirst = False
                if sid >= len(shdrs):
                    continue
                sh = shdrs[sid]
                if sh["type"] & 0x8000 or sh["end"] <= sh["start"] or sh["end"] > smpl_frames:
                    continue
                ikey, ivel = range_of(ig, True, 43), range_of(ig, True, 44)
                key = (max(pkey[0], ikey[0]), min(pkey[1], ikey[1]))
                vel = (max(pvel[0], ivel[0]), min(pvel[1], ivel[1]))
                if key[0] > key[1] or vel[0] > vel[1]:
                    continue
                g = [sat16(ig[op] + pg[op]) if ps[op] and op not in ABSOLUTE_ONLY else ig[op]
                     for op in range(GEN_COUNT)]
                used.add(sid)
                pre_regions.append(len(regions))
                regions.append(region_of(g, sid, key, vel))
        if len(pre_regions) > start:
            presets.append((u16(phdr[pi], 22), u16(phdr[pi], 20), start, len(pre_regions) - start))
    if not presets:
        raise ValueError(f"{path}: no usable presets")
    # pool, src/sf2.rs:585-703
    pool_rate = sample_rate if resample_pool else 0
    if pool_rate:
        while True:
            est = sum(int((shdrs[s]["end"] - shdrs[s]["start"]) * float(pool_rate)
                          / float(max(shdrs[s]["rate"], 1))) + POOL_GUARD for s in used)
            if est * 2 <= budget or pool_rate <= 8000:
                break
            pool_rate //= 2
    pool, samples, remap = [], [], {}
    at = 0
    for sid in sorted(used):
        sh = shdrs[sid]
        data = np.frombuffer(raw, "<i2", sh["end"] - sh["start"], smpl_off + sh["start"] * 2)
        src_rate = max(sh["rate"], 1)
        ls, le = max(sh["sloop"] - sh["start"], 0), max(sh["eloop"] - sh["start"], 0)
        if pool_rate and pool_rate != src_rate:
            ratio = pool_rate / src_rate
            data = resample_i16(data, ratio)
            ls, le = int(round_half_away(ls * ratio)), int(round_half_away(le * ratio))
        n = len(data)
        ls, le = min(ls, max(n - 1, 0)), min(le, n)
        remap[sid] = len(samples)
        samples.append(dict(start=at, len=n, loop_start=ls, loop_end=le,
                            rate=pool_rate if pool_rate else src_rate,
                            correction_cents=F(sh["corr"]),
                            resample_ratio=F(pool_rate) / F(src_rate) if pool_rate else F(1),
                            root_key=min(sh["pitch"], 127)))
        pool.append(np.asarray(data, np.int16))
        pool.append(np.zeros(POOL_GUARD, np.int16))
        at += n + POOL_GUARD
    for r in regions:
        r["sample"] = remap.get(r["sample"], 0xFFFFFFFF)
    return (np.concatenate(pool) if pool else np.zeros(0, np.int16), samples, regions, presets,
            pre_regions, pool_rate)


def round_half_away(x: float) -> float:
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def sinc_table(ratio: float) -> np.ndarray:
    """SincBank::new, src/resample.rs:18-49: (512, 32) float32."""
    taps, phases = 32, 512
    cutoff = (ratio if ratio < 1.0 else 1.0) * 0.95
    half = taps / 2.0
    table = np.zeros((phases, taps), np.float32)
    for p in range(phases):
        frac = p / phases
        s_ = 0.0
        for k in range(taps):
            x = (k - half + 1.0 - frac) * cutoff
            s = 1.0 if abs(x) < 1e-9 else math.sin(math.pi * x) / (math.pi * x)
            s *= cutoff
            t = min(max((k - frac + 1.0) / taps, 0.0), 1.0)
            w = 0.42 - 0.5 * math.cos(2.0 * math.pi * t) + 0.08 * math.cos(4.0 * math.pi * t)
            table[p, k] = F(s * w)
            s_ += s * w
        if abs(s_) > 1e-9:
            table[p] = table[p] * F(1.0 / s_)
    return table


def resample_i16(src, ratio: float) -> np.ndarray:
    """resample_i16 / SincBank::process, src/resample.rs:51-108."""
    src = np.asarray(src, np.int16)
    if abs(ratio - 1.0) < 1e-12:
        return src.copy()
    if src.size == 0:
        return np.zeros(0, np.int16)
    table = sinc_table(ratio)
    n = src.size
    out_len = max(int(round_half_away(n * ratio)), 1)
    j = np.arange(out_len, dtype=np.float64)
    t = j * (1.0 / ratio)
    i = np.floor(t)
    p = np.minimum((( t - i) * 512).astype(np.int64), 511)
    base = i.astype(np.int64) - 16 + 1
    srcf = src.astype(np.float32)
    acc = np.zeros(out_len, np.float32)
    for k in range(32):
        idx = np.clip(base + k, 0, n - 1)
        acc = acc + table[p, k] * srcf[idx]
    return np.clip(acc, F(-32768), F(32767)).astype(np.int16)
