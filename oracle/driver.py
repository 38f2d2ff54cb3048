"""TEST INFRASTRUCTURE ONLY: a pure-Python restatement of the reference's driver
block loop -- src/driver.rs with the bank it builds voices from (src/bank.rs),
the note-off gate and controller rows (src/voice.rs:212-658) and the SMF reader
and tempo clock (src/midi.rs) -- for checking the library's native driver
block by block on small files. Never imported by the product.

Arithmetic: Python ints for the integer work (masked where the reference wraps),
Python floats for the f64 work (one rounding per operation, as Rust's f64), and
numpy.float32 scalars for every f32 value. The transcendental functions are
glibc's, called through ctypes, as the reference's Rust reaches them on Linux:
`2f32.powf(x)` / `2f64.powf(x)` as exp2f / exp2 (LLVM rewrites a power of the
constant 2 into exp2), `10f32.powf(x)` as powf, sin/cos/log10 directly.

One choice is the restatement's own: where loudness ranking keeps the `q` best
candidates of a time stratum, the reference leaves their order to its selection
routine ("only has to be deterministic", src/driver.rs:718-733); here, as in
the library, the winners keep their arrival order. The admitted set is the
reference's either way.
"""
from __future__ import annotations

import ctypes
import heapq
import math

import numpy as np

from . import post

F = np.float32
M32 = 0xFFFFFFFF

_libm = ctypes.CDLL("libm.so.6")


def _bind(name, res, args):
    fn = getattr(_libm, name)
    fn.restype, fn.argtypes = res, args
    return fn


_exp2f = _bind("exp2f", ctypes.c_float, [ctypes.c_float])
_powf = _bind("powf", ctypes.c_float, [ctypes.c_float, ctypes.c_float])
_log10f = _bind("log10f", ctypes.c_float, [ctypes.c_float])
_cosf = _bind("cosf", ctypes.c_float, [ctypes.c_float])
_sinf = _bind("sinf", ctypes.c_float, [ctypes.c_float])
_roundf = _bind("roundf", ctypes.c_float, [ctypes.c_float])
_exp2 = _bind("exp2", ctypes.c_double, [ctypes.c_double])
_sin = _bind("sin", ctypes.c_double, [ctypes.c_double])


def exp2f(x): return F(_exp2f(float(x)))
def powf(x, y): return F(_powf(float(x), float(y)))
def log10f(x): return F(_log10f(float(x)))
def cosf(x): return F(_cosf(float(x)))
def sinf(x): return F(_sinf(float(x)))
def roundf(x): return F(_roundf(float(x)))
def fmax(a, b): return b if a != a else (a if b != b else (a if a >= b else b))
def fmin(a, b): return b if a != a else (a if b != b else (a if a <= b else b))
def clamp(v, lo, hi): return lo if v < lo else (hi if v > hi else v)


def f32_bits(x) -> int:
    return int(np.array([x], np.float32).view(np.uint32)[0])


def bits_f32(u: int):
    return np.array([u & M32], np.uint32).view(np.float32)[0]


def as_u32(x: float) -> int:
    """Rust's saturating `f as u32`."""
    if not x > 0:
        return 0
    return M32 if x >= 4294967296.0 else int(x)


def as_u64(x: float) -> int:
    if not x > 0:
        return 0
    return (1 << 64) - 1 if x >= 18446744073709551616.0 else int(x)


def as_i64(x: float) -> int:
    if x != x:
        return 0
    return max(min(int(x), (1 << 63) - 1), -(1 << 63)) if math.isfinite(x) else (
        (1 << 63) - 1 if x > 0 else -(1 << 63))


# ---------------------------------------------------------------------------
# src/bank.rs
# ---------------------------------------------------------------------------
PI32 = F(math.pi)


def time_mod(v: int):
    """TimeMod::from_controller, src/bank.rs:243-252 -> (scale, add)."""
    d = F(v) - F(64)
    if d <= 0:
        return exp2f(d / F(8.0)), F(0)
    return F(1), F(6.0e-5) * d * d * d


def param_mod(cc71, cc72, cc73, cc74, cc75) -> dict:
    """ParamMod::from_controllers, src/bank.rs:336-350."""
    return {"cutoff": (F(cc74) - F(64)) * F(75.0),
            "q_cb": fmax(F(cc71) - F(64), F(0)) * (F(240.0) / F(63.0)),
            "attack": time_mod(cc73), "decay": time_mod(cc75), "release": time_mod(cc72)}


NEUTRAL = param_mod(64, 64, 64, 64, 64)


def velocity_atten_cb(vel: int, veltrack):
    """src/bank.rs:492-511"""
    v = F(max(vel, 1)) / F(127.0)
    t = clamp(F(veltrack), F(-100), F(100)) / F(100)
    if t == F(1):
        return clamp(F(-400.0) * log10f(v), F(0), F(960))
    gain = F(1) - t * (F(1) - v * v) if t >= 0 else F(1) + t * v * v
    if gain <= 0:
        return F(960)
    return clamp(F(-200.0) * log10f(gain), F(0), F(960))


def cb_to_gain(cb):
    return powf(F(10.0), -cb / F(200.0))


def biquad_lowpass(fc, q_cb, sr):
    """src/bank.rs:963-979"""
    q_db = q_cb / F(10.0) - F(3.01)
    q = fmax(powf(F(10.0), q_db / F(20.0)), F(0.001))
    gain = np.sqrt(F(1 / math.sqrt(2)) / q, dtype=np.float32)
    w0 = F(2.0) * PI32 * clamp(fc / sr, F(1e-5), F(0.49))
    s, c = sinf(w0), cosf(w0)
    alpha = s / (F(2.0) * q)
    a0 = F(1) + alpha
    return (gain * (F(1) - c) * F(0.5) / a0, gain * (F(1) - c) / a0, F(-2.0) * c / a0,
            (F(1) - alpha) / a0)


class Bank:
    """src/bank.rs: layout, variants, admission tables, preview and build."""

    def __init__(self, settings: dict, samples, regions, presets, preset_regions, pool_rate=0):
        self.s = settings
        self.sr = settings["sample_rate"]
        self.pool_rate = pool_rate
        self.samples = [{n: samples[n][i] for n in samples.dtype.names} for i in range(len(samples))]
        self.regions = [{n: regions[n][i] for n in regions.dtype.names} for i in range(len(regions))]
        self.presets = []
        index = []
        for p, d in enumerate(presets):
            regs = [int(x) for x in preset_regions[int(d["first"]):int(d["first"]) + int(d["count"])]]
            per_key = [[] for _ in range(128)]
            for ri in regs:
                r = self.regions[ri]
                for k in range(int(r["key_lo"]), min(int(r["key_hi"]), 127) + 1):
                    per_key[k].append(ri)
            self.presets.append(per_key)
            index.append(((int(d["bank"]), int(d["program"])), p))
        self.index = sorted(index, key=lambda e: e[0])  # stable
        # layout, src/bank.rs:689-721
        n = 0
        self.layout = []
        for r in self.regions:
            per_key = int(r["keynum_to_decay"]) != 0 or int(r["keynum_to_hold"]) != 0
            span = min(max(int(r["vel_hi"]) - int(r["vel_lo"]), 0) + 1, 128) \
                if F(r["filter_veltrack_cents"]) != 0 else 1
            self.layout.append((n, 1 if per_key else 0, span))
            n += span * (128 if per_key else 1)
        self.params = self.build_variant(NEUTRAL)

    # -- pitch and gain, shared by the preview tables and build_voice --
    def _ratio(self, r, s, key):
        eff_key = int(r["fixed_key"]) & 0xFF if r["fixed_key"] >= 0 else key
        root = float(r["root_key_override"]) if r["root_key_override"] >= 0 else float(s["root_key"])
        cents = ((float(eff_key) - root) * float(r["scale_tuning"]) + float(r["coarse_tune"]) * 100.0
                 + float(r["fine_tune"]) + float(F(s["correction_cents"])))
        ratio = _exp2(cents / 1200.0)
        rate = float(s["rate"]) if self.pool_rate == 0 else float(self.pool_rate)
        return ratio * (rate / float(self.sr)), eff_key

    def _gains(self, r, vel):
        eff_vel = int(r["fixed_vel"]) & 0xFF if r["fixed_vel"] >= 0 else vel
        atten = F(r["attenuation_cb"]) + velocity_atten_cb(eff_vel, r["amp_veltrack"])
        gain = cb_to_gain(atten) * F(self.s["master_volume"])
        theta = (clamp(F(r["pan"]), F(-1), F(1)) + F(1)) * F(0.5) * (PI32 * F(0.5))
        return gain * cosf(theta), gain * sinf(theta)

    def _sample(self, r):
        i = int(r["sample"])
        return self.samples[i] if i < len(self.samples) else None

    def find_preset(self, bank, program):
        """src/bank.rs:529-546 (first match where a pair repeats)."""
        for want in ((bank, program), (0, program)):
            for key, p in self.index:
                if key == want:
                    return p
        if bank == 128:
            for key, p in self.index:
                if key[0] == 128:
                    return p
        return self.index[0][1] if self.index else -1

    def preview_note_on(self, preset, key, vel, max_layers):
        """src/bank.rs:793-834: the layers note_on would build, with their gains."""
        if not 0 <= preset < len(self.presets):
            return []
        key = min(key, 127)
        out = []
        for ri in self.presets[preset][key]:
            if len(out) >= max_layers:
                break
            r = self.regions[ri]
            if vel < r["vel_lo"] or vel > r["vel_hi"]:
                continue
            s = self._sample(r)
            if s is None or s["len"] == 0:
                continue
            ratio, _ = self._ratio(r, s, key)
            if not (math.isfinite(ratio) and ratio > 0):
                continue
            gl, gr = self._gains(r, vel)
            out.append((ri, gl, gr, as_u32(float(F(r["delay"]) * F(self.sr)))))
        return out

    def build_layer(self, ri, key, vel):
        """src/bank.rs:585-687 -> dict or None."""
        if ri >= len(self.regions):
            return None
        r = self.regions[ri]
        key = min(key, 127)
        s = self._sample(r)
        if s is None or s["len"] == 0:
            return None
        ratio, eff_key = self._ratio(r, s, key)
        if not (math.isfinite(ratio) and ratio > 0):
            return None
        gl, gr = self._gains(r, vel)
        L = int(s["len"])

        def scale(v):
            return as_i64(float(roundf(F(v) * F(s["resample_ratio"]))))
        start = clamp(scale(int(r["addr_start"])), 0, L - 1)
        smp_len = clamp(L + scale(int(r["addr_end"])), 1, L)
        ls = clamp(int(s["loop_start"]) + scale(int(r["addr_loop_start"])), 0, smp_len - 1)
        le = clamp(int(s["loop_end"]) + scale(int(r["addr_loop_end"])), 0, smp_len)
        base, stride, span = self.layout[ri]
        vi = min(max(vel - int(r["vel_lo"]), 0), span - 1) if span > 1 else 0
        params = base + vi * (128 if stride else 1) + (eff_key if stride else 0)
        mode = int(r["loop_mode"])
        flags = {0: 0, 1: 1, 2: 3}[mode]
        if mode != 0 and le <= ls + 1:
            flags = 0
        return {"phase": start << 32, "step": as_u64(ratio * 4294967296.0),
                "smp_base": int(s["start"]), "smp_len": smp_len, "loop_start": ls,
                "loop_end": le, "flags": flags, "params": params, "region": ri,
                "gain_l": gl, "gain_r": gr, "delay": as_u32(float(F(r["delay"]) * F(self.sr)))}

    def build_variant(self, m) -> list:
        """src/bank.rs:836-853: one RegionParams tuple per entry."""
        out = []
        for ri, r in enumerate(self.regions):
            _, stride, span = self.layout[ri]
            for v in range(span):
                vel = min(int(r["vel_lo"]) + v, 127)
                keys = range(128) if stride else (60,)
                out.extend(self.make_params(r, k, vel, m) for k in keys)
        return out

    def make_params(self, r, key, vel, m):
        """src/bank.rs:876-955 -> (attack_rate, attack_end, decay_coef, decay_target,
        sustain, release_coef, b0, b1, a1, a2, flags)."""
        sr = F(self.sr)
        kd = F(60) - F(key)
        hold = F(r["hold"]) * exp2f(F(int(r["keynum_to_hold"])) * kd / F(1200))
        decay = F(r["decay"]) * exp2f(F(int(r["keynum_to_decay"])) * kd / F(1200))

        def floor_t(t, base):
            return fmax(t, fmin(F(0.004), base))
        (asc, aad), (dsc, dad), (rsc, rad) = m["attack"], m["decay"], m["release"]
        attack_frames = fmax((F(r["attack"]) * asc + aad) * sr, F(0))
        hold_frames = fmax(hold * sr, F(0))
        decay_frames = fmax(floor_t(decay * dsc + dad, decay) * sr, F(1))
        release = F(r["release"])
        release_frames = fmax(floor_t(release * rsc + rad, release) * sr, F(1))
        rate = F(1) if attack_frames < 1 else F(1) / attack_frames
        end = F(1) + hold_frames * rate
        sustain = clamp(F(r["sustain"]), F(0), F(1))

        def coef(curve, frames):
            return powf(F(10.0), F(-100.0) / (F(20.0) * frames)) if curve == 0 else F(1) / frames
        fc_cents = F(r["filter_fc_cents"]) + F(r["filter_veltrack_cents"]) * (F(vel) / F(127)) \
            + m["cutoff"]
        fc = F(8.176) * exp2f(fc_cents / F(1200))
        use = bool(self.s["filter_enabled"]) and fc_cents < 13500 and fc < sr * F(0.49) and fc > 20
        bq = biquad_lowpass(fc, F(r["filter_q_cb"]) + m["q_cb"], sr) if use else (F(1), F(0), F(0), F(0))
        return (rate, end, coef(self.s["decay_curve"], decay_frames),
                fmax(sustain, F(self.s["env_floor"])), sustain,
                coef(self.s["release_curve"], release_frames), *bq, 1 if use else 0)


def params_array(rows, dtype) -> np.ndarray:
    out = np.zeros(len(rows), dtype)
    names = ("attack_rate", "attack_end", "decay_coef", "decay_target", "sustain", "release_coef",
             "b0", "b1", "a1", "a2", "flags")
    for j, n in enumerate(names):
        out[n] = [row[j] for row in rows]
    return out


# ---------------------------------------------------------------------------
# src/midi.rs
# ---------------------------------------------------------------------------
NOTE_ON, NOTE_OFF, CC, PROGRAM, BEND, TEMPO, OTHER = range(7)


def _track_events(data: bytes):
    """TrackReader::next_event over one chunk: yields (tick, event tuple)."""
    pos, tick, running = 0, 0, 0

    def byte():
        nonlocal pos
        if pos >= len(data):
            raise EOFError
        pos += 1
        return data[pos - 1]

    def varlen():
        v = 0
        for _ in range(8):
            b = byte()
            v = (v << 7) | (b & 0x7F)
            if not b & 0x80:
                return v
        raise EOFError
    try:
        while True:
            tick += varlen()
            status = byte()
            if status < 0x80:
                first, status = status, running
                if status < 0x80:
                    return
            else:
                if status < 0xF0:
                    running = status
                elif status == 0xFF:
                    running = 0
                first = 0 if status in (0xFF, 0xF0, 0xF7) else byte()
            ch, hi = status & 0x0F, status & 0xF0
            if hi == 0x80:
                byte()
                yield tick, (NOTE_OFF, ch, first & 0x7F, 0, 0)
            elif hi == 0x90:
                v = byte()
                yield tick, ((NOTE_OFF, ch, first & 0x7F, 0, 0) if v == 0
                             else (NOTE_ON, ch, first & 0x7F, v & 0x7F, 0))
            elif hi == 0xA0:
                byte()
                yield tick, (OTHER, 0, 0, 0, 0)
            elif hi == 0xB0:
                v = byte()
                yield tick, (CC, ch, first & 0x7F, v & 0x7F, 0)
            elif hi == 0xC0:
                yield tick, (PROGRAM, ch, first & 0x7F, 0, 0)
            elif hi == 0xD0:
                yield tick, (OTHER, 0, 0, 0, 0)
            elif hi == 0xE0:
                msb = byte()
                yield tick, (BEND, ch, 0, 0, (((msb & 0x7F) << 7) | (first & 0x7F)) - 8192)
            elif status == 0xFF:
                meta, ln = byte(), varlen()
                if meta == 0x2F:
                    if pos + ln > len(data):
                        return
                    yield tick, (OTHER, 0, 0, 0, 0)
                    return
                if meta == 0x51 and ln == 3:
                    a, b, c = byte(), byte(), byte()
                    yield tick, (TEMPO, 0, 0, 0, (a << 16) | (b << 8) | c)
                    continue
                if pos + ln > len(data):
                    return
                pos += ln
                yield tick, (OTHER, 0, 0, 0, 0)
            elif status in (0xF0, 0xF7):
                ln = varlen()
                if pos + ln > len(data):
                    return
                pos += ln
                yield tick, (OTHER, 0, 0, 0, 0)
            else:
                yield tick, (OTHER, 0, 0, 0, 0)
    except EOFError:
        return


def read_midi(path: str):
    """MidiStream::open + next (src/midi.rs:259-361): (division, [(tick, event)])."""
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < 14:
        raise ValueError(f"{path}: too short for an SMF header")
    if raw[:4] != b"MThd":
        raise ValueError(f"{path}: not a standard MIDI file")
    hdr_len = int.from_bytes(raw[4:8], "big")
    div_raw = int.from_bytes(raw[12:14], "big", signed=True)
    if div_raw > 0:
        division = ("ppq", div_raw)
    else:
        division = ("smpte", (-(div_raw >> 8)) & 0xFF, div_raw & 0xFF)
    tracks, off = [], 8 + hdr_len
    while off + 8 <= len(raw):
        tag, ln = raw[off:off + 4], int.from_bytes(raw[off + 4:off + 8], "big")
        start = off + 8
        ln = min(ln, max(len(raw) - start, 0))
        if tag == b"MTrk":
            tracks.append(raw[start:start + ln])
        elif ln == 0:
            break
        off = start + ln
    if not tracks:
        raise ValueError(f"{path}: no MTrk chunks")
    its = [_track_events(t) for t in tracks]
    heap, pend = [], [None] * len(its)
    for i, it in enumerate(its):
        e = next(it, None)
        if e is not None:
            pend[i] = e[1]
            heapq.heappush(heap, (e[0], i))
    out = []
    while heap:
        tick, i = heapq.heappop(heap)
        out.append((tick, pend[i]))
        e = next(its[i], None)
        if e is not None:
            pend[i] = e[1]
            heapq.heappush(heap, (e[0], i))
    return division, out


class TempoClock:
    """src/midi.rs:363-418"""

    def __init__(self, division, sample_rate):
        self.div, self.sr = division, float(sample_rate)
        self.base_tick, self.base_frame = 0, 0.0
        self._raw(500_000)

    def _raw(self, us):
        if self.div[0] == "ppq":
            self.fpt = (us * 1e-6) * self.sr / float(max(self.div[1], 1))
        else:
            self.fpt = self.sr / (float(max(self.div[1], 1)) * float(max(self.div[2], 1)))

    def set_tempo(self, tick, us):
        if self.div[0] != "ppq":
            return
        self.base_frame = self.frame_at(tick)
        self.base_tick = tick
        self._raw(us)

    def frame_at(self, tick):
        return self.base_frame + float(tick - self.base_tick) * self.fpt


# ---------------------------------------------------------------------------
# src/voice.rs: gate and controller rows
# ---------------------------------------------------------------------------
class Gates:
    """GateTable, src/voice.rs:219-452."""

    def __init__(self, tiles, tile_frames):
        self.tiles, self.tf = tiles, tile_frames
        self.rows = np.zeros((tiles, 2048), np.uint32)
        self.on = [0] * 2048
        self.off = [0] * 2048
        self.pend = [0] * 2048
        self.held = [0] * 2048
        self.sus = self.sost = 0
        self.cursor = 0

    @staticmethod
    def slot(ch, key):
        return (ch & 15) * 128 + (key & 127)

    def _advance(self, frame):
        tile = frame // self.tf
        while self.cursor <= tile and self.cursor < self.tiles:
            self.rows[self.cursor] = self.off
            self.cursor += 1

    def end_block(self):
        self._advance(self.tiles * self.tf)

    def sustained(self, ch):
        return bool(self.sus >> (ch & 15) & 1)

    def note_on(self, ch, key, frame):
        self._advance(frame)
        s = self.slot(ch, key)
        self.on[s] = (self.on[s] + 1) & M32
        return self.on[s]

    def note_off(self, ch, key, frame):
        self._advance(frame)
        s = self.slot(ch, key)
        if self.sost >> (ch & 15) & 1 and self.held[s] > 0:
            if (self.off[s] + self.pend[s]) & M32 != self.on[s]:
                self.pend[s] += 1
                self.held[s] -= 1
            return
        if self.sustained(ch):
            if (self.off[s] + self.pend[s]) & M32 != self.on[s]:
                self.pend[s] += 1
        elif self.off[s] != self.on[s]:
            self.off[s] = (self.off[s] + 1) & M32

    def _flush(self, ch):
        for s in range((ch & 15) * 128, (ch & 15) * 128 + 128):
            if self.pend[s]:
                self.off[s] = (self.off[s] + self.pend[s]) & M32
                self.pend[s] = 0

    def set_sustain(self, ch, down, frame):
        if down == self.sustained(ch):
            return
        self._advance(frame)
        bit = 1 << (ch & 15)
        if down:
            self.sus |= bit
            return
        self.sus &= ~bit
        if self.sost & bit:
            return
        self._flush(ch)

    def set_sostenuto(self, ch, down, frame):
        bit = 1 << (ch & 15)
        if down == bool(self.sost & bit):
            return
        self._advance(frame)
        base = (ch & 15) * 128
        if down:
            self.sost |= bit
            for s in range(base, base + 128):
                self.held[s] = (self.on[s] - self.off[s] - self.pend[s]) & M32
            return
        self.sost &= ~bit
        for s in range(base, base + 128):
            self.held[s] = 0
        if self.sustained(ch):
            return
        self._flush(ch)

    def all_notes_off(self, ch, frame):
        self._advance(frame)
        base = (ch & 15) * 128
        if self.sustained(ch):
            for s in range(base, base + 128):
                self.pend[s] = (self.on[s] - self.off[s]) & M32
            return
        for s in range(base, base + 128):
            self.off[s] = self.on[s]

    def all_sound_off(self, ch, frame):
        self._advance(frame)
        for s in range((ch & 15) * 128, (ch & 15) * 128 + 128):
            self.off[s] = self.on[s]
            self.pend[s] = 0
            self.held[s] = 0

    def reset_controllers(self, ch, frame):
        self.set_sostenuto(ch, False, frame)
        self.set_sustain(ch, False, frame)

    def sounding(self):
        return sum((a - b) & M32 for a, b in zip(self.on, self.off))


BEND_ONE = 1 << 24
ONE_BITS = f32_bits(F(1))


class Chans:
    """ChannelTable, src/voice.rs:481-658: [bend, gain_l, gain_r, variant] x 16 per tile."""

    def __init__(self, tiles, tile_frames):
        self.tiles, self.tf = tiles, tile_frames
        self.now = [BEND_ONE, ONE_BITS, ONE_BITS, 0] * 16
        self.rows = np.tile(np.array(self.now, np.uint32), (tiles, 1))
        self.cursor = 0
        self.flags = (False, False, False)

    def _advance(self, frame):
        tile = frame // self.tf
        while self.cursor <= tile and self.cursor < self.tiles:
            self.rows[self.cursor] = self.now
            self.cursor += 1

    def end_block(self):
        self._advance(self.tiles * self.tf)

    def _set(self, i, v, frame):
        if self.now[i] == v:
            return
        self._advance(frame)
        self.now[i] = v

    def set_bend(self, ch, factor, frame):
        self._set((ch & 15) * 4, factor, frame)

    def set_variant(self, ch, v, frame):
        self._set((ch & 15) * 4 + 3, v, frame)

    def set_gain(self, ch, l, r, frame):
        b = (ch & 15) * 4
        lb, rb = f32_bits(l), f32_bits(r)
        if self.now[b + 1] == lb and self.now[b + 2] == rb:
            return
        self._advance(frame)
        self.now[b + 1], self.now[b + 2] = lb, rb

    def modulate_bend(self, ch, tile, factor):
        v = (int(self.rows[tile, (ch & 15) * 4]) * factor >> 24) & M32
        self.rows[tile, (ch & 15) * 4] = max(v, 1)

    def modulate_gain(self, ch, tile, factor):
        for f in (1, 2):
            i = (ch & 15) * 4 + f
            self.rows[tile, i] = f32_bits(bits_f32(int(self.rows[tile, i])) * factor)

    def refresh_active(self):
        e = self.rows.reshape(-1, 4)
        self.flags = (bool((e[:, 0] != BEND_ONE).any()),
                      bool(((e[:, 1] != ONE_BITS) | (e[:, 2] != ONE_BITS)).any()),
                      bool((e[:, 3] != 0).any()))


# ---------------------------------------------------------------------------
# src/driver.rs
# ---------------------------------------------------------------------------
def bend_factor(semitones: float) -> int:
    """src/fixed.rs:99-112"""
    if not math.isfinite(semitones):
        return BEND_ONE
    f = _exp2(semitones / 12.0) * float(BEND_ONE)
    if f <= 0:
        return 1
    if f >= 4294967295.0:
        return M32
    return int(math.floor(f + 0.5))  # round half away from zero, f > 0


def mix48(i: int) -> int:
    """src/voice.rs:168-177"""
    m = (1 << 48) - 1
    z = i & m
    z = (z ^ (z >> 24)) & m
    z = (z * 0x9E3779B97F4B) & m
    z = (z ^ (z >> 23)) & m
    z = (z * 0xBF58476D1CE5) & m
    return (z ^ (z >> 25)) & m


def rank_gain_q(gl, gr) -> int:
    """src/voice.rs:203-210"""
    g = F(abs(gl)) + F(abs(gr))
    if g != g or g <= 0:
        return 0
    return (f32_bits(g) >> 16) & 0x7FFF


def rank_key(gl, gr, index) -> int:
    return (rank_gain_q(gl, gr) << 48) | mix48(index)


def admit_take(cfg, live, queued) -> int:
    """Config::admit_take, src/config.rs:120-131"""
    cap = cfg.max_voices
    want = min(queued, cap)
    if live + want <= cap:
        return want
    ms = max((cfg.max_voices * cfg.max_steal_percent // 100) & M32, 1)
    need = 0 if cfg.steal_rule == 1 else min(live + want - cap, live, ms)
    return min(want, cap - (live - need))


DEFERRED = -1


class Driver:
    """Driver (src/driver.rs:240-1192), split as begin / admit / end."""

    def __init__(self, cfg, bank: Bank, division, events):
        self.cfg, self.bank = cfg, bank
        self.events = list(events)  # [(tick, (kind, ch, a, b, value))]
        self.at = 0
        self.clock = TempoClock(division, cfg.sample_rate)
        tiles = cfg.block_frames // cfg.gate_frames
        self.gates = Gates(tiles, cfg.gate_frames)
        self.chan = Chans(tiles, cfg.gate_frames)
        self.bank_msb = [0] * 16
        self.bend_val = [0] * 16
        self.bend_range = [2.0] * 16
        self.rpn = [0] * 16
        self.vol = [min(cfg.default_channel_volume, 127)] * 16
        self.expr = [127] * 16
        self.pan = [64] * 16
        self.sound = [[64] * 5 for _ in range(16)]
        self.mod = [0] * 16
        self.vib_rate = [64] * 16
        self.vib_depth = [64] * 16
        self.trem = [0] * 16
        self.soft = [0] * 16
        self.lsb_mod = [0] * 16
        self.lsb_vol = [0] * 16
        self.lsb_pan = [0] * 16
        self.lsb_expr = [0] * 16
        self.lfo = [0.0] * 16
        self.variants = [(32,) * 5]
        self.seen_states = [(32,) * 5]
        self.cur_variant = [0] * 16
        self.variant_used = 1
        self.variant_seen = [0]
        self.variant_clock = 0
        self.pending_variants = []
        self.spawn_variants = False
        self.next_id = 1
        self.block_index = 0
        self.pending = None
        self.stream_done = self.end_sent = False
        self.tail_left = int(60.0 * cfg.sample_rate / cfg.block_frames)
        self.deferred = []
        self.stats = dict(param_variants=0, variant_fallbacks=0, variant_states=0,
                          variant_rebuilds=0, frames=0, blocks=0, events=0, dropped=0, notes=0,
                          voices_spawned=0, peak=F(0), clipped=0, last_live=0, last_want=0,
                          last_take=0, last_stolen=0, last_want_energy=0, last_take_energy=0)
        self.preset = [0] * 16
        for c in range(16):
            self._refresh_preset(c, 0)
        self.limiter = None
        if cfg.limiter and cfg.limiter_mode == 1:
            self.limiter = post.Omni(cfg.sample_rate)
        elif cfg.limiter and cfg.limiter_mode == 2:
            self.limiter = post.Brickwall(cfg.sample_rate, 10.0 ** (cfg.limiter_ceiling_db / 20.0),
                                          cfg.limiter_lookahead_ms, cfg.limiter_release_ms,
                                          cfg.limiter_sustain_ms, cfg.limiter_true_peak)

    def _refresh_preset(self, c, program):
        p = self.bank.find_preset(128 if c == 9 else self.bank_msb[c], program)
        self.preset[c] = max(p, 0)

    def _next_event(self):
        if self.at >= len(self.events):
            return None
        self.at += 1
        return self.events[self.at - 1]

    # -- begin: src/driver.rs:450-555 --
    def begin_block(self):
        F_ = self.cfg.block_frames
        start, end = self.block_index * F_, (self.block_index + 1) * F_
        self.gates.cursor = 0
        self.chan.cursor = 0
        self.variant_used = 1
        for v in self.cur_variant:
            self.variant_used |= 1 << v
        self.spawn = []
        self.notes = []
        self.cands = []  # [key, note, region, id_off]
        self.deferred_now = []
        self.first_id = self.next_id
        self.spawn_variants = False
        i = 0
        while i < len(self.deferred):
            if self.deferred[i][0] < end:
                frame, cmd = self.deferred[i]
                self.deferred[i] = self.deferred[-1]
                self.deferred.pop()
                cmd = dict(cmd, start_rel=max(frame - start, 0) & M32)
                if cmd["variant"]:
                    self.variant_used |= 1 << cmd["variant"]
                    self.spawn_variants = True
                self.cands.append([rank_key(cmd["gain_l"], cmd["gain_r"], len(self.cands)),
                                   DEFERRED, len(self.deferred_now), 0])
                self.deferred_now.append(cmd)
            else:
                i += 1
        while True:
            if self.pending is not None:
                tick, ev = self.pending
                self.pending = None
            else:
                e = self._next_event()
                if e is None:
                    self.stream_done = True
                    break
                tick, ev = e
            if ev[0] == TEMPO:
                self.clock.set_tempo(tick, ev[4])
                self.stats["events"] += 1
                continue
            frame = as_u64(self.clock.frame_at(tick))
            if frame >= end:
                self.pending = (tick, ev)
                break
            rel = min(max(frame - start, 0) & M32, F_ - 1)
            self.stats["events"] += 1
            self._event(ev, rel, start)
        if self.stream_done and not self.end_sent:
            for ch in range(16):
                self.gates.all_sound_off(ch, 0)
            self.end_sent = True
        uploads = [(i, self.bank.build_variant(m)) for i, m in self.pending_variants]
        self.pending_variants = []
        self.gates.end_block()
        self.chan.end_block()
        self._modulate()
        self.chan.refresh_active()
        b, g, v = self.chan.flags
        self.flags = (b, g, v or self.spawn_variants)
        return uploads

    # -- admission: src/driver.rs:556-594, 680-783 --
    def admit(self, live, stolen_before):
        live &= M32
        want = len(self.cands)
        take = admit_take(self.cfg, live, want)
        self.stats["dropped"] += want - take
        self.stolen_before = stolen_before
        self.stats.update(last_live=live, last_want=want, last_take=take)
        c = self.cands
        if 0 < take < want:
            if self.cfg.admit_rule == 1:
                for i in range(take):
                    c[i] = c[i * want // take]
            else:
                strata = min(take, 64)
                for s in range(strata):
                    lo, hi = s * want // strata, (s + 1) * want // strata
                    olo, ohi = s * take // strata, (s + 1) * take // strata
                    q = ohi - olo
                    if q == 0:
                        continue
                    if q < hi - lo:
                        best = set(sorted(range(lo, hi), key=lambda j: -c[j][0])[:q])
                        seg = [c[j] for j in range(lo, hi) if j in best] + \
                              [c[j] for j in range(lo, hi) if j not in best]
                        c[lo:hi] = seg
                    for j in range(q):
                        c[olo + j], c[lo + j] = c[lo + j], c[olo + j]
        self.stats["last_want_energy"] = sum((x[0] >> 48) & 0x7FFF for x in c)
        self.stats["last_take_energy"] = sum((x[0] >> 48) & 0x7FFF for x in c[:take])
        for key, note, region, id_off in c[:min(take, len(c))]:
            if note == DEFERRED:
                self.spawn.append(self.deferred_now[region])
                continue
            ch, k, vel, variant, rel, ordinal = self.notes[note]
            v = self.bank.build_layer(region, k, vel)
            if v is None:
                continue
            sr = rel if v["delay"] == 0 else (rel + v["delay"]) & M32
            self.spawn.append(self._cmd(v, variant, Gates.slot(ch, k), ordinal, sr,
                                        self.first_id + id_off))
        self.stats["voices_spawned"] += want
        return self.spawn

    @staticmethod
    def _cmd(v, variant, slot, ordinal, start_rel, nid):
        return {"phase_lo": v["phase"] & M32, "phase_hi": v["phase"] >> 32,
                "step_lo": v["step"] & M32, "step_hi": (v["step"] >> 32) & M32,
                "smp_base": v["smp_base"], "smp_len": v["smp_len"],
                "loop_start": v["loop_start"], "loop_end": v["loop_end"], "flags": v["flags"],
                "params": v["params"], "variant": variant, "region": v["region"],
                "gate_slot": slot, "ordinal": ordinal, "start_rel": start_rel,
                "note_id_lo": nid & M32, "note_id_hi": nid >> 32, "gain_l": v["gain_l"],
                "gain_r": v["gain_r"]}

    # -- end: src/driver.rs:596-641 --
    def end_block(self, out: np.ndarray, stolen_after_spawn, st):
        self.stats["last_stolen"] = stolen_after_spawn - self.stolen_before
        if self.limiter is not None:
            self.limiter.process(out)
        self.stats["clipped"] += post.clamp_block(out)
        self.stats["peak"] = fmax(self.stats["peak"], F(st["peak"]))
        self.stats["frames"] += self.cfg.block_frames
        self.stats["blocks"] += 1
        self.block_index += 1
        if not self.stream_done or self.deferred:
            return True
        if st["active_voices"] > 0:
            self.tail_left = max(self.tail_left - 1, 0)
            return self.tail_left != 0
        return False

    # -- events: src/driver.rs:785-978 --
    def _event(self, ev, rel, start):
        kind, ch, a, b, value = ev
        c = ch & 15
        if kind == NOTE_ON:
            ordinal = self.gates.note_on(ch, a, rel)
            self.stats["notes"] += 1
            variant = self.cur_variant[c]
            note = len(self.notes)
            recorded = False
            for region, gl, gr, delay in self.bank.preview_note_on(self.preset[c], a, b,
                                                                   self.cfg.max_layers):
                nid = self.next_id
                self.next_id += 1
                if delay and start + rel + delay >= start + self.cfg.block_frames:
                    v = self.bank.build_layer(region, a, b)
                    if v is not None:
                        self.deferred.append((start + rel + delay,
                                              self._cmd(v, variant, Gates.slot(ch, a), ordinal,
                                                        rel, nid)))
                    continue
                if not recorded:
                    self.notes.append((c, a, b, variant, rel, ordinal))
                    recorded = True
                self.cands.append([rank_key(gl, gr, len(self.cands)), note, region,
                                   nid - self.first_id])
                if variant:
                    self.variant_used |= 1 << variant
                    self.spawn_variants = True
        elif kind == NOTE_OFF:
            self.gates.note_off(ch, a, rel)
        elif kind == CC:
            self._cc(ch, a, b, rel)
        elif kind == PROGRAM:
            self._refresh_preset(c, a)
        elif kind == BEND:
            self.bend_val[c] = value
            self._bend(c, rel)

    def _cc(self, ch, num, val, rel):
        c = ch & 15
        rpn0 = self.rpn[c] == 0
        if num == 0:
            self.bank_msb[c] = val
        elif num == 1:
            self.mod[c], self.lsb_mod[c] = val, 0
        elif num == 33:
            self.lsb_mod[c] = val
        elif num in (39, 42, 43):
            {39: self.lsb_vol, 42: self.lsb_pan, 43: self.lsb_expr}[num][c] = val
            self._gain(c, rel)
        elif num == 32:
            pass  # bank select LSB: kept by the reference, read by nothing
        elif num == 6 and rpn0:
            self.bend_range[c] = float(val)
            self._bend(c, rel)
        elif num == 38 and rpn0:
            self.bend_range[c] = math.trunc(self.bend_range[c]) + val / 100.0
            self._bend(c, rel)
        elif num == 96 and rpn0:
            self.bend_range[c] = min(self.bend_range[c] + 1.0, 127.0)
            self._bend(c, rel)
        elif num == 97 and rpn0:
            self.bend_range[c] = max(self.bend_range[c] - 1.0, 0.0)
            self._bend(c, rel)
        elif num in (98, 99):
            self.rpn[c] = 0x4000
        elif num in (100, 101):
            keep = 0 if self.rpn[c] == 0x4000 else self.rpn[c]
            self.rpn[c] = (keep & 0x3F80) | val if num == 100 else (keep & 0x7F) | (val << 7)
        elif num in (7, 10, 11):
            msb, lsb = {7: (self.vol, self.lsb_vol), 10: (self.pan, self.lsb_pan),
                        11: (self.expr, self.lsb_expr)}[num]
            msb[c], lsb[c] = val, 0
            self._gain(c, rel)
        elif num == 64:
            self.gates.set_sustain(ch, val >= 64, rel)
        elif num == 66:
            self.gates.set_sostenuto(ch, val >= 64, rel)
        elif num == 67:
            self.soft[c] = val
            self._gain(c, rel)
        elif 71 <= num <= 75:
            self._sound(ch, num - 71, val, rel)
        elif num == 76:
            self.vib_rate[c] = val
        elif num == 77:
            self.vib_depth[c] = val
        elif num == 92:
            self.trem[c] = val
        elif num == 120:
            self.gates.all_sound_off(ch, rel)
        elif num == 121:
            self._reset(ch, rel)
        elif 123 <= num <= 127:
            self.gates.all_notes_off(ch, rel)

    def _sound(self, ch, which, val, rel):
        """src/driver.rs:980-1056"""
        c = ch & 15
        if self.sound[c][which] == val:
            return
        self.sound[c][which] = val
        want = tuple(x >> 1 for x in self.sound[c])
        if want not in self.seen_states:
            self.seen_states.append(want)
            self.stats["variant_states"] = len(self.seen_states)
        cap = min(max(self.cfg.max_param_variants, 1), 63)
        if want in self.variants:
            idx = self.variants.index(want)
        elif len(self.variants) < cap:
            idx = len(self.variants)
            self.variants.append(want)
            self.variant_seen.append(0)
            self.stats["param_variants"] = idx + 1
            self._queue_variant(idx, c)
        else:
            free = [i for i in range(1, len(self.variants)) if not self.variant_used >> i & 1]
            if free:
                idx = min(free, key=lambda i: (self.variant_seen[i], i))
                self.variants[idx] = want
                self._queue_variant(idx, c)
                self.stats["variant_rebuilds"] += 1
            else:
                self.stats["variant_fallbacks"] += 1
                dist = [sum(abs(p - q) for p, q in zip(v, want)) for v in self.variants]
                idx = dist.index(min(dist))
        self.variant_clock += 1
        self.variant_seen[idx] = self.variant_clock
        self.variant_used |= 1 << idx
        self.cur_variant[c] = idx
        self.chan.set_variant(ch, idx, rel)

    def _queue_variant(self, i, c):
        m = param_mod(*self.sound[c])
        self.pending_variants = [p for p in self.pending_variants if p[0] != i] + [(i, m)]

    def _reset(self, ch, rel):
        """src/driver.rs:1073-1108"""
        c = ch & 15
        self.bank_msb[c] = 0
        self.bend_range[c] = 2.0
        self.bend_val[c] = 0
        self.rpn[c] = 0x4000
        self.vol[c] = min(self.cfg.default_channel_volume, 127)
        self.expr[c], self.pan[c] = 127, 64
        for a in (self.soft, self.mod, self.lsb_mod, self.lsb_vol, self.lsb_pan, self.lsb_expr,
                  self.trem):
            a[c] = 0
        self.vib_rate[c] = self.vib_depth[c] = 64
        for w in range(5):
            self._sound(ch, w, 64, rel)
        self._bend(c, rel)
        self._gain(c, rel)
        self.gates.reset_controllers(ch, rel)

    def _modulate(self):
        """src/driver.rs:1110-1144"""
        tiles = self.chan.tiles
        ts = self.cfg.gate_frames / self.cfg.sample_rate
        for c in range(16):
            vib = (self.mod[c] * 128.0 + self.lsb_mod[c]) / 16383.0 * (self.vib_depth[c] / 64.0) * 50.0
            trem = self.trem[c] / 127.0 * 0.25
            rate = 5.0 * _exp2((self.vib_rate[c] - 64.0) / 32.0)
            if vib <= 0 and trem <= 0:
                x = self.lfo[c] + rate * ts * tiles
                self.lfo[c] = x - math.trunc(x)
                continue
            ph = self.lfo[c]
            for t in range(tiles):
                s = _sin(ph * 6.283185307179586)
                if vib > 0:
                    self.chan.modulate_bend(c, t, bend_factor(vib * s / 100.0))
                if trem > 0:
                    self.chan.modulate_gain(c, t, F(1.0 - trem + trem * s))
                x = ph + rate * ts
                ph = x - math.trunc(x)
            self.lfo[c] = ph

    def _bend(self, c, rel):
        self.chan.set_bend(c, bend_factor((self.bend_val[c] / 8192.0) * self.bend_range[c]), rel)

    def _gain(self, c, rel):
        """src/driver.rs:1161-1188"""
        def fine(msb, lsb):
            return F(msb) / F(127) if lsb == 0 else (F(msb) * F(128) + F(lsb)) / F(16383)
        v, e = fine(self.vol[c], self.lsb_vol[c]), fine(self.expr[c], self.lsb_expr[c])
        soft = F(1) - F(0.5) * (F(self.soft[c]) / F(127))
        amp = (v * v) * (e * e) * soft
        if self.pan[c] == 64 and self.lsb_pan[c] == 0:
            l = r = F(1)
        else:
            th = fine(self.pan[c], self.lsb_pan[c]) * F(math.pi / 2)
            l, r = cosf(th) * F(math.sqrt(2)), sinf(th) * F(math.sqrt(2))
        self.chan.set_gain(c, amp * l, amp * r, rel)


def spawn_array(cmds, dtype) -> np.ndarray:
    out = np.zeros(len(cmds), dtype)
    for n in dtype.names:
        out[n] = [c[n] for c in cmds]
    return out
