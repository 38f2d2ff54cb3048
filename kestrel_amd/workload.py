"""Synthetic banks, note streams and a driver-lite that turns them into the
per-block inputs of the `Backend` trait: gate rows, channel rows and spawn
commands, for the backend-only workloads; and the inputs of the native driver
(kestrel_amd/driver.py): a bank description and dense SMF files.

The driver-lite reproduces the parts of `Driver::next_block` (src/driver.rs:
434-642) the backend's inputs depend on, vectorised for millions of notes, so
the CPU oracle and the gfx950 backend are fed identical blocks:

* GateTable / ChannelTable row semantics (src/voice.rs:219-658), ported
  faithfully (`GateTable`, `ChannelTable`) and, for large note counts, built
  vectorised per block (`BlockBuilder`), which a test checks against the port.
* note ids in event order, gate slot `channel * 128 + key`, per-slot ordinals,
  `start_rel` = frame offset in the block (src/driver.rs:751-783, 785-967).
* `build_voice`'s pitch, gain and loop arithmetic (src/bank.rs:596-686) and
  `make_params`' envelope and filter design (src/bank.rs:876-979), in float32.

Admission ranking (`rank_candidates`, src/driver.rs:701-742) is the driver's
and is not reproduced: every note-on of a block is handed to `spawn`, and the
backend's own rule (prefix under AdmitRule::Loudest, `spawn_pick` under Even)
decides what an oversubscribed block keeps -- on both backends alike.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from .abi import PARAMS_DTYPE, SPAWN_DTYPE
from .config import Config

GATE_SLOTS = 2048
BEND_CHANNELS = 16
CHAN_FIELDS = 4
CHAN_BEND, CHAN_GAIN_L, CHAN_GAIN_R, CHAN_VARIANT = 0, 1, 2, 3
BEND_ONE = 1 << 24
ONE_F32_BITS = int(np.float32(1.0).view(np.uint32))
VF_LOOP, VF_LOOP_UNTIL_RELEASE = 1, 2
RP_FILTER = 1
POOL_GUARD = 8  # zero frames after every sample, src/sf2.rs:583
F32 = np.float32


# ---------------------------------------------------------------------------
# coefficient design (float32, src/bank.rs:452-511, 876-979)
# ---------------------------------------------------------------------------

def biquad_lowpass(fc: float, q_cb: float, sr: float) -> tuple:
    q_default = F32(0.707106769)
    pi = F32(3.14159274)
    q_db = F32(q_cb) / F32(10.0) - F32(3.01)
    q = max(F32(np.power(F32(10.0), q_db / F32(20.0))), F32(0.001))
    gain = F32(np.sqrt(q_default / q))
    w0 = F32(2.0) * pi * F32(min(max(F32(fc) / F32(sr), F32(1.0e-5)), F32(0.49)))
    sin_w0, cos_w0 = F32(np.sin(w0)), F32(np.cos(w0))
    alpha = sin_w0 / (F32(2.0) * q)
    a0 = F32(1.0) + alpha
    b0 = gain * (F32(1.0) - cos_w0) * F32(0.5) / a0
    b1 = gain * (F32(1.0) - cos_w0) / a0
    a1 = F32(-2.0) * cos_w0 / a0
    a2 = (F32(1.0) - alpha) / a0
    return F32(b0), F32(b1), F32(a1), F32(a2)


def make_params(attack_s, hold_s, decay_s, sustain, release_s, fc_cents, q_cb, sr,
                exp_decay=True, exp_release=True, env_floor=1e-5, filter_enabled=True):
    """One RegionParams, key scaling and controller modifiers neutral."""
    sr = F32(sr)
    attack_frames = max(F32(attack_s) * sr, F32(0.0))
    hold_frames = max(F32(hold_s) * sr, F32(0.0))
    decay_frames = max(F32(decay_s) * sr, F32(1.0))
    release_frames = max(F32(release_s) * sr, F32(1.0))
    attack_rate = F32(1.0) if attack_frames < 1.0 else F32(1.0) / attack_frames
    attack_end = F32(1.0) + hold_frames * attack_rate
    sus = F32(min(max(F32(sustain), F32(0.0)), F32(1.0)))

    def coef(frames, exp):
        if exp:
            return F32(np.power(F32(10.0), F32(-100.0) / (F32(20.0) * frames)))
        return F32(1.0) / frames

    fc = F32(8.176) * F32(np.power(F32(2.0), F32(fc_cents) / F32(1200.0)))
    use_filter = filter_enabled and fc_cents < 13500.0 and fc < sr * F32(0.49) and fc > 20.0
    b = biquad_lowpass(float(fc), q_cb, float(sr)) if use_filter else (1.0, 0.0, 0.0, 0.0)
    p = np.zeros((), PARAMS_DTYPE)
    p["attack_rate"], p["attack_end"] = attack_rate, attack_end
    p["decay_coef"], p["decay_target"] = coef(decay_frames, exp_decay), max(sus, F32(env_floor))
    p["sustain"], p["release_coef"] = sus, coef(release_frames, exp_release)
    p["b0"], p["b1"], p["a1"], p["a2"] = b
    p["flags"] = RP_FILTER if use_filter else 0
    return p


def hz_to_cents(hz: float) -> float:
    # inverse of cents_to_hz, src/bank.rs:465-468
    return 1200.0 * math.log2(hz / 8.176)


def q_to_cb(q: float) -> float:
    # inverse of biquad_lowpass's Q mapping: q = 10^((q_cb/10 - 3.01)/20)
    return 10.0 * (20.0 * math.log10(q) + 3.01)


def fixed_from_f64(v: np.ndarray) -> np.ndarray:
    """src/fixed.rs:84-96, vectorised."""
    scaled = np.asarray(v, np.float64) * 4294967296.0
    out = np.zeros(scaled.shape, np.uint64)
    ok = np.isfinite(scaled) & (scaled > 0.0)
    big = scaled >= 18446744073709551616.0
    out[ok & ~big] = scaled[ok & ~big].astype(np.uint64)
    out[big] = np.uint64(0xFFFFFFFFFFFFFFFF)
    return out


def velocity_gain(vel: np.ndarray) -> np.ndarray:
    """cb_to_gain(velocity_atten_cb(vel, 100)) in f32 (src/bank.rs:461-505)."""
    v = np.maximum(vel, 1).astype(F32) / F32(127.0)
    cb = np.clip(F32(-400.0) * np.log10(v).astype(F32), F32(0.0), F32(960.0)).astype(F32)
    return np.power(F32(10.0), -cb / F32(200.0)).astype(F32)


# ---------------------------------------------------------------------------
# banks
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class Bank:
    pool: np.ndarray          # int16, every sample followed by POOL_GUARD zeros
    params: np.ndarray        # PARAMS_DTYPE, one entry per region (f32 host design)
    smp_start: np.ndarray     # per region (the left channel of a stereo region)
    smp_len: np.ndarray
    loop_start: np.ndarray
    loop_end: np.ndarray
    loop_flags: np.ndarray
    root_key: np.ndarray
    key_lo: np.ndarray        # per region key range (inclusive)
    key_hi: np.ndarray
    instruments: int
    zones: int
    design: np.ndarray = None     # DESIGN_DTYPE per region: the inputs of `params`
    smp_start_r: np.ndarray = None  # right channel of stereo regions, else None

    @property
    def stereo(self) -> bool:
        return self.smp_start_r is not None

    @property
    def n_regions(self) -> int:
        return len(self.smp_start)

    @property
    def max_sample_len(self) -> int:
        return int(self.smp_len.max()) if len(self.smp_len) else 1

    def region_of(self, instrument: np.ndarray, key: np.ndarray) -> np.ndarray:
        span = (self.key_hi[0] - self.key_lo[0] + 1)
        zone = np.clip((key.astype(np.int64) - int(self.key_lo[0])) // int(span), 0,
                       self.zones - 1)
        return (instrument.astype(np.int64) % self.instruments) * self.zones + zone


def sine_noise_sample(rng: np.random.Generator, n: int, freq: float, rate: int,
                      amp: float = 0.5, noise_db: float = -20.0) -> np.ndarray:
    t = np.arange(n, dtype=np.float64)
    phase0 = rng.uniform(0, 2 * np.pi)
    s = amp * np.sin(2 * np.pi * freq * t / rate + phase0)
    s += amp * 10 ** (noise_db / 20.0) * rng.uniform(-1.0, 1.0, n) * math.sqrt(3.0) / math.sqrt(2.0)
    return np.clip(np.round(s * 32767.0), -32767, 32767).astype(np.int16)


def saw_noise_sample(rng: np.random.Generator, n: int, freq: float, rate: int,
                     amp: float = 0.4, noise_db: float = -20.0, harmonics: int = 24) -> np.ndarray:
    """Band-limited sawtooth (additive, partials below Nyquist) plus noise."""
    t = np.arange(n, dtype=np.float64)
    s = np.zeros(n)
    phase0 = rng.uniform(0, 2 * np.pi)
    for k in range(1, harmonics + 1):
        if k * freq >= rate / 2:
            break
        s += np.sin(2 * np.pi * k * freq * t / rate + k * phase0) / k
    s *= amp * 2 / np.pi
    s += amp * 10 ** (noise_db / 20.0) * rng.uniform(-1.0, 1.0, n)
    return np.clip(np.round(s * 32767.0), -32767, 32767).astype(np.int16)


def make_bank(seed: int, instruments: int = 16, zones: int = 4, key_lo: int = 21,
              key_hi: int = 108, frames: int = 32768, loop=(8192, 30000),
              loop_mode: int = VF_LOOP | VF_LOOP_UNTIL_RELEASE, rate: int = 48000,
              adsr=(0.005, 0.100, 0.6, 0.150), cutoff_hz=(2000.0, 12000.0),
              q=0.7071067811865476, wave: str = "sine",
              filter_enabled: bool = True, stereo: bool = False) -> Bank:
    """`instruments` x `zones` regions over [key_lo, key_hi], one int16 sample
    per region (two, left and right, when `stereo`) at its zone's root pitch,
    seeded. `q` is a fixed resonance or a (lo, hi) range drawn log-uniform."""
    from .abi import DESIGN_DTYPE
    rng = np.random.default_rng(seed)
    nkeys = key_hi - key_lo + 1
    span = -(-nkeys // zones)
    chunks, starts, starts_r, lens, roots, klo, khi, params = [], [], [], [], [], [], [], []
    designs = []
    off = 0
    a, d, s, r = adsr
    gen = saw_noise_sample if wave == "saw" else sine_noise_sample
    for inst in range(instruments):
        for z in range(zones):
            lo = key_lo + z * span
            hi = min(lo + span - 1, key_hi)
            root = (lo + hi) // 2
            freq = 440.0 * 2 ** ((root - 69) / 12)
            for ch in range(2 if stereo else 1):
                chunks.append(gen(rng, frames, freq, rate))
                chunks.append(np.zeros(POOL_GUARD, np.int16))
                (starts_r if ch else starts).append(off)
                off += frames + POOL_GUARD
            lens.append(frames)
            roots.append(root)
            klo.append(lo)
            khi.append(hi)
            fc = rng.uniform(*cutoff_hz)
            qq = q if np.isscalar(q) else math.exp(rng.uniform(math.log(q[0]), math.log(q[1])))
            fc_cents, q_cb = hz_to_cents(fc), q_to_cb(qq)
            params.append(make_params(a, 0.0, d, s, r, fc_cents, q_cb, rate,
                                      filter_enabled=filter_enabled))
            designs.append((a, 0.0, d, s, r, fc_cents, q_cb, int(filter_enabled), 0))
    n = len(starts)
    flags = loop_mode if loop[1] > loop[0] + 1 else 0
    return Bank(
        pool=np.concatenate(chunks) if chunks else np.zeros(2, np.int16),
        params=np.array(params, dtype=PARAMS_DTYPE),
        smp_start=np.array(starts, np.uint32), smp_len=np.array(lens, np.uint32),
        loop_start=np.full(n, loop[0], np.uint32), loop_end=np.full(n, loop[1], np.uint32),
        loop_flags=np.full(n, flags, np.uint32), root_key=np.array(roots, np.int64),
        key_lo=np.array(klo, np.int64), key_hi=np.array(khi, np.int64),
        instruments=instruments, zones=zones, design=np.array(designs, dtype=DESIGN_DTYPE),
        smp_start_r=np.array(starts_r, np.uint32) if stereo else None)


# ---------------------------------------------------------------------------
# notes
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class Notes:
    """A note list sorted by onset frame."""
    on: np.ndarray     # onset frame (int64)
    off: np.ndarray    # note-off frame (> on)
    ch: np.ndarray     # channel 0..15 (= instrument)
    key: np.ndarray
    vel: np.ndarray

    def __len__(self):
        return len(self.on)


def _finish(on, dur_frames, ch, key, vel) -> Notes:
    on = np.asarray(on, np.int64)
    dur = np.maximum(np.asarray(dur_frames, np.int64), 1)
    order = np.argsort(on, kind="stable")
    return Notes(on[order], (on + dur)[order], np.asarray(ch, np.int64)[order],
                 np.asarray(key, np.int64)[order], np.asarray(vel, np.int64)[order])


def poisson_notes(seed: int, n: int, seconds: float, rate: int = 48000, key_lo=21, key_hi=108,
                  channels=16, mean_dur_s=0.2) -> Notes:
    """Poisson onsets, uniform keys and velocities, exponential durations."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(seconds / n, n)
    t = np.cumsum(gaps)
    t = t[t < seconds]
    n = len(t)
    return _finish(np.floor(t * rate), np.round(rng.exponential(mean_dur_s, n) * rate),
                   rng.integers(0, channels, n), rng.integers(key_lo, key_hi + 1, n),
                   rng.integers(1, 128, n))


def burst_notes(seed: int, seconds: float, rate: int = 48000, base_rate_hz: float = 300_000.0,
                burst_notes_n: int = 50_000, burst_len_s: float = 0.020,
                burst_every_s: float = 0.5, dur_range_s=(0.5, 8.0), channels=16,
                key_lo=0, key_hi=127) -> Notes:
    """A Poisson background plus dense 20 ms clusters; log-uniform durations."""
    rng = np.random.default_rng(seed)
    n_bg = rng.poisson(base_rate_hz * seconds)
    t = rng.uniform(0.0, seconds, n_bg)
    starts = np.arange(0.0, seconds, burst_every_s)
    tb = (starts[:, None] + rng.uniform(0.0, burst_len_s, (len(starts), burst_notes_n))).ravel()
    t = np.concatenate([t, tb[tb < seconds]])
    n = len(t)
    lo, hi = math.log(dur_range_s[0]), math.log(dur_range_s[1])
    dur = np.exp(rng.uniform(lo, hi, n))
    return _finish(np.floor(t * rate), np.round(dur * rate), rng.integers(0, channels, n),
                   rng.integers(key_lo, key_hi + 1, n), rng.integers(1, 128, n))


# ---------------------------------------------------------------------------
# GateTable / ChannelTable, ported from src/voice.rs:219-658
# ---------------------------------------------------------------------------

class GateTable:
    def __init__(self, cfg: Config):
        self.tiles = cfg.block_frames // cfg.gate_frames
        self.rows = np.zeros(self.tiles * GATE_SLOTS, np.uint32)
        self.on_count = np.zeros(GATE_SLOTS, np.uint32)
        self.off_count = np.zeros(GATE_SLOTS, np.uint32)
        self.pending_off = np.zeros(GATE_SLOTS, np.uint32)
        self.sustain = 0
        self.sostenuto = 0
        self.sost_held = np.zeros(GATE_SLOTS, np.uint32)
        self.cursor = 0
        self.tile_frames = cfg.gate_frames

    @staticmethod
    def slot(ch: int, key: int) -> int:
        return (ch & 15) * 128 + (key & 127)

    def begin_block(self):
        self.cursor = 0

    def _advance_to(self, frame: int):
        tile = frame // self.tile_frames
        while self.cursor <= tile and self.cursor < self.tiles:
            b = self.cursor * GATE_SLOTS
            self.rows[b:b + GATE_SLOTS] = self.off_count
            self.cursor += 1

    def released(self, slot: int, ordinal: int) -> bool:
        return self.off_count[slot % GATE_SLOTS] >= ordinal

    def note_on(self, ch: int, key: int, frame: int) -> int:
        self._advance_to(frame)
        s = self.slot(ch, key)
        self.on_count[s] = (int(self.on_count[s]) + 1) & 0xFFFFFFFF
        return int(self.on_count[s])

    def note_off(self, ch: int, key: int, frame: int):
        self._advance_to(frame)
        s = self.slot(ch, key)
        if self.sostenuto & (1 << (ch & 15)) and self.sost_held[s] > 0:
            if (int(self.off_count[s]) + int(self.pending_off[s])) & 0xFFFFFFFF != self.on_count[s]:
                self.pending_off[s] += 1
                self.sost_held[s] -= 1
            return
        if self.sustained(ch):
            if (int(self.off_count[s]) + int(self.pending_off[s])) & 0xFFFFFFFF != self.on_count[s]:
                self.pending_off[s] += 1
        elif self.off_count[s] != self.on_count[s]:
            self.off_count[s] = (int(self.off_count[s]) + 1) & 0xFFFFFFFF

    def sustained(self, ch: int) -> bool:
        return bool(self.sustain & (1 << (ch & 15)))

    def set_sustain(self, ch: int, down: bool, frame: int):
        if down == self.sustained(ch):
            return
        self._advance_to(frame)
        bit = 1 << (ch & 15)
        if down:
            self.sustain |= bit
            return
        self.sustain &= ~bit
        if self.sostenuto & bit:
            return
        self._flush_pending(ch)

    def set_sostenuto(self, ch: int, down: bool, frame: int):
        bit = 1 << (ch & 15)
        if down == bool(self.sostenuto & bit):
            return
        self._advance_to(frame)
        base = (ch & 15) * 128
        if down:
            self.sostenuto |= bit
            sl = slice(base, base + 128)
            self.sost_held[sl] = (self.on_count[sl] - self.off_count[sl] - self.pending_off[sl])
            return
        self.sostenuto &= ~bit
        self.sost_held[base:base + 128] = 0
        if self.sustained(ch):
            return
        self._flush_pending(ch)

    def _flush_pending(self, ch: int):
        sl = slice((ch & 15) * 128, (ch & 15) * 128 + 128)
        self.off_count[sl] += self.pending_off[sl]
        self.pending_off[sl] = 0

    def all_notes_off(self, ch: int, frame: int):
        self._advance_to(frame)
        sl = slice((ch & 15) * 128, (ch & 15) * 128 + 128)
        if self.sustained(ch):
            self.pending_off[sl] = self.on_count[sl] - self.off_count[sl]
            return
        self.off_count[sl] = self.on_count[sl]

    def all_sound_off(self, ch: int, frame: int):
        self._advance_to(frame)
        sl = slice((ch & 15) * 128, (ch & 15) * 128 + 128)
        self.off_count[sl] = self.on_count[sl]
        self.pending_off[sl] = 0
        self.sost_held[sl] = 0

    def reset_controllers(self, ch: int, frame: int):
        self.set_sostenuto(ch, False, frame)
        self.set_sustain(ch, False, frame)

    def end_block(self):
        while self.cursor < self.tiles:
            b = self.cursor * GATE_SLOTS
            self.rows[b:b + GATE_SLOTS] = self.off_count
            self.cursor += 1

    def row(self, tile: int) -> np.ndarray:
        return self.rows[tile * GATE_SLOTS:(tile + 1) * GATE_SLOTS]

    def sounding(self) -> int:
        return int((self.on_count - self.off_count).astype(np.uint64).sum())


class ChannelTable:
    def __init__(self, cfg: Config):
        self.tiles = cfg.block_frames // cfg.gate_frames
        w = BEND_CHANNELS * CHAN_FIELDS
        self.now = np.zeros(w, np.uint32)
        self.now[CHAN_BEND::CHAN_FIELDS] = BEND_ONE
        self.now[CHAN_GAIN_L::CHAN_FIELDS] = ONE_F32_BITS
        self.now[CHAN_GAIN_R::CHAN_FIELDS] = ONE_F32_BITS
        self.rows = np.tile(self.now, self.tiles)
        self.cursor = 0
        self.tile_frames = cfg.gate_frames
        self.bend_active = self.gain_active = self.variant_active = False

    def begin_block(self):
        self.cursor = 0

    def _advance_to(self, frame: int):
        tile = frame // self.tile_frames
        w = BEND_CHANNELS * CHAN_FIELDS
        while self.cursor <= tile and self.cursor < self.tiles:
            self.rows[self.cursor * w:(self.cursor + 1) * w] = self.now
            self.cursor += 1

    def set_bend(self, ch: int, factor: int, frame: int):
        i = (ch & 15) * CHAN_FIELDS + CHAN_BEND
        if self.now[i] == factor:
            return
        self._advance_to(frame)
        self.now[i] = factor

    def set_gain(self, ch: int, l: float, r: float, frame: int):
        b = (ch & 15) * CHAN_FIELDS
        lb, rb = int(F32(l).view(np.uint32)), int(F32(r).view(np.uint32))
        if self.now[b + CHAN_GAIN_L] == lb and self.now[b + CHAN_GAIN_R] == rb:
            return
        self._advance_to(frame)
        self.now[b + CHAN_GAIN_L] = lb
        self.now[b + CHAN_GAIN_R] = rb

    def set_variant(self, ch: int, variant: int, frame: int):
        i = (ch & 15) * CHAN_FIELDS + CHAN_VARIANT
        if self.now[i] == variant:
            return
        self._advance_to(frame)
        self.now[i] = variant

    def end_block(self):
        w = BEND_CHANNELS * CHAN_FIELDS
        while self.cursor < self.tiles:
            self.rows[self.cursor * w:(self.cursor + 1) * w] = self.now
            self.cursor += 1

    def refresh_active(self):
        e = self.rows.reshape(-1, CHAN_FIELDS)
        self.bend_active = bool((e[:, CHAN_BEND] != BEND_ONE).any())
        self.gain_active = bool(((e[:, CHAN_GAIN_L] != ONE_F32_BITS)
                                 | (e[:, CHAN_GAIN_R] != ONE_F32_BITS)).any())
        self.variant_active = bool((e[:, CHAN_VARIANT] != 0).any())

    @staticmethod
    def channel_of(gate_slot: int) -> int:
        return (gate_slot >> 7) & 15


def bend_factor(semitones: float) -> int:
    """src/fixed.rs:159-171"""
    if not math.isfinite(semitones):
        return BEND_ONE
    f = 2.0 ** (semitones / 12.0) * BEND_ONE
    if f <= 0.0:
        return 1
    if f >= 4294967295.0:
        return 0xFFFFFFFF
    return int(math.floor(f + 0.5))


# ---------------------------------------------------------------------------
# driver-lite: notes -> per-block Backend inputs
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class BlockInput:
    gates: np.ndarray       # tiles * 2048 u32
    chans: np.ndarray       # tiles * 64 u32
    flags: tuple            # (bend, gain, variant)
    cmds: np.ndarray        # SPAWN_DTYPE

    @property
    def chan_flag_word(self) -> int:
        b, g, v = self.flags
        return (1 if b else 0) | (2 if g else 0) | (4 if v else 0)


class BlockBuilder:
    """Vectorised driver-lite: the same rows GateTable/ChannelTable would
    publish for a stream of note-ons and note-offs without pedals, and the
    spawn commands `materialise`/`make_cmd` would build for every note-on."""

    def __init__(self, cfg: Config, bank: Bank, notes: Notes, bends=None):
        self.cfg, self.bank, self.notes = cfg, bank, notes
        self.F = cfg.block_frames
        self.tiles = cfg.tiles()
        self.block = 0
        self.next_id = 0
        self.on_count = np.zeros(GATE_SLOTS, np.uint64)
        self.off_count = np.zeros(GATE_SLOTS, np.uint64)
        self.off_sorted = np.sort(notes.off, kind="stable")
        self.off_order = np.argsort(notes.off, kind="stable")
        self.on_ptr = 0
        self.off_ptr = 0
        # Optional per-channel bend curve: callable (block, tiles) -> [tiles, 16] 8.24 factors.
        self.bends = bends

    def next_block(self) -> BlockInput:
        cfg, F, gf = self.cfg, self.F, self.cfg.gate_frames
        start, end = self.block * F, (self.block + 1) * F
        n = self.notes
        # note-ons in [start, end)
        i0 = self.on_ptr
        i1 = int(np.searchsorted(n.on, end, side="left"))
        self.on_ptr = i1
        j0 = self.off_ptr
        j1 = int(np.searchsorted(self.off_sorted, end, side="left"))
        self.off_ptr = j1
        offs = self.off_order[j0:j1]
        off_rel = (n.off[offs] - start).astype(np.int64)
        off_slot = (n.ch[offs] & 15) * 128 + (n.key[offs] & 127)
        # Gate rows: row t = off counts after every off at frame < t * gate_frames.
        cnt = np.zeros((self.tiles + 1, GATE_SLOTS), np.uint64)
        np.add.at(cnt, (off_rel // gf + 1, off_slot), 1)
        rows = self.off_count[None, :] + np.cumsum(cnt, axis=0)[: self.tiles]
        self.off_count += cnt.sum(axis=0)
        gates = rows.astype(np.uint32).ravel()

        # Spawn commands, in event order.
        idx = np.arange(i0, i1)
        m = len(idx)
        cmds = np.zeros(m, SPAWN_DTYPE)
        if m:
            ch, key, vel = n.ch[idx], n.key[idx], n.vel[idx]
            slot = (ch & 15) * 128 + (key & 127)
            # ordinal = on_count before the block + rank among same-slot ons.
            order = np.argsort(slot, kind="stable")
            s_sorted = slot[order]
            first = np.r_[0, np.flatnonzero(np.diff(s_sorted)) + 1]
            grp = np.repeat(first, np.diff(np.r_[first, m]))
            rank = np.empty(m, np.int64)
            rank[order] = np.arange(m) - grp
            ordinal = self.on_count[slot] + rank.astype(np.uint64) + 1
            np.add.at(self.on_count, slot, 1)
            b = self.bank
            region = b.region_of(ch, key)
            ratio = np.exp2((key - b.root_key[region]) / 12.0)
            step = fixed_from_f64(ratio)
            gain = velocity_gain(vel)
            theta = F32((0.0 + 1.0) * 0.5) * F32(F32(np.pi) * F32(0.5))
            cmds["phase_lo"] = 0
            cmds["phase_hi"] = 0
            cmds["step_lo"] = (step & np.uint64(0xFFFFFFFF)).astype(np.uint32)
            cmds["step_hi"] = (step >> np.uint64(32)).astype(np.uint32)
            cmds["smp_base"] = b.smp_start[region]
            cmds["smp_len"] = b.smp_len[region]
            cmds["loop_start"] = b.loop_start[region]
            cmds["loop_end"] = b.loop_end[region]
            cmds["flags"] = b.loop_flags[region]
            cmds["params"] = region
            cmds["variant"] = 0
            cmds["region"] = region
            cmds["gate_slot"] = slot
            cmds["ordinal"] = ordinal.astype(np.uint32)
            cmds["start_rel"] = (n.on[idx] - start).astype(np.uint32)
            ids = np.arange(self.next_id, self.next_id + m, dtype=np.uint64)
            cmds["note_id_lo"] = (ids & np.uint64(0xFFFFFFFF)).astype(np.uint32)
            cmds["note_id_hi"] = (ids >> np.uint64(32)).astype(np.uint32)
            cmds["gain_l"] = gain * F32(np.cos(theta))
            cmds["gain_r"] = gain * F32(np.sin(theta))
            self.next_id += m
            if b.stereo:
                # A stereo sample is two hard-panned mono voices
                # (src/sfz.rs:10-11, 364-368): left then right, each its own id.
                two = np.repeat(cmds, 2)
                two["smp_base"][1::2] = b.smp_start_r[region]
                two["gain_l"][0::2] = gain
                two["gain_r"][0::2] = 0.0
                two["gain_l"][1::2] = 0.0
                two["gain_r"][1::2] = gain
                ids2 = np.arange(self.next_id - m, self.next_id - m + 2 * m, dtype=np.uint64)
                two["note_id_lo"] = (ids2 & np.uint64(0xFFFFFFFF)).astype(np.uint32)
                two["note_id_hi"] = (ids2 >> np.uint64(32)).astype(np.uint32)
                self.next_id += m
                cmds = two

        w = BEND_CHANNELS * CHAN_FIELDS
        chan = np.zeros((self.tiles, BEND_CHANNELS, CHAN_FIELDS), np.uint32)
        chan[:, :, CHAN_BEND] = BEND_ONE
        chan[:, :, CHAN_GAIN_L] = ONE_F32_BITS
        chan[:, :, CHAN_GAIN_R] = ONE_F32_BITS
        bend_active = False
        if self.bends is not None:
            f = np.asarray(self.bends(self.block, self.tiles), np.uint32)
            chan[:, :, CHAN_BEND] = f
            bend_active = bool((f != BEND_ONE).any())
        self.block += 1
        return BlockInput(gates, chan.reshape(self.tiles * w), (bend_active, False, False), cmds)


def random_walk_bends(seed: int, semitone_range: float = 2.0, step_st: float = 0.02):
    """Per-channel pitch-bend random walks clamped to +-range semitones, one
    value per gate tile, as 8.24 factors (src/fixed.rs:159-171)."""
    rng = np.random.default_rng(seed)
    state = np.zeros(BEND_CHANNELS)

    def curve(block: int, tiles: int) -> np.ndarray:
        out = np.empty((tiles, BEND_CHANNELS), np.uint32)
        for t in range(tiles):
            state[:] = np.clip(state + rng.normal(0.0, step_st, BEND_CHANNELS),
                               -semitone_range, semitone_range)
            out[t] = [bend_factor(x) for x in state]
        return out

    return curve


# ---------------------------------------------------------------------------
# the configurations
# ---------------------------------------------------------------------------

def config_c1(**kw) -> Config:
    """Plain method, small (CPU-runnable): F=256, 4096 voices, linear."""
    base = dict(block_frames=256, max_voices=4096, interpolation=1, max_steal_percent=25)
    base.update(kw)
    return Config(**base)


def config_c2(**kw) -> Config:
    """Plain method at saturation: F=512, 1,048,576 voices, steal 10 %, linear."""
    base = dict(block_frames=512, max_voices=1 << 20, interpolation=1, max_steal_percent=10)
    base.update(kw)
    return Config(**base)


def config_c3(**kw) -> Config:
    """Precision form: F=512, 2,097,152 voices, cubic, compensated mixdown.
    Steal rule Oldest, so f32 and f64 renders of the same stream keep the same
    voices (Quietest ranks on the envelope level, which the precisions round
    differently)."""
    base = dict(block_frames=512, max_voices=1 << 21, interpolation=2, max_steal_percent=10,
                steal_rule=0, compensated=True)
    base.update(kw)
    return Config(**base)


def workload_c1(seed: int = 0x4B530001, cfg: Config | None = None):
    cfg = cfg or config_c1()
    bank = make_bank(seed, instruments=16, zones=4, key_lo=21, key_hi=108)
    notes = poisson_notes(seed + 1, 10_000, 10.0, cfg.sample_rate)
    return cfg, bank, notes


def workload_c2(seed: int = 0x4B530002, seconds: float = 6.0, cfg: Config | None = None,
                **note_kw):
    cfg = cfg or config_c2()
    bank = make_bank(seed, instruments=16, zones=4, key_lo=0, key_hi=127)
    notes = burst_notes(seed + 1, seconds, cfg.sample_rate, **note_kw)
    return cfg, bank, notes


def workload_c3(seed: int = 0x4B530003, seconds: float = 6.0, cfg: Config | None = None,
                **note_kw):
    """Stereo band-limited saw + noise samples of 65536 frames (each note two
    hard-panned voices), resonant lowpass with Q log-uniform in 0.5-20, keys
    0-127, burst note-ons; bends come from `random_walk_bends` (+-2 semitones)."""
    cfg = cfg or config_c3()
    bank = make_bank(seed, instruments=16, zones=4, key_lo=0, key_hi=127, frames=65536,
                     loop=(16384, 60000), wave="saw", q=(0.5, 20.0), stereo=True)
    notes = burst_notes(seed + 1, seconds, cfg.sample_rate, **note_kw)
    return cfg, bank, notes


# ---------------------------------------------------------------------------
# batching form: many independent renders of uneven size
# ---------------------------------------------------------------------------

def c4_specs(seed: int = 0x4B530004, n: int = 64) -> list[dict]:
    """`n` render descriptions: pool 4,096-524,288 voices (log-uniform), note
    count 1e3-5e6 (log-uniform) over 5-60 s, 4-64 instruments, sample length
    4k-128k frames, linear or cubic."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        out.append(dict(
            index=i, seed=seed + 101 * (i + 1),
            max_voices=int(round(math.exp(rng.uniform(math.log(4096), math.log(524288))))),
            notes=int(round(math.exp(rng.uniform(math.log(1e3), math.log(5e6))))),
            seconds=float(rng.uniform(5.0, 60.0)),
            instruments=int(rng.integers(4, 65)),
            frames=int(2 ** int(rng.integers(12, 18))),
            interpolation=int(rng.choice([1, 2]))))
    return out


def c4_render(spec: dict, window_s: float, block_frames: int = 512, waves_per_group: int = 1,
              sort_digit_bits: int = 9):
    """(cfg, bank, notes) for one batch member: its first `window_s` seconds of
    Poisson note-ons at its own rate, exponential durations (mean 1 s)."""
    cfg = Config(block_frames=block_frames, max_voices=spec["max_voices"],
                 interpolation=spec["interpolation"], max_steal_percent=10, compensated=True,
                 waves_per_group=waves_per_group, sort_digit_bits=sort_digit_bits)
    frames = spec["frames"]
    bank = make_bank(spec["seed"], instruments=spec["instruments"], zones=4, key_lo=0,
                     key_hi=127, frames=frames, loop=(frames // 4, frames - frames // 16))
    rate = spec["notes"] / spec["seconds"]
    n = max(1, int(round(rate * window_s)))
    notes = poisson_notes(spec["seed"] + 1, n, window_s, cfg.sample_rate, key_lo=0, key_hi=127,
                          mean_dur_s=1.0)
    return cfg, bank, notes


# ---------------------------------------------------------------------------
# extreme pool
# ---------------------------------------------------------------------------

def config_c5(**kw) -> Config:
    """Extreme pool: F=512, 67,108,864 voices, steal 5 %, cubic, compensated."""
    base = dict(block_frames=512, max_voices=1 << 26, interpolation=2, max_steal_percent=5,
                compensated=True)
    base.update(kw)
    return Config(**base)


def workload_c5(seed: int = 0x4B530005, seconds: float = 0.25, cfg: Config | None = None,
                note_rate_hz: float = 4_000_000.0):
    """128 instruments x 4 zones of stereo 65536-frame samples (each note two
    voices); Poisson note-ons at `note_rate_hz` plus 50 ms bursts every 0.5 s,
    exponential durations (mean 1 s), 128 channels (= instruments)."""
    cfg = cfg or config_c5()
    bank = make_bank(seed, instruments=128, zones=4, key_lo=0, key_hi=127, frames=65536,
                     loop=(16384, 60000), stereo=True)
    rng = np.random.default_rng(seed + 1)
    n_bg = rng.poisson(note_rate_hz * seconds)
    t = rng.uniform(0.0, seconds, n_bg)
    starts = np.arange(0.0, seconds, 0.5)
    tb = (starts[:, None] + rng.uniform(0.0, 0.05, (len(starts), int(note_rate_hz * 0.05)))).ravel()
    t = np.concatenate([t, tb[tb < seconds]])
    n = len(t)
    notes = _finish(np.floor(t * cfg.sample_rate),
                    np.maximum(np.round(rng.exponential(1.0, n) * cfg.sample_rate), 1),
                    rng.integers(0, 128, n), rng.integers(0, 128, n), rng.integers(1, 128, n))
    return cfg, bank, notes


def synthetic_pool(cfg: Config, bank: Bank, n: int, seed: int, chunk: int = 1 << 22) -> np.ndarray:
    """A full pool in the 24-field SoA layout, as if `n` notes had been playing
    for a while: random regions and keys, phases inside the sample, stages and
    levels spread over attack/decay/sustain/release, unique note ids 0..n-1,
    ordinals that no gate row of a later window reaches."""
    from .abi import FIELD
    out = np.zeros((24, n), np.uint32)
    rng = np.random.default_rng(seed)
    params = bank.params
    for lo in range(0, n, chunk):
        hi = min(n, lo + chunk)
        m = hi - lo
        r = rng.integers(0, bank.n_regions, m)
        key = rng.integers(bank.key_lo[r], bank.key_hi[r] + 1)
        step = fixed_from_f64(np.exp2((key - bank.root_key[r]) / 12.0))
        right = (rng.random(m) < 0.5) if bank.stereo else np.zeros(m, bool)
        base = np.where(right, bank.smp_start_r[r] if bank.stereo else 0, bank.smp_start[r])
        idx = rng.integers(0, bank.loop_end[r])
        u = rng.random(m)
        stage = np.select([u < 0.05, u < 0.15, u < 0.85], [0, 1, 2], 3).astype(np.uint32)
        sus = params["sustain"][r].astype(np.float32)
        level = np.select([stage == 0, stage == 1, stage == 2],
                          [rng.random(m), sus + (1 - sus) * rng.random(m), sus],
                          sus * rng.random(m) + 1e-4).astype(np.float32)
        gain = velocity_gain(rng.integers(1, 128, m))
        gl = np.where(right, 0.0, gain).astype(np.float32) if bank.stereo else gain * F32(0.70710677)
        gr = np.where(right, gain, 0.0).astype(np.float32) if bank.stereo else gain * F32(0.70710677)
        ch = rng.integers(0, 16, m)
        ids = np.arange(lo, hi, dtype=np.uint64)
        f = out[:, lo:hi]
        f[FIELD["phase_lo"]] = rng.integers(0, 1 << 32, m, dtype=np.uint64).astype(np.uint32)
        f[FIELD["phase_hi"]] = idx.astype(np.uint32)
        f[FIELD["step_lo"]] = (step & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        f[FIELD["step_hi"]] = (step >> np.uint64(32)).astype(np.uint32)
        f[FIELD["smp_base"]] = base
        f[FIELD["smp_len"]] = bank.smp_len[r]
        f[FIELD["loop_start"]] = bank.loop_start[r]
        f[FIELD["loop_end"]] = bank.loop_end[r]
        f[FIELD["flags"]] = bank.loop_flags[r]
        f[FIELD["env_stage"]] = stage
        f[FIELD["env_level"]] = level.view(np.uint32)
        f[FIELD["gain_l"]] = np.asarray(gl, np.float32).view(np.uint32)
        f[FIELD["gain_r"]] = np.asarray(gr, np.float32).view(np.uint32)
        f[FIELD["params"]] = r
        f[FIELD["region"]] = r
        f[FIELD["gate_slot"]] = (ch * 128 + key).astype(np.uint32)
        f[FIELD["ordinal"]] = 0x7FFFFFFF
        f[FIELD["note_lo"]] = (ids & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        f[FIELD["note_hi"]] = (ids >> np.uint64(32)).astype(np.uint32)
    return out


# ---------------------------------------------------------------------------
# inputs of the native driver (kestrel_amd/driver.py): a bank description in
# the C ABI's region layout and a dense standard MIDI file
# ---------------------------------------------------------------------------

def driver_bank_desc(seed: int, instruments: int = 16, zones: int = 4, frames: int = 32768,
                     adsr=(0.005, 0.100, 0.6, 0.150), stereo: bool = False):
    """(pool, samples, regions, presets, preset_regions): `make_bank`'s samples as
    SampleInfo / Region records (one preset per instrument, program = instrument;
    the last one also as the drum kit in bank 128), filters on half the regions.
    Stereo banks get two hard-panned regions per zone, as src/sf2.rs lays out a
    stereo pair."""
    from .abi import PRESET_DTYPE, REGION_DTYPE, SAMPLE_DTYPE
    b = make_bank(seed, instruments=instruments, zones=zones, key_lo=0, key_hi=127,
                  frames=frames, adsr=adsr, stereo=stereo)
    rng = np.random.default_rng(seed + 7)
    n = len(b.smp_start)
    chans = 2 if stereo else 1
    samples = np.zeros(n * chans, SAMPLE_DTYPE)
    regions = np.zeros(n * chans, REGION_DTYPE)
    a, d, s, r = adsr
    for i in range(n):
        for c in range(chans):
            j = i * chans + c
            samples[j] = ((b.smp_start_r if c else b.smp_start)[i], b.smp_len[i], b.loop_start[i],
                          b.loop_end[i], 48000, 0.0, 1.0, int(b.root_key[i]))
            reg = regions[j:j + 1]
            reg["sample"] = j
            reg["key_lo"], reg["key_hi"] = int(b.key_lo[i]), int(b.key_hi[i])
            reg["vel_lo"], reg["vel_hi"] = 0, 127
            reg["root_key_override"] = reg["fixed_key"] = reg["fixed_vel"] = -1
            reg["scale_tuning"] = 100
            reg["amp_veltrack"] = 100.0
            reg["pan"] = (-1.0 if c == 0 else 1.0) if stereo else float(rng.uniform(-0.5, 0.5))
            reg["loop_mode"] = 2 if b.loop_flags[i] else 0
            reg["attack"], reg["decay"], reg["sustain"], reg["release"] = a, d, s, r
            reg["attenuation_cb"] = float(rng.uniform(0, 60))
            reg["filter_fc_cents"] = float(b.design["fc_cents"][i]) if i % 2 else 13500.0
            reg["filter_q_cb"] = float(b.design["q_cb"][i])
    presets = np.zeros(instruments + 1, PRESET_DTYPE)
    per = zones * chans
    for p in range(instruments):
        presets[p] = (0, p, p * per, per)
    presets[instruments] = (128, 0, (instruments - 1) * per, per)
    return (b.pool, samples, regions, presets, np.arange(n * chans, dtype=np.uint32))


def dense_midi(path: str, seed: int, notes: int, seconds: float, ppq: int = 960,
               channels: int = 16, note_ms=(2.0, 60.0), key_range=(12, 115)) -> int:
    """Write a format-1 SMF of `notes` short notes spread uniformly over
    `seconds` at 120 bpm, one track per channel, each opening with a program
    change, volume, pan and expression; returns the file size. Events are
    written with running status. Vectorised, so tens of millions of notes
    take seconds."""
    rng = np.random.default_rng(seed)
    tps = ppq * 2.0  # ticks per second at 120 bpm
    on = np.sort(rng.uniform(0, seconds * tps, notes)).astype(np.int64)
    dur = np.maximum(1, (rng.uniform(*note_ms, notes) * 1e-3 * tps).astype(np.int64))
    ch = rng.integers(0, channels, notes)
    key = rng.integers(key_range[0], key_range[1], notes)
    vel = rng.integers(30, 128, notes)
    tracks = []
    for c in range(channels):
        m = ch == c
        t_on, t_off, k, v = on[m], on[m] + dur[m], key[m], vel[m]
        tick = np.concatenate([t_on, t_off])
        kind = np.concatenate([np.ones(t_on.size, np.int64), np.zeros(t_off.size, np.int64)])
        kk = np.concatenate([k, k])
        vv = np.concatenate([v, np.zeros_like(v)])  # note-off as note-on velocity 0
        order = np.lexsort((kind, tick))  # offs before ons on a shared tick
        tick, kk, vv = tick[order], kk[order], vv[order]
        delta = np.diff(np.concatenate([[0], tick]))
        head = bytes([0, 0xC0 | c, c % 8, 0, 0xB0 | c, 7, 110, 0, 10, int(rng.integers(20, 108)),
                      0, 11, 120])
        body = _varlen_bytes(delta, kk, vv, 0x90 | c)
        tracks.append(head + body + bytes([0, 0xFF, 0x2F, 0]))
    with open(path, "wb") as f:
        f.write(b"MThd" + (6).to_bytes(4, "big") + (1).to_bytes(2, "big")
                + len(tracks).to_bytes(2, "big") + ppq.to_bytes(2, "big"))
        for t in tracks:
            f.write(b"MTrk" + len(t).to_bytes(4, "big") + t)
        return f.tell()


def _varlen_bytes(delta: np.ndarray, key: np.ndarray, vel: np.ndarray, status: int) -> bytes:
    """Delta-time varlen + running-status note messages, vectorised: the first
    message carries the status byte, the rest reuse it."""
    delta = delta.astype(np.int64)
    nb = np.ones(delta.size, np.int64)
    for k in range(1, 5):
        nb += delta >= (1 << (7 * k))
    total = nb + 2
    total[0] += 1
    out = np.zeros(int(total.sum()), np.uint8)
    pos = np.concatenate([[0], np.cumsum(total)[:-1]])
    for k in range(5):
        sel = nb > k
        shift = 7 * (nb[sel] - 1 - k)
        byte = (delta[sel] >> shift) & 0x7F
        out[pos[sel] + k] = byte | np.where(k < nb[sel] - 1, 0x80, 0)
    p = pos + nb
    out[p[0]] = status
    p[0] += 1
    out[p] = key
    out[p + 1] = vel
    return out.tobytes()
