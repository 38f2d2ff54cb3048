"""TEST INFRASTRUCTURE ONLY: a pure-Python restatement of the reference's
post-render step -- the two limiters and the hard clamp of src/limiter.rs and
the comparison of `kestrel null` (src/bin/kestrel.rs:597-646) -- for checking
the library's host code bit for bit on short signals. Python floats are IEEE
doubles with one rounding per operation, as Rust's f64 arithmetic is; every
f32 store goes through numpy.float32.
"""
from __future__ import annotations

import collections
import math

import numpy as np

F32 = np.float32


class Omni:
    """src/limiter.rs:14-118"""

    def __init__(self, sample_rate: int):
        self.ll = self.lr = 1.0
        self.vl = self.vr = 0.0
        self.attack = 100.0
        self.falloff = sample_rate / 3.0
        self.min_thresh = 0.4
        self.strength = 1.0
        self.first = True

    def _follow(self, s, x):
        if s > x:
            return (s * self.falloff + x) / (self.falloff + 1.0)
        return (s * self.attack + x) / (self.attack + 1.0)

    def process(self, buf: np.ndarray) -> None:
        for i in range(0, len(buf) - 1, 2):
            il, ir = float(buf[i]), float(buf[i + 1])
            self.ll = self._follow(self.ll, abs(il))
            self.lr = self._follow(self.lr, abs(ir))
            if self.ll < self.min_thresh:
                self.ll = self.min_thresh
            if self.lr < self.min_thresh:
                self.lr = self.min_thresh
            l = il / (self.ll * self.strength + 2.0 * (1.0 - self.strength)) / 2.0
            r = ir / (self.lr * self.strength + 2.0 * (1.0 - self.strength)) / 2.0
            if not self.first:
                self.vl = self._follow(self.vl, abs(il - l))
                self.vr = self._follow(self.vr, abs(ir - r))
            self.first = False
            buf[i], buf[i + 1] = F32(l), F32(r)


def polyphase():
    """src/limiter.rs:448-471"""
    taps, phases = 8, 4
    n = float(taps * phases)
    out = []
    for p in range(phases):
        row = []
        for t in range(taps):
            x = t - (taps / 2.0 - 1.0) - p / phases
            sinc = 1.0 if abs(x) < 1e-12 else math.sin(math.pi * x) / (math.pi * x)
            k = float(t * phases + p)
            w = 0.42 - 0.5 * math.cos(2.0 * math.pi * k / (n - 1.0)) + \
                0.08 * math.cos(4.0 * math.pi * k / (n - 1.0))
            row.append(sinc * w)
        s = 0.0
        for c in row:
            s += c
        if abs(s) > 1e-12:
            row = [c / s for c in row]
        out.append(row)
    return out


class Brickwall:
    """src/limiter.rs:203-446"""

    def __init__(self, sample_rate, ceiling, lookahead_ms, release_ms, sustain_ms, true_peak):
        sr = float(sample_rate)
        self.ceiling = ceiling
        self.look = max(int(round_half_away(lookahead_ms * 1e-3 * sr)), 1)
        rel = max(release_ms * 1e-3 * sr, 1.0)
        if sustain_ms > 0.0:
            a = max(sustain_ms * 1e-3 * sr, 1.0)
            self.s_atk, self.s_rel = 1.0 - math.exp(-1.0 / a), 1.0 - math.exp(-1.0 / (a * 4.0))
        else:
            self.s_atk = self.s_rel = 0.0
        self.rel = 1.0 - math.exp(-1.0 / rel)
        self.tp = bool(true_peak)
        self.delay_frames = self.look + (3 if self.tp else 0)
        self.delay = [0.0] * (2 * self.delay_frames)
        self.dpos = 0
        self.dq = collections.deque()
        self.ring = [1.0] * (self.look + 1)
        self.pos = 0
        self.sum = float(self.look + 1)
        self.g_slow = self.g_fast = 1.0
        self.hist = [[0.0] * 8, [0.0] * 8]
        self.poly = polyphase()
        self.idx = 0

    def detect(self, l, r):
        if not self.tp:
            return max(abs(l), abs(r))
        peak = 0.0
        for ch, v in enumerate((l, r)):
            h = self.hist[ch]
            h[1:] = h[:-1]
            h[0] = v
            for ph in self.poly:
                acc = 0.0
                for c, x in zip(ph, h):
                    acc += c * x
                if abs(acc) > peak:
                    peak = abs(acc)
        return peak

    def process(self, buf: np.ndarray) -> None:
        for i in range(0, len(buf) - 1, 2):
            il, ir = float(buf[i]), float(buf[i + 1])
            d = self.dpos * 2
            ol, orr = self.delay[d], self.delay[d + 1]
            self.delay[d], self.delay[d + 1] = il, ir
            self.dpos = (self.dpos + 1) % self.delay_frames
            det = self.detect(il, ir)
            while self.dq and self.dq[-1][0] <= det:
                self.dq.pop()
            self.dq.append((det, self.idx))
            oldest = max(self.idx - self.look, 0)
            while self.dq and self.dq[0][1] < oldest:
                self.dq.popleft()
            env = self.dq[0][0] if self.dq else 0.0
            gr = self.ceiling / env if env > self.ceiling else 1.0
            self.sum += gr - self.ring[self.pos]
            self.ring[self.pos] = gr
            self.pos = (self.pos + 1) % len(self.ring)
            ma = self.sum / float(len(self.ring))
            if self.s_atk > 0.0:
                c = self.s_atk if ma < self.g_slow else self.s_rel
                self.g_slow += (ma - self.g_slow) * c
                self.g_slow = min(max(self.g_slow, 1e-9), 1.0)
            req = min(ma / self.g_slow, 1.0)
            if req < self.g_fast:
                self.g_fast = req
            else:
                self.g_fast += (req - self.g_fast) * self.rel
                if self.g_fast > req:
                    self.g_fast = req
            gain = self.g_slow * self.g_fast
            if gain > ma:
                gain = ma
            buf[i], buf[i + 1] = F32(ol * gain), F32(orr * gain)
            self.idx += 1


def round_half_away(x: float) -> float:
    """f64::round: halves away from zero."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def clamp_block(buf: np.ndarray) -> int:
    """src/limiter.rs:120-129"""
    n = 0
    for i in range(len(buf)):
        v = buf[i]
        if v > 1.0 or v < -1.0:
            n += 1
            buf[i] = F32(1.0) if v > 1.0 else F32(-1.0)
    return n


def null(a: np.ndarray, b: np.ndarray) -> dict:
    """src/bin/kestrel.rs:597-646"""
    n = min(len(a), len(b))
    peak = ssq = pref = 0.0
    worst = 0
    for i in range(n):
        d = float(abs(F32(a[i]) - F32(b[i])))
        if d > peak:
            peak, worst = d, i
        ssq += d * d
        pref = max(pref, float(abs(a[i])))
    return {"compared": n, "peak_ref": pref, "peak_diff": peak,
            "rms_diff": math.sqrt(ssq / max(n, 1)), "worst": worst,
            "identical": bool(np.array_equal(a[:n], b[:n]))}
