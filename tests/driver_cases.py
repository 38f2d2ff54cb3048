"""Shared inputs for the driver tests: a synthetic instrument bank that reaches
every branch of src/bank.rs the driver uses, random event streams over every
controller the driver acts on, a minimal SMF writer, and the block-by-block
comparison of the library's driver against the oracle's.
"""
from __future__ import annotations

import numpy as np

from kestrel_amd.abi import (EVENT_DTYPE, PARAMS_DTYPE, PRESET_DTYPE, REGION_DTYPE, SAMPLE_DTYPE,
                             SPAWN_DTYPE)
from kestrel_amd.config import Config

GUARD = 8  # zero frames after every sample (src/sf2.rs:581-583)


def make_bank_desc(seed: int, n_samples: int = 12, n_regions: int = 48, n_presets: int = 10,
                   pool_rate: int = 0):
    """(pool, samples, regions, presets, preset_regions, pool_rate) with edge cases:
    an empty sample, a region naming a missing sample, velocity splits, fixed
    keys and velocities, every loop mode, address offsets under resampling,
    delays long enough to defer, key scaling, veltracked filters, a drum bank
    and more overlapping layers than max_layers."""
    rng = np.random.default_rng(seed)
    pool, samples = [], np.zeros(n_samples, SAMPLE_DTYPE)
    at = 0
    for i in range(n_samples):
        n = 0 if i == 3 else int(rng.integers(64, 3000))
        t = np.arange(n)
        x = 0.6 * np.sin(2 * np.pi * t * rng.uniform(0.005, 0.05)) + 0.2 * rng.standard_normal(n)
        data = np.clip(x * 20000, -32768, 32767).astype(np.int16)
        pool.append(data)
        pool.append(np.zeros(GUARD, np.int16))
        ls = int(rng.integers(0, max(n // 2, 1)))
        le = int(rng.integers(ls, n + 1)) if n else 0
        samples[i] = (at, n, ls, le, int(rng.choice([22050, 44100, 48000])),
                      float(rng.uniform(-40, 40)), float(rng.choice([1.0, 1.0, 0.5, 1.7])),
                      int(rng.integers(36, 84)))
        at += n + GUARD
    regions = np.zeros(n_regions, REGION_DTYPE)
    for i in range(n_regions):
        r = regions[i:i + 1]
        r["sample"] = n_samples + 1 if i == 5 else int(rng.integers(0, n_samples))
        klo = int(rng.integers(0, 100))
        r["key_lo"], r["key_hi"] = klo, min(127, klo + int(rng.integers(0, 60)))
        if rng.random() < 0.3:
            vlo = int(rng.integers(0, 100))
            r["vel_lo"], r["vel_hi"] = vlo, min(127, vlo + int(rng.integers(10, 60)))
        else:
            r["vel_lo"], r["vel_hi"] = 0, 127
        r["root_key_override"] = int(rng.integers(30, 90)) if rng.random() < 0.3 else -1
        r["fixed_key"] = int(rng.integers(30, 90)) if rng.random() < 0.1 else -1
        r["fixed_vel"] = int(rng.integers(1, 128)) if rng.random() < 0.1 else -1
        r["coarse_tune"] = int(rng.integers(-12, 13)) if rng.random() < 0.3 else 0
        r["fine_tune"] = int(rng.integers(-50, 51)) if rng.random() < 0.3 else 0
        r["scale_tuning"] = int(rng.choice([100, 100, 100, 50, 0]))
        per_key = rng.random() < 0.15
        r["keynum_to_hold"] = int(rng.choice([0, 40])) if per_key else 0
        r["keynum_to_decay"] = -60 if per_key and int(r["keynum_to_hold"][0]) == 0 else 0
        r["attenuation_cb"] = float(rng.uniform(0, 150))
        r["amp_veltrack"] = float(rng.choice([100.0, 100.0, 0.0, 50.0, -50.0]))
        r["pan"] = float(rng.choice([0.0, -1.0, 1.0, rng.uniform(-1, 1)]))
        r["loop_mode"] = int(rng.integers(0, 3))
        if rng.random() < 0.3:
            r["addr_start"] = int(rng.integers(0, 40))
            r["addr_end"] = -int(rng.integers(0, 40))
            r["addr_loop_start"] = int(rng.integers(-10, 10))
            r["addr_loop_end"] = int(rng.integers(-10, 10))
        r["exclusive_class"] = int(rng.integers(0, 3))
        r["delay"] = float(rng.choice([0.0, 0.0, 0.0, 0.002, 0.03]))
        r["attack"] = float(rng.choice([0.0, 0.001, 0.02, 0.3]))
        r["hold"] = float(rng.choice([0.0, 0.05]))
        r["decay"] = float(rng.choice([0.001, 0.2, 2.0]))
        r["sustain"] = float(rng.uniform(0, 1))
        r["release"] = float(rng.choice([0.001, 0.003, 0.1, 1.0]))
        r["filter_fc_cents"] = float(rng.choice([13500.0, rng.uniform(5000, 12500)]))
        r["filter_q_cb"] = float(rng.choice([0.0, rng.uniform(0, 200)]))
        # a veltracked filter expands the region's params along velocity; not
        # on top of key scaling, which would make a 128 x 128 block
        r["filter_veltrack_cents"] = 0.0 if per_key else float(rng.choice([0.0, 0.0, 2400.0]))
    # Presets: lists of region indices; the last one stacks every region.
    presets = np.zeros(n_presets, PRESET_DTYPE)
    pre_regions = []
    for p in range(n_presets):
        if p == n_presets - 1:
            regs = list(range(n_regions))
        else:
            regs = list(rng.choice(n_regions, size=int(rng.integers(2, 10)), replace=False))
        bank = 128 if p in (2, 7) else int(rng.choice([0, 0, 1]))
        presets[p] = (bank, p if p < n_presets - 1 else 99, len(pre_regions), len(regs))
        pre_regions.extend(int(x) for x in regs)
    return (np.concatenate(pool), samples, regions, presets,
            np.array(pre_regions, np.uint32), pool_rate)


APPLIED_CC = (0, 1, 6, 7, 10, 11, 32, 33, 38, 39, 42, 43, 64, 66, 67, 71, 72, 73, 74, 75, 76, 77,
              92, 96, 97, 98, 99, 100, 101, 120, 121, 123, 124, 125, 126, 127, 2, 5, 91)


def random_events(seed: int, n_notes: int, ticks: int, *, cc_rate: float = 0.4,
                  channels=range(16), long_notes: float = 0.3) -> np.ndarray:
    """A tick-ordered stream: note pairs (some overlapping restrikes), every
    controller the driver acts on (and a few it ignores), programs, bends,
    tempo changes and events it skips."""
    rng = np.random.default_rng(seed)
    ev = []
    chs = list(channels)
    for _ in range(n_notes):
        t = int(rng.integers(0, ticks))
        ch, key, vel = int(rng.choice(chs)), int(rng.integers(20, 110)), int(rng.integers(1, 128))
        dur = int(rng.integers(1, ticks // 4 if rng.random() < long_notes else 40))
        ev.append((t, 0, ch, key, vel, 0))
        ev.append((t + dur, 1, ch, key, 0, 0))
    for _ in range(int(n_notes * cc_rate)):
        t = int(rng.integers(0, ticks))
        ch = int(rng.choice(chs))
        r = rng.random()
        if r < 0.7:
            num = int(rng.choice(APPLIED_CC))
            val = int(rng.integers(0, 128))
            if num in (71, 72, 73, 74, 75):
                val = int(rng.choice([0, 20, 40, 64, 64, 80, 100, 127, int(rng.integers(0, 128))]))
            if num in (6, 100, 101):
                val = int(rng.choice([0, 0, 1, 12, 24]))
            ev.append((t, 2, ch, num, val, 0))
        elif r < 0.8:
            ev.append((t, 3, ch, int(rng.integers(0, 12)), 0, 0))
        elif r < 0.93:
            ev.append((t, 4, ch, 0, 0, int(rng.integers(-8192, 8192))))
        elif r < 0.97:
            ev.append((t, 5, 0, 0, 0, int(rng.choice([300_000, 500_000, 750_000, 1_200_000]))))
        else:
            ev.append((t, 6, 0, 0, 0, 0))
    ev.sort(key=lambda e: e[0])  # stable: ties keep their generation order
    out = np.zeros(len(ev), EVENT_DTYPE)
    for i, (t, k, ch, a, b, v) in enumerate(ev):
        out[i] = (t, k, ch, a, b, 0, v, 0)
    return out


def events_to_oracle(ev: np.ndarray):
    return [(int(e["tick"]), (int(e["kind"]), int(e["ch"]), int(e["a"]), int(e["b"]),
                              int(e["value"]))) for e in ev]


# ---- SMF writing (the format of src/midi.rs:420-490, written independently) ----
def varlen(v: int) -> bytes:
    out = [v & 0x7F]
    v >>= 7
    while v:
        out.append((v & 0x7F) | 0x80)
        v >>= 7
    return bytes(reversed(out))


def event_bytes(e, running: list | None = None) -> bytes:
    """One channel or meta message for an EVENT_DTYPE row. With `running`
    (a one-element list holding the last status), channel messages use
    running status where they can."""
    k, ch, a, b, v = int(e["kind"]), int(e["ch"]), int(e["a"]), int(e["b"]), int(e["value"])
    if k == 5:
        msg = bytes([0xFF, 0x51, 3]) + v.to_bytes(3, "big")
    elif k == 6:
        msg = bytes([0xFF, 0x01, 2]) + b"hi"
    else:
        if k == 0:
            status, data = 0x90 | ch, bytes([a, b])
        elif k == 1:
            status, data = 0x80 | ch, bytes([a, 64])
        elif k == 2:
            status, data = 0xB0 | ch, bytes([a, b])
        elif k == 3:
            status, data = 0xC0 | ch, bytes([a])
        else:
            raw = v + 8192
            status, data = 0xE0 | ch, bytes([raw & 0x7F, raw >> 7])
        if running is not None and running[0] == status:
            return data
        if running is not None:
            running[0] = status
        return bytes([status]) + data
    if running is not None:
        running[0] = 0  # a meta event clears running status
    return msg


def track_bytes(ev: np.ndarray, use_running: bool = True, end: bool = True) -> bytes:
    out, last, running = bytearray(), 0, [0]
    for e in ev:
        t = int(e["tick"])
        out += varlen(t - last) + event_bytes(e, running if use_running else None)
        last = t
    if end:
        out += varlen(0) + bytes([0xFF, 0x2F, 0])
    return bytes(out)


def write_smf(path: str, tracks: list[bytes], ppq: int = 480, fmt: int = 1,
              extra: bytes = b"") -> None:
    with open(path, "wb") as f:
        f.write(b"MThd" + (6).to_bytes(4, "big") + fmt.to_bytes(2, "big")
                + len(tracks).to_bytes(2, "big") + (ppq & 0xFFFF).to_bytes(2, "big"))
        for t in tracks:
            f.write(b"MTrk" + len(t).to_bytes(4, "big") + t)
        f.write(extra)


# ---- the comparison harness ----
def driver_config(**kw) -> Config:
    base = dict(sample_rate=48000, block_frames=256, gate_frames=32, max_voices=300,
                max_steal_percent=10, max_param_variants=4, max_layers=3, limiter=True,
                limiter_mode=2)
    base.update(kw)
    return Config(**base)


def oracle_settings(cfg: Config) -> dict:
    return dict(sample_rate=cfg.sample_rate, decay_curve=cfg.decay_curve,
                release_curve=cfg.release_curve, filter_enabled=cfg.filter_enabled,
                env_floor=np.float32(cfg.env_floor), master_volume=cfg.master_volume)


def u32(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a).view(np.uint32)


def run_pair(cfg: Config, desc, lib_driver, ora_driver, blocks: int, seed: int = 0):
    """Drive both drivers block by block against a simulated backend (a live
    count that fills and drains, a running stolen counter, a noisy rendered
    block with overs) and compare everything they publish. Returns the number
    of blocks that admitted voices."""
    from oracle import driver as OD
    rng = np.random.default_rng(seed)
    live, stolen, busy = 0, 0, 0
    for b in range(blocks):
        up_l = lib_driver.begin_block()
        up_o = ora_driver.begin_block()
        assert [i for i, _ in up_l] == [i for i, _ in up_o], f"block {b}: variant slots"
        for (_, pl), (_, po) in zip(up_l, up_o):
            assert np.array_equal(u32(pl), u32(OD.params_array(po, PARAMS_DTYPE))), \
                f"block {b}: variant table"
        g, c, flags = lib_driver.rows()
        assert np.array_equal(g, ora_driver.gates.rows.reshape(-1)), f"block {b}: gate rows"
        assert np.array_equal(c, ora_driver.chan.rows.reshape(-1)), f"block {b}: channel rows"
        assert flags == ora_driver.flags, f"block {b}: controller flags"
        cmds = lib_driver.admit(live, stolen)
        ocmds = OD.spawn_array(ora_driver.admit(live, stolen), SPAWN_DTYPE)
        assert cmds.size == ocmds.size, f"block {b}: {cmds.size} spawns vs {ocmds.size}"
        assert cmds.tobytes() == ocmds.tobytes(), f"block {b}: spawn commands"
        busy += cmds.size > 0
        taken = cmds.size
        stolen_after = stolen + max(0, live + taken - cfg.max_voices)
        live = max(0, min(cfg.max_voices, live + taken) - int(rng.integers(1, 2 + live // 6)))
        stolen = stolen_after
        out = (rng.standard_normal(cfg.block_frames * 2) * 0.6).astype(np.float32)
        st = dict(active_voices=live, stolen=stolen, dropped=0, peak=float(np.abs(out).max()))
        out_o = out.copy()
        more_l = lib_driver.end_block(out, stolen_after, st)
        more_o = ora_driver.end_block(out_o, stolen_after, st)
        assert np.array_equal(out.view(np.uint32), out_o.view(np.uint32)), f"block {b}: output"
        assert more_l == more_o, f"block {b}: more"
        sl, so = lib_driver.stats(), ora_driver.stats
        for k, v in so.items():
            assert float(sl[k]) == float(v), f"block {b}: stats {k}: {sl[k]} vs {v}"
        if not more_l:
            break
    return busy
