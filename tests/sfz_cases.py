"""SFZ instruments and WAV samples for the loader tests: a WAV writer for every
sample format the reader takes (PCM 8/16/24/32, float 32/64, the extensible
header) with an optional `smpl` chunk, and a generator of random SFZ files
over the opcodes, headers and text forms the loader handles.
"""
from __future__ import annotations

import os
import struct

import numpy as np

FORMATS = ("pcm8", "pcm16", "pcm24", "pcm32", "f32", "f64", "ext16")


def write_wav(path: str, x: np.ndarray, rate: int, fmt: str = "pcm16", smpl=None,
              junk: bool = False) -> None:
    """x: float frames in [-1, 1], shape (n,) or (n, channels). smpl: (root
    note, pitch fraction, [(loop start, loop end), ...])."""
    x = np.asarray(x, np.float64)
    if x.ndim == 1:
        x = x[:, None]
    ch = x.shape[1]
    flat = x.reshape(-1)
    tag, bits = {"pcm8": (1, 8), "pcm16": (1, 16), "pcm24": (1, 24), "pcm32": (1, 32),
                 "f32": (3, 32), "f64": (3, 64), "ext16": (0xFFFE, 16)}[fmt]
    if fmt == "pcm8":
        data = np.clip(np.round(flat * 127 + 128), 0, 255).astype(np.uint8).tobytes()
    elif fmt in ("pcm16", "ext16"):
        data = np.clip(np.round(flat * 32767), -32768, 32767).astype("<i2").tobytes()
    elif fmt == "pcm24":
        v = np.clip(np.round(flat * 8388607), -8388608, 8388607).astype(np.int32)
        b = v.astype("<i4").view(np.uint8).reshape(-1, 4)[:, :3]
        data = b.tobytes()
    elif fmt == "pcm32":
        data = np.clip(np.round(flat * 2147483647), -2147483648, 2147483647).astype("<i4").tobytes()
    elif fmt == "f32":
        data = flat.astype("<f4").tobytes()
    else:
        data = flat.astype("<f8").tobytes()
    block = ch * bits // 8
    fmt_b = struct.pack("<HHIIHH", tag, ch, rate, rate * block, block, bits)
    if tag == 0xFFFE:
        guid = struct.pack("<H", 1) + bytes.fromhex("000000001000800000aa00389b71")
        fmt_b += struct.pack("<HHI", 22, bits, 3 if ch == 2 else 4) + guid
    chunks = [(b"fmt ", fmt_b)]
    if junk:
        chunks.append((b"LIST", b"INFOISFT\x05\0\0\0test\0"))
    if smpl is not None:
        # The reference reads the root key and pitch fraction at offsets 20
        # and 24 of the chunk (src/wav.rs:274-276), which the WAV layout
        # calls the SMPTE format and offset; the unity note and fraction
        # fields at 12 and 16 carry other values here so a reader that
        # follows the layout instead is caught.
        root, frac, loops = smpl
        s = struct.pack("<9I", 0, 0, 0, (root + 7) % 128, frac ^ 0x5555, root, frac, len(loops), 0)
        for a, e in loops:
            s += struct.pack("<6I", 0, 0, a, e, 0, 0)
        chunks.append((b"smpl", s))
    chunks.append((b"data", data))
    body = b"WAVE" + b"".join(c + struct.pack("<I", len(p)) + p + (b"\0" if len(p) & 1 else b"")
                              for c, p in chunks)
    with open(path, "wb") as f:
        f.write(b"RIFF" + struct.pack("<I", len(body)) + body)


NOTE = ["c", "c#", "d", "eb", "e", "f", "f#", "g", "ab", "a", "bb", "b"]


def key_text(rng, k: int) -> str:
    if rng.random() < 0.4:
        return f"{NOTE[k % 12]}{k // 12 - 1}"
    return str(k)


def random_sfz(dirpath: str, seed: int) -> str:
    """An instrument directory: WAVs in every format (mono and stereo, with and
    without smpl loops, mixed rates), an SFZ with <control> default_path,
    <global>/<master>/<group>/<region> inheritance, an #include, note names,
    aliases (key=, loopmode=, pitch=), comments, several headers on a line,
    file names with spaces and in the wrong case, disabled and missing samples."""
    rng = np.random.default_rng(seed)
    os.makedirs(os.path.join(dirpath, "Samples"), exist_ok=True)
    names = []
    for i in range(7):
        n = int(rng.integers(300, 2500))
        rate = int(rng.choice([22050, 44100, 48000, 96000]))
        t = np.arange(n) / rate
        x = 0.7 * np.sin(2 * np.pi * rng.uniform(100, 2000) * t) + rng.normal(0, 0.05, n)
        if i % 3 == 0:
            x = np.stack([x, 0.5 * np.roll(x, 7)], axis=1)
        smpl = None
        if rng.random() < 0.6:
            a = int(rng.integers(0, n // 2))
            smpl = (int(rng.integers(30, 90)), int(rng.integers(0, 2 ** 32)),
                    [(a, int(rng.integers(a, n)))])
        name = f"Tone {i}.wav" if i == 2 else f"tone{i}.wav"
        write_wav(os.path.join(dirpath, "Samples", name), np.clip(x, -1, 1), rate,
                  FORMATS[i % len(FORMATS)], smpl, junk=i == 4)
        names.append(name)
    names[5] = names[5].upper()  # a Windows-authored reference to an existing file

    def region(i):
        ops = [f"sample={names[i % len(names)]}"]
        if rng.random() < 0.5:
            lo = int(rng.integers(0, 90))
            ops += [f"lokey={key_text(rng, lo)}", f"hikey={key_text(rng, min(127, lo + int(rng.integers(0, 30))))}"]
        elif rng.random() < 0.5:
            ops.append(f"key={key_text(rng, int(rng.integers(20, 100)))}")
        if rng.random() < 0.4:
            lo = int(rng.integers(0, 100))
            ops += [f"lovel={lo}", f"hivel={min(127, lo + int(rng.integers(5, 60)))}"]
        for op, gen in (("pitch_keycenter", lambda: key_text(rng, int(rng.integers(30, 90)))),
                        ("tune", lambda: f"{rng.uniform(-50, 50):.2f}"),
                        ("pitch", lambda: str(int(rng.integers(-30, 30)))),
                        ("transpose", lambda: str(int(rng.integers(-12, 13)))),
                        ("volume", lambda: f"{rng.uniform(-20, 6):.1f}"),
                        ("pan", lambda: f"{rng.uniform(-120, 120):.1f}"),
                        ("amp_veltrack", lambda: str(int(rng.choice([0, 50, 100, -100])))),
                        ("pitch_keytrack", lambda: str(int(rng.choice([0, 50, 100, 1200])))),
                        ("loop_mode", lambda: str(rng.choice(["loop_continuous", "loop_sustain",
                                                              "one_shot", "no_loop"]))),
                        ("loopmode", lambda: "loop_continuous"),
                        ("offset", lambda: str(int(rng.integers(-5, 200)))),
                        ("end", lambda: str(int(rng.integers(100, 3000)))),
                        ("loop_start", lambda: str(int(rng.integers(0, 300)))),
                        ("loop_end", lambda: str(int(rng.integers(300, 2000)))),
                        ("ampeg_delay", lambda: f"{rng.uniform(0, 0.05):.3f}"),
                        ("ampeg_attack", lambda: f"{rng.uniform(0, 0.5):.3f}"),
                        ("ampeg_hold", lambda: f"{rng.uniform(0, 0.2):.3f}"),
                        ("ampeg_decay", lambda: f"{rng.uniform(0, 3):.3f}"),
                        ("ampeg_sustain", lambda: f"{rng.uniform(0, 120):.1f}"),
                        ("ampeg_release", lambda: f"{rng.uniform(0, 2):.3f}"),
                        ("cutoff", lambda: f"{rng.uniform(30, 15000):.1f}"),
                        ("resonance", lambda: f"{rng.uniform(0, 20):.2f}"),
                        ("fil_veltrack", lambda: str(int(rng.choice([0, 2400, 9600, 12000])))),
                        ("fil_type", lambda: str(rng.choice(["lpf_2p", "hpf_2p", "lpf_1p"]))),
                        ("group", lambda: str(int(rng.integers(0, 4)))),
                        ("off_by", lambda: str(int(rng.integers(0, 4)))),
                        ("amplfo_freq", lambda: "5")):
            if rng.random() < 0.18:
                ops.append(f"{op}={gen()}")
        return ops

    inc = ["// included regions", "<group> volume=-3 ampeg_release=0.4"]
    for i in range(4):
        inc.append("<region> " + " ".join(region(i + 3)))
    with open(os.path.join(dirpath, "extra.sfzh"), "w") as f:
        f.write("\n".join(inc) + "\n")
    lines = ["// a random instrument", "<control> default_path=Samples\\", "#define $X 1",
             "<global> ampeg_attack=0.005 amp_veltrack=80 // trailing comment",
             "<master> tune=10", "<group> lovel=0 hivel=127 fil_type=lpf_2p"]
    for i in range(10):
        if rng.random() < 0.2:
            lines.append(f"<group> lokey=0 hikey=127 pitch_keycenter={int(rng.integers(40, 80))}")
        if rng.random() < 0.15:
            lines.append(f"<master> volume={rng.uniform(-6, 0):.1f}")
        ops = region(i)
        if i == 7:
            ops.append("end=-1")  # disabled
        if rng.random() < 0.3:
            lines.append("<region> " + ops[0])
            lines.append("   " + " ".join(ops[1:]))
        else:
            lines.append("<region> " + " ".join(ops))
    lines.append("<region> sample=missing.wav lokey=60")  # unreadable: dropped
    lines.append('#include "extra.sfzh"')
    lines.append("<global> volume=-1 <region> sample=tone1.wav key=c5 <curve> curve_index=1")
    lines.append("<region> sample=Tone 2.wav lokey=10 hikey=20 offset=5")
    path = os.path.join(dirpath, "inst.sfz")
    with open(path, "w") as f:
        f.write("\r\n".join(lines) + "\r\n")
    return path
