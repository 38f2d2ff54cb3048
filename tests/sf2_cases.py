"""SoundFont 2 files for the loader tests, written from the SoundFont 2.01
format (RIFF `sfbk`: INFO, sdta/smpl, pdta with phdr/pbag/pmod/pgen,
inst/ibag/imod/igen and shdr, each hydra list closed by its terminal record),
plus a generator of random soundfonts that reach every branch of the loader.
"""
from __future__ import annotations

import struct

import numpy as np

GAP = 46  # zero frames the format asks for after each sample in smpl


def _name(s: str) -> bytes:
    return s.encode()[:19].ljust(20, b"\0")


def _chunk(cid: bytes, payload: bytes) -> bytes:
    out = cid + struct.pack("<I", len(payload)) + payload
    return out + (b"\0" if len(payload) & 1 else b"")


def write_sf2(path: str, samples: list, instruments: list, presets: list,
              drop: tuple = (), form: bytes = b"sfbk") -> None:
    """samples: dicts(name, data int16, rate, root, correction, loop=(start, end)
    relative, type=1, and optionally start/end overriding the written range);
    instruments: dicts(name, zones=[dict(gens=[(op, amount)], sample=index|None)]);
    presets: dicts(name, bank, program, zones=[dict(gens=[...], instrument=index|None)]).
    A zone with sample/instrument None has no terminal generator (a global zone
    when first). `drop` names pdta chunks to leave out."""
    smpl, ranges = [], []
    at = 0
    for s in samples:
        d = np.asarray(s["data"], np.int16)
        ranges.append((at, at + d.size))
        smpl.append(d)
        smpl.append(np.zeros(GAP, np.int16))
        at += d.size + GAP
    smpl_b = np.concatenate(smpl).astype("<i2").tobytes() if smpl else b""
    phdr, pbag, pgen = bytearray(), bytearray(), bytearray()
    for p in presets:
        phdr += _name(p["name"]) + struct.pack("<HHHIII", p["program"], p["bank"], len(pbag) // 4,
                                               0, 0, 0)
        for z in p["zones"]:
            pbag += struct.pack("<HH", len(pgen) // 4, 0)
            for op, amt in z["gens"]:
                pgen += struct.pack("<Hh", op, amt)
            if z.get("instrument") is not None:
                pgen += struct.pack("<HH", 41, z["instrument"])
    phdr += _name("EOP") + struct.pack("<HHHIII", 0, 0, len(pbag) // 4, 0, 0, 0)
    pbag += struct.pack("<HH", len(pgen) // 4, 0)
    pgen += struct.pack("<HH", 0, 0)
    inst, ibag, igen = bytearray(), bytearray(), bytearray()
    for i in instruments:
        inst += _name(i["name"]) + struct.pack("<H", len(ibag) // 4)
        for z in i["zones"]:
            ibag += struct.pack("<HH", len(igen) // 4, 0)
            for op, amt in z["gens"]:
                igen += struct.pack("<Hh", op, amt)
            if z.get("sample") is not None:
                igen += struct.pack("<HH", 53, z["sample"])
    inst += _name("EOI") + struct.pack("<H", len(ibag) // 4)
    ibag += struct.pack("<HH", len(igen) // 4, 0)
    igen += struct.pack("<HH", 0, 0)
    shdr = bytearray()
    for s, (a, b) in zip(samples, ranges):
        a, b = s.get("start", a), s.get("end", b)
        ls, le = s.get("loop", (0, 0))
        shdr += _name(s["name"]) + struct.pack("<5IBbHH", a, b, a + ls, a + le, s["rate"],
                                               s["root"], s.get("correction", 0), 0,
                                               s.get("type", 1))
    shdr += _name("EOS") + bytes(26)
    pdta_parts = [(b"phdr", phdr), (b"pbag", pbag), (b"pmod", bytes(10)), (b"pgen", pgen),
                  (b"inst", inst), (b"ibag", ibag), (b"imod", bytes(10)), (b"igen", igen),
                  (b"shdr", shdr)]
    pdta = b"pdta" + b"".join(_chunk(c, bytes(p)) for c, p in pdta_parts if c not in drop)
    info = b"INFO" + _chunk(b"ifil", bytes([2, 0, 1, 0])) + _chunk(b"isng", b"EMU8000\0") \
        + _chunk(b"INAM", b"kst test\0\0")
    sdta = b"sdta" + (_chunk(b"smpl", smpl_b) if b"smpl" not in drop else b"")
    body = form + _chunk(b"LIST", info) + _chunk(b"LIST", sdta) + _chunk(b"LIST", pdta)
    with open(path, "wb") as f:
        f.write(b"RIFF" + struct.pack("<I", len(body)) + body)


def sine(freq: float, rate: int, n: int, amp: float = 0.5) -> np.ndarray:
    t = np.arange(n) / rate
    return np.round(np.sin(2 * np.pi * freq * t) * amp * 32767).astype(np.int16)


# generator ranges the random zones draw from: (op, lo, hi)
INST_GENS = [(0, 0, 60), (1, -60, 0), (2, -20, 20), (3, -20, 20), (4, 0, 1), (8, 1500, 13500),
             (9, 0, 300), (17, -500, 500), (33, -12000, -2000), (34, -12000, 1000),
             (35, -12000, -1000), (36, -8000, 2000), (37, 0, 1440), (38, -12000, 1500),
             (39, -100, 100), (40, -100, 100), (48, 0, 300), (51, -24, 24), (52, -99, 99),
             (54, 0, 3), (56, 0, 100), (57, -5, 300), (58, 30, 90), (46, 30, 90), (47, 1, 127)]
PRESET_GENS = [(8, -2000, 2000), (17, -300, 300), (34, -1200, 1200), (38, -1200, 1200),
               (48, 0, 200), (51, -12, 12), (52, -50, 50), (0, 0, 50), (54, 0, 3), (58, 30, 90),
               (43, 0, 0), (44, 0, 0)]


def _range(rng) -> int:
    lo = int(rng.integers(0, 100))
    hi = min(127, lo + int(rng.integers(0, 60)))
    return (hi << 8) | lo if rng.random() < 0.9 else (lo << 8) | hi  # sometimes reversed


def random_sf2(path: str, seed: int) -> None:
    """Samples at mixed rates (one ROM, one past the end of smpl, one unused);
    instruments and presets with global zones, late non-terminal zones, ranges
    (some reversed, some disjoint), repeated generators, absolute-only
    generators at preset level, and a drum kit in bank 128."""
    rng = np.random.default_rng(seed)
    samples = []
    for i in range(9):
        rate = int(rng.choice([22050, 32000, 44100, 48000]))
        n = int(rng.integers(200, 3000))
        d = sine(float(rng.uniform(100, 3000)), rate, n) + (rng.standard_normal(n) * 500).astype(np.int16)
        ls = int(rng.integers(0, n // 2))
        s = dict(name=f"s{i}", data=d, rate=rate, root=int(rng.integers(30, 90)),
                 correction=int(rng.integers(-50, 50)), loop=(ls, int(rng.integers(ls, n + 1))))
        samples.append(s)
    samples[6]["type"] = 0x8001  # ROM
    samples[7]["end"] = 10 ** 8  # past the end of smpl
    instruments = []
    for i in range(7):
        zones = []
        if rng.random() < 0.6:
            zones.append(dict(gens=[(int(op), int(rng.integers(lo, hi + 1)))
                                    for op, lo, hi in rng.permutation(INST_GENS)[:4]], sample=None))
        for _ in range(int(rng.integers(1, 5))):
            gens = [(43, _range(rng)), (44, _range(rng))] if rng.random() < 0.8 else []
            for op, lo, hi in rng.permutation(INST_GENS)[:int(rng.integers(2, 9))]:
                gens.append((int(op), int(rng.integers(lo, hi + 1))))
            if rng.random() < 0.2 and gens:
                gens.append(gens[0])  # a repeated generator: the later one wins
            zones.append(dict(gens=gens, sample=int(rng.integers(0, 9 if i else 8))))
        if rng.random() < 0.3:
            zones.append(dict(gens=[(48, 100)], sample=None))  # not first: ignored
        instruments.append(dict(name=f"i{i}", zones=zones))
    presets = []
    for p in range(8):
        zones = []
        if rng.random() < 0.5:
            zones.append(dict(gens=[(int(op), int(rng.integers(lo, hi + 1)) if op not in (43, 44)
                                     else _range(rng)) for op, lo, hi in
                                    rng.permutation(PRESET_GENS)[:3]], instrument=None))
        for _ in range(int(rng.integers(1, 4))):
            gens = []
            for op, lo, hi in rng.permutation(PRESET_GENS)[:int(rng.integers(0, 5))]:
                gens.append((int(op), _range(rng) if op in (43, 44) else int(rng.integers(lo, hi + 1))))
            zones.append(dict(gens=gens, instrument=int(rng.integers(0, 8))))  # 7: missing
        presets.append(dict(name=f"p{p}", bank=128 if p == 5 else int(rng.choice([0, 0, 1])),
                            program=p, zones=zones))
    write_sf2(path, samples, instruments, presets)


def simple_sf2(path: str, rate: int = 48000) -> np.ndarray:
    """One preset of one looped sine zone, instant envelope, no filter: a note
    renders exactly the sample, repeated. Returns the sample."""
    data = sine(480.0, rate, 1000)  # 10 whole cycles at 48 kHz
    s = dict(name="sine", data=data, rate=rate, root=69, loop=(0, data.size))
    inst = dict(name="i", zones=[dict(gens=[(54, 1)], sample=0)])
    pre = dict(name="p", bank=0, program=0, zones=[dict(gens=[], instrument=0)])
    write_sf2(path, [s], [inst], [pre])
    return data
