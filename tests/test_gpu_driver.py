"""The whole chain on the device: a MIDI file through the library's driver into
the HIP backend and its post-render step, against the oracle's driver feeding
the oracle's CPU backend (src/driver.rs:434-642 over src/cpu.rs).

Bar: the pool bit-exact after every block (which pins admission, ids, ordinals,
variants, steals and compaction end to end), the driver counters equal, and the
limited, clamped output null at -100 dB or better against the oracle's.
"""
import math

import numpy as np
import pytest

import driver_cases as DC

pytestmark = pytest.mark.gpu

NULL_DB = -100.0


@pytest.mark.parametrize("admit_rule", [0, 1])
def test_midi_file_through_driver_and_device(tmp_path, admit_rule):
    from kestrel_amd.driver import Bank, Driver
    from kestrel_amd.backend import HipSynth
    from oracle import driver as OD
    from oracle import oracle

    cfg = DC.driver_config(block_frames=256, gate_frames=32, max_voices=400,
                           max_steal_percent=10, max_layers=4, max_param_variants=6,
                           admit_rule=admit_rule, compensated=True, sort_voices=True)
    desc = DC.make_bank_desc(90 + admit_rule, n_samples=16, n_regions=64, n_presets=12)
    ev = DC.random_events(91 + admit_rule, 6000, 600, cc_rate=0.25)
    path = str(tmp_path / "burst.mid")
    DC.write_smf(path, [DC.track_bytes(ev[0::2]), DC.track_bytes(ev[1::2])], ppq=480)
    pool, samples, regions, presets, pre, rate = desc
    lb = Bank(cfg, pool, samples, regions, presets, pre, pool_rate=rate)
    ob = OD.Bank(DC.oracle_settings(cfg), samples, regions, presets, pre, pool_rate=rate)
    gpu = HipSynth(cfg, lb.pool, lb.params, lb.n_regions, lb.max_sample_len)
    cpu = oracle.Oracle(cfg, lb.pool, lb.params)
    cpu.set_sorted(lb.n_regions, lb.max_sample_len, cfg.sort_voices)
    ld = Driver.open(cfg, lb, path)
    div, oev = OD.read_midi(path)
    od = OD.Driver(cfg, ob, div, oev)

    worst, peak_live, blocks = -math.inf, 0, 0
    out = np.zeros(cfg.block_frames * 2, np.float32)
    more = True
    while more and blocks < 360:
        print(f"admit_rule {admit_rule} block {blocks}", flush=True)
        more = ld.next_block(gpu, out)
        for idx, table in od.begin_block():
            cpu.set_params_variant(idx, OD.params_array(table, cpu_params_dtype()))
        cpu.set_gates(od.gates.rows.reshape(-1))
        cpu.set_channels(od.chan.rows.reshape(-1), *od.flags)
        st = cpu.stats()
        cmds = OD.spawn_array(od.admit(st["active_voices"], st["stolen"]), spawn_dtype())
        cpu.spawn(cmds)
        stolen_after = cpu.stats()["stolen"]
        oc = cpu.render()
        more_o = od.end_block(oc, stolen_after, cpu.stats())
        assert more == more_o, blocks
        pg, pc = gpu.pool_export(), cpu.pool_export()
        assert pg.shape == pc.shape and np.array_equal(pg, pc), f"block {blocks}: pool"
        peak = float(np.abs(oc).max())
        d = float(np.abs(out.astype(np.float64) - oc).max())
        if peak > 0:
            worst = max(worst, -math.inf if d == 0 else 20 * math.log10(d / peak))
        peak_live = max(peak_live, gpu.stats()["active_voices"])
        blocks += 1
    sl = ld.stats()
    for k, v in od.stats.items():
        if k != "peak":
            assert sl[k] == v, k
    print(f"{blocks} blocks, {sl['notes']} notes, {sl['dropped']} layers dropped, "
          f"peak live {peak_live}, worst null {worst:.1f} dB")
    assert peak_live > 300 and sl["dropped"] > 0
    assert worst <= NULL_DB
    gpu.close()


def cpu_params_dtype():
    from kestrel_amd.abi import PARAMS_DTYPE
    return PARAMS_DTYPE


def spawn_dtype():
    from kestrel_amd.abi import SPAWN_DTYPE
    return SPAWN_DTYPE


def test_soundfont_through_driver_and_device(tmp_path):
    """An SF2 file loaded by the library (pool resampled to the output rate)
    and a MIDI file through its driver and the device, against the same
    bank contents through the oracle's driver and CPU backend."""
    import sf2_cases as S
    from kestrel_amd.driver import Bank, Driver
    from kestrel_amd.backend import HipSynth
    from oracle import driver as OD
    from oracle import oracle

    cfg = DC.driver_config(block_frames=256, gate_frames=32, max_voices=300, max_layers=4,
                           compensated=True)
    sf2 = str(tmp_path / "bank.sf2")
    S.random_sf2(sf2, 3)
    lb = Bank.load_sf2(cfg, sf2)
    c = lb.contents()
    ob = OD.Bank(DC.oracle_settings(cfg), c["samples"], c["regions"], c["presets"],
                 c["preset_regions"], pool_rate=c["pool_rate"])
    ev = DC.random_events(5, 2500, 500, cc_rate=0.2, channels=range(8))
    mid = str(tmp_path / "song.mid")
    DC.write_smf(mid, [DC.track_bytes(ev)], ppq=480)
    gpu = HipSynth(cfg, lb.pool, lb.params, lb.n_regions, lb.max_sample_len)
    cpu = oracle.Oracle(cfg, lb.pool, lb.params)
    cpu.set_sorted(lb.n_regions, lb.max_sample_len, cfg.sort_voices)
    ld = Driver.open(cfg, lb, mid)
    div, oev = OD.read_midi(mid)
    od = OD.Driver(cfg, ob, div, oev)
    out = np.zeros(cfg.block_frames * 2, np.float32)
    worst, live = -math.inf, 0
    for b in range(220):
        print(f"sf2 block {b}", flush=True)
        more = ld.next_block(gpu, out)
        for idx, table in od.begin_block():
            cpu.set_params_variant(idx, OD.params_array(table, cpu_params_dtype()))
        cpu.set_gates(od.gates.rows.reshape(-1))
        cpu.set_channels(od.chan.rows.reshape(-1), *od.flags)
        st = cpu.stats()
        cpu.spawn(OD.spawn_array(od.admit(st["active_voices"], st["stolen"]), spawn_dtype()))
        stolen_after = cpu.stats()["stolen"]
        oc = cpu.render()
        assert more == od.end_block(oc, stolen_after, cpu.stats()), b
        pg, pc = gpu.pool_export(), cpu.pool_export()
        assert pg.shape == pc.shape and np.array_equal(pg, pc), f"block {b}: pool"
        peak = float(np.abs(oc).max())
        if peak > 0:
            d = float(np.abs(out.astype(np.float64) - oc).max())
            worst = max(worst, -math.inf if d == 0 else 20 * math.log10(d / peak))
        live = max(live, gpu.stats()["active_voices"])
        if not more:
            break
    print(f"sf2: {lb.n_regions} regions, pool {c['pool'].size} frames at {c['pool_rate']} Hz, "
          f"peak live {live}, worst null {worst:.1f} dB")
    assert live > 100 and worst <= NULL_DB
    gpu.close()
