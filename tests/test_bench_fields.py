"""bench.py's counter-profile fields: a committed profile speaks for a run
only on its own workload and within 2 % of its mean voice count."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _profile(name):
    with open(os.path.join(ROOT, "profiles", name)) as f:
        return json.load(f)


def test_committed_profiles_name_their_workload_and_voices():
    for wl, voices in (("c2", 1_000_000), ("c5", 60_000_000)):
        for kind in ("traffic", "valu"):
            d = _profile(f"render_{kind}_{wl}.json")
            assert d["workload"] == wl and d["mean_voices"] > voices
            assert d["kernel"] == "k_render_fast" and "command" in d


def test_valu_fields_follow_the_profile_and_this_runs_render_time():
    d = _profile("render_valu_c2.json")
    mv = d["mean_voices"]
    vpf, frac, src = bench.measured_valu("c2", mv, 0.5)
    assert vpf == d["valu_per_voice_frame"]
    # count / (1024 SIMDs x 2.4 GHz / 4 x 0.5 ms)
    assert abs(frac - d["sq_insts_valu_per_launch"] / (1024 * 2.4e9 / 4 * 0.5e-3)) < 1e-4
    assert src.startswith("profiles/render_valu_c2.json")
    # Twice the time, half the fraction; 1 % more voices, 1 % more instructions.
    assert abs(bench.measured_valu("c2", mv, 1.0)[1] - frac / 2) < 1e-4
    assert abs(bench.measured_valu("c2", mv * 1.01, 0.5)[1] - frac * 1.01) < 1e-4
    # Per voice-frame: a wave64 instruction serves 64 voices.
    assert abs(vpf - d["sq_insts_valu_per_launch"] * 64 / (mv * d["frames"])) < 1e-2


def test_profiles_of_another_workload_or_size_give_nothing():
    mv = _profile("render_valu_c2.json")["mean_voices"]
    assert bench.measured_valu("c3", mv, 0.5) == (None, None, None)
    assert bench.measured_valu("c2", mv * 0.97, 0.5) == (None, None, None)
    assert bench.measured_valu("c2", mv, None) == (None, None, None)
    mt = _profile("render_traffic_c2.json")["mean_voices"]
    assert bench.measured_traffic("c2", mt)[0] == _profile("render_traffic_c2.json")["bytes_per_launch"]
    assert bench.measured_traffic("c2", mt * 1.05) == (None, None)
    assert bench.measured_traffic("c1", mt) == (None, None)


def test_more_than_one_gpu_exits_with_a_message():
    import subprocess
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "measures one GPU" in r.stdout
