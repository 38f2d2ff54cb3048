// kst_bank.cpp -- the flattened instrument bank the driver builds voices from
// (src/bank.rs): params table layout and design, the admission preview tables,
// note-on preview and layer build, preset lookup.
//
// Floating point follows the reference operation for operation (f32 where it
// is f32, f64 where it is f64, no contraction: the library is built with
// -ffp-contract=off). Its transcendental calls are the ones the reference's
// Rust lowers to on Linux: `2f32.powf(x)` and `2f64.powf(x)` become exp2f /
// exp2 (LLVM rewrites a power of the constant 2 into exp2), `10f32.powf(x)`
// stays powf (glibc's exp10 is not used on Linux), `x.sin_cos()` is sincosf.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "kst_host.h"

namespace kst {
namespace {

constexpr float CC_CUTOFF_CENTS_PER_STEP = 75.0f;        // src/bank.rs:290
constexpr float CC_Q_CB_PER_STEP = 240.0f / 63.0f;       // src/bank.rs:298
constexpr float CC_ENV_STEPS_PER_OCTAVE = 8.0f;          // src/bank.rs:306
constexpr float CC_ENV_ADD_PER_CUBE = 6.0e-5f;           // src/bank.rs:318
constexpr float CC_MIN_FADE_SECS = 0.004f;               // src/bank.rs:326
constexpr float PI_F32 = 3.14159265358979323846f;
constexpr uint32_t VF_LOOP = 1u, VF_LOOP_UNTIL_RELEASE = 2u;  // src/bank.rs:13-15

uint32_t loop_flags(uint32_t mode) {  // src/bank.rs:24-31
    return mode == KST_LOOP_CONTINUOUS      ? VF_LOOP
           : mode == KST_LOOP_UNTIL_RELEASE ? (VF_LOOP | VF_LOOP_UNTIL_RELEASE)
                                            : 0u;
}

// src/bank.rs:243-252: scale below the neutral 64, add above it.
void time_mod(uint8_t v, float &scale, float &add) {
    const float d = (float)v - 64.0f;
    if (d <= 0.0f) {
        scale = exp2f(d / CC_ENV_STEPS_PER_OCTAVE);
        add = 0.0f;
    } else {
        scale = 1.0f;
        add = CC_ENV_ADD_PER_CUBE * d * d * d;
    }
}

float cents_to_hz(float cents) { return 8.176f * exp2f(cents / 1200.0f); }  // :462-464
float cb_to_gain(float cb) { return powf(10.0f, -cb / 200.0f); }           // :467-469
float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// src/bank.rs:492-511
float velocity_atten_cb(uint8_t vel, float veltrack) {
    const float v = (float)std::max<uint8_t>(vel, 1) / 127.0f;
    const float t = clampf(veltrack, -100.0f, 100.0f) / 100.0f;
    if (t == 1.0f) return clampf(-400.0f * log10f(v), 0.0f, 960.0f);
    const float gain = t >= 0.0f ? 1.0f - t * (1.0f - v * v) : 1.0f + t * v * v;
    if (gain <= 0.0f) return 960.0f;
    return clampf(-200.0f * log10f(gain), 0.0f, 960.0f);
}

// src/bank.rs:963-979, RBJ low-pass normalised against the default Q.
void biquad_lowpass(float fc, float q_cb, float sr, float &b0, float &b1, float &a1, float &a2) {
    const float q_default = 0.70710678118654752f;
    const float q_db = q_cb / 10.0f - 3.01f;
    const float q = std::fmax(powf(10.0f, q_db / 20.0f), 0.001f);
    const float gain = sqrtf(q_default / q);
    const float w0 = 2.0f * PI_F32 * clampf(fc / sr, 1.0e-5f, 0.49f);
    float sin_w0, cos_w0;
    sincosf(w0, &sin_w0, &cos_w0);
    const float alpha = sin_w0 / (2.0f * q);
    const float a0 = 1.0f + alpha;
    b0 = gain * (1.0f - cos_w0) * 0.5f / a0;
    b1 = gain * (1.0f - cos_w0) / a0;
    a1 = -2.0f * cos_w0 / a0;
    a2 = (1.0f - alpha) / a0;
}

// Fixed::from_f64, src/fixed.rs:27-37
uint64_t fixed_from_f64(double v) { return sat_u64(v * 4294967296.0); }

int64_t clamp_i64(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

}  // namespace

ParamMod ParamMod::from_controllers(uint8_t cc71, uint8_t cc72, uint8_t cc73, uint8_t cc74,
                                    uint8_t cc75) {  // src/bank.rs:336-350
    ParamMod m;
    m.cutoff_cents = ((float)cc74 - 64.0f) * CC_CUTOFF_CENTS_PER_STEP;
    m.q_cb = std::fmax((float)cc71 - 64.0f, 0.0f) * CC_Q_CB_PER_STEP;
    time_mod(cc73, m.attack_scale, m.attack_add);
    time_mod(cc75, m.decay_scale, m.decay_add);
    time_mod(cc72, m.release_scale, m.release_add);
    return m;
}

Bank::Bank(const BankSettings &s, const int16_t *pool_in, size_t pool_len, uint32_t pool_rate_in,
           const kst_sample_info *smp, size_t n_samples, const kst_region *reg, size_t n_regions,
           const kst_preset *pre, size_t n_presets, const uint32_t *pre_regions,
           size_t n_pre_regions)
    : pool(pool_in, pool_in + pool_len), pool_rate(pool_rate_in), samples(smp, smp + n_samples),
      regions(reg, reg + n_regions), set(s), preset_desc(pre, pre + n_presets),
      preset_regions(pre_regions, pre_regions + n_pre_regions) {
    if (set.sample_rate == 0) throw std::invalid_argument("sample_rate must be non-zero");
    for (size_t i = 0; i < n_samples; ++i) {
        const kst_sample_info &x = samples[i];
        if ((uint64_t)x.start + x.len > pool_len)
            throw std::invalid_argument("sample " + std::to_string(i) + " runs past the pool (" +
                                        std::to_string(x.start) + " + " + std::to_string(x.len) +
                                        " > " + std::to_string(pool_len) + ")");
        max_sample_len = std::max(max_sample_len, x.len);
    }
    for (size_t i = 0; i < n_regions; ++i)
        if (regions[i].loop_mode > KST_LOOP_UNTIL_RELEASE)
            throw std::invalid_argument("region " + std::to_string(i) + ": loop_mode " +
                                        std::to_string(regions[i].loop_mode) + " is not 0, 1 or 2");
    // Presets and their key index (src/bank.rs:364-382, 514-527).
    presets_.resize(n_presets);
    for (size_t p = 0; p < n_presets; ++p) {
        const kst_preset &d = pre[p];
        if ((uint64_t)d.first + d.count > n_pre_regions)
            throw std::invalid_argument("preset " + std::to_string(p) +
                                        " names regions past the preset_regions list");
        if (d.bank > 0xFFFF || d.program > 0xFFFF)
            throw std::invalid_argument("preset " + std::to_string(p) + ": bank and program are u16");
        PresetEntry &e = presets_[p];
        e.bank = d.bank;
        e.program = d.program;
        std::vector<std::vector<uint32_t>> per_key(128);
        for (uint32_t j = 0; j < d.count; ++j) {
            const uint32_t r = pre_regions[d.first + j];
            if (r >= n_regions)
                throw std::invalid_argument("preset " + std::to_string(p) + " names region " +
                                            std::to_string(r) + " of " + std::to_string(n_regions));
            const kst_region &x = regions[r];
            for (uint32_t k = x.key_lo; k <= std::min<uint32_t>(x.key_hi, 127); ++k) per_key[k].push_back(r);
        }
        e.key_index.reserve(129);
        for (auto &list : per_key) {
            e.key_index.push_back((uint32_t)e.key_regions.size());
            e.key_regions.insert(e.key_regions.end(), list.begin(), list.end());
        }
        e.key_index.push_back((uint32_t)e.key_regions.size());
        index_.emplace_back(((uint64_t)e.bank << 16) | e.program, (uint32_t)p);
    }
    std::stable_sort(index_.begin(), index_.end(),
                     [](const auto &a, const auto &b) { return a.first < b.first; });

    // Params layout (src/bank.rs:689-721): a (velocity, key) grid per region.
    params_base.resize(n_regions);
    params_stride.resize(n_regions);
    params_vel_span.resize(n_regions);
    uint64_t n = 0;
    for (size_t i = 0; i < n_regions; ++i) {
        const kst_region &r = regions[i];
        const bool per_key = r.keynum_to_decay != 0 || r.keynum_to_hold != 0;
        const uint32_t vel_span =
            r.filter_veltrack_cents != 0.0f
                ? std::min<uint32_t>((uint32_t)(r.vel_hi > r.vel_lo ? r.vel_hi - r.vel_lo : 0) + 1, 128)
                : 1;
        params_base[i] = (uint32_t)n;
        params_stride[i] = per_key ? 1 : 0;
        params_vel_span[i] = vel_span;
        n += (uint64_t)vel_span * (per_key ? 128 : 1);
        if (n > UINT32_MAX) throw std::invalid_argument("params table exceeds 2^32 entries");
    }
    build_variant(ParamMod{}, params);

    // Admission tables (src/bank.rs:723-779): exact, not approximate.
    gain_table_.assign(n_regions * 256, 0.0f);
    key_ok_.assign((n_regions * 128 + 63) / 64, 0);
    delay_frames_.resize(n_regions);
    const float sr = (float)set.sample_rate;
    for (size_t ri = 0; ri < n_regions; ++ri) {
        const kst_region &r = regions[ri];
        delay_frames_[ri] = sat_u32(r.delay * sr);
        const float theta = (clampf(r.pan, -1.0f, 1.0f) + 1.0f) * 0.5f * (PI_F32 * 0.5f);
        const float pan_l = cosf(theta), pan_r = sinf(theta);
        for (uint32_t v = 0; v < 128; ++v) {
            const uint8_t eff_vel = r.fixed_vel >= 0 ? (uint8_t)r.fixed_vel : (uint8_t)v;
            const float atten = r.attenuation_cb + velocity_atten_cb(eff_vel, r.amp_veltrack);
            const float gain = cb_to_gain(atten) * set.master_volume;
            gain_table_[(ri * 128 + v) * 2] = gain * pan_l;
            gain_table_[(ri * 128 + v) * 2 + 1] = gain * pan_r;
        }
        if (r.sample >= samples.size() || samples[r.sample].len == 0) continue;
        const kst_sample_info &x = samples[r.sample];
        for (uint32_t k = 0; k < 128; ++k) {
            const uint8_t eff_key = r.fixed_key >= 0 ? (uint8_t)r.fixed_key : (uint8_t)k;
            const double root = r.root_key_override >= 0 ? (double)r.root_key_override : (double)x.root_key;
            const double cents = ((double)eff_key - root) * (double)r.scale_tuning +
                                 (double)r.coarse_tune * 100.0 + (double)r.fine_tune +
                                 (double)x.correction_cents;
            double ratio = exp2(cents / 1200.0);
            ratio *= (pool_rate == 0 ? (double)x.rate : (double)pool_rate) / (double)set.sample_rate;
            if (std::isfinite(ratio) && ratio > 0.0) {
                const size_t bit = ri * 128 + k;
                key_ok_[bit / 64] |= 1ull << (bit % 64);
            }
        }
    }
    // Layer build tables: every (region, key) that builds, through the same
    // build_voice, so a lookup returns exactly what the call would.
    plan_.assign(n_regions, RegionPlan{});
    step_.assign(n_regions * 128, 0);
    for (size_t ri = 0; ri < n_regions; ++ri) {
        bool first = true;
        for (uint32_t k = 0; k < 128; ++k) {
            if (!key_ok((uint32_t)ri, (uint8_t)k)) continue;
            Layer v;
            build_voice(regions[ri], (uint32_t)ri, (uint8_t)k, 64, v);
            step_[ri * 128 + k] = v.step;
            if (first) {
                plan_[ri] = RegionPlan{(uint32_t)(v.phase >> 32), v.smp_len, v.loop_start,
                                       v.loop_end, v.flags};
                first = false;
            }
        }
    }
}

int64_t Bank::find_preset(uint32_t bank, uint32_t program) const {  // src/bank.rs:529-546
    auto find = [&](uint64_t key) -> int64_t {
        auto it = std::lower_bound(index_.begin(), index_.end(), key,
                                   [](const auto &e, uint64_t k) { return e.first < k; });
        return it != index_.end() && it->first == key ? (int64_t)it->second : -1;
    };
    int64_t p = find(((uint64_t)bank << 16) | program);
    if (p >= 0) return p;
    p = find(program);
    if (p >= 0) return p;
    if (bank == 128)
        for (const auto &e : index_)
            if ((e.first >> 16) == 128) return e.second;
    return index_.empty() ? -1 : (int64_t)index_.front().second;
}

void Bank::preview_note_on(uint32_t preset, uint8_t key, uint8_t vel, size_t max_layers,
                           std::vector<Preview> &out) const {  // src/bank.rs:793-834
    if (preset >= presets_.size()) return;
    const PresetEntry &p = presets_[preset];
    key = std::min<uint8_t>(key, 127);
    size_t layers = 0;
    for (uint32_t j = p.key_index[key]; j < p.key_index[key + 1]; ++j) {
        if (layers >= max_layers) break;
        const uint32_t ri = p.key_regions[j];
        const kst_region &r = regions[ri];
        if (vel < r.vel_lo || vel > r.vel_hi) continue;
        if (!key_ok(ri, key)) continue;
        const size_t g = ((size_t)ri * 128 + vel) * 2;
        out.push_back(Preview{ri, gain_table_[g], gain_table_[g + 1], delay_frames_[ri]});
        ++layers;
    }
}

bool Bank::build_layer(uint32_t region, uint8_t key, uint8_t vel, Layer &o) const {
    if (region >= regions.size()) return false;  // src/bank.rs:585-594
    key = std::min<uint8_t>(key, 127);
    if (vel > 127) return build_voice(regions[region], region, key, vel, o);
    if (!key_ok(region, key)) return false;  // exactly where build_voice gives up
    const kst_region &r = regions[region];
    const RegionPlan &p = plan_[region];
    const uint8_t eff_key = r.fixed_key >= 0 ? (uint8_t)r.fixed_key : key;
    const uint32_t span = params_vel_span[region];
    const uint32_t vi = span > 1 ? std::min<uint32_t>(vel > r.vel_lo ? vel - r.vel_lo : 0, span - 1) : 0;
    const size_t g = ((size_t)region * 128 + vel) * 2;
    o.phase = (uint64_t)p.start_offset << 32;
    o.step = step_[(size_t)region * 128 + key];
    o.smp_base = samples[r.sample].start;
    o.smp_len = p.smp_len;
    o.loop_start = p.loop_start;
    o.loop_end = p.loop_end;
    o.flags = p.flags;
    o.params = params_base[region] + vi * (params_stride[region] ? 128u : 1u) +
               (params_stride[region] ? (uint32_t)eff_key : 0u);
    o.region = region;
    o.gain_l = gain_table_[g];
    o.gain_r = gain_table_[g + 1];
    o.delay_frames = delay_frames_[region];
    o.exclusive_class = r.exclusive_class;
    return true;
}

bool Bank::build_voice(const kst_region &r, uint32_t ri, uint8_t key, uint8_t vel,
                       Layer &o) const {  // src/bank.rs:596-687
    if (r.sample >= samples.size()) return false;
    const kst_sample_info &s = samples[r.sample];
    if (s.len == 0) return false;
    const uint8_t eff_key = r.fixed_key >= 0 ? (uint8_t)r.fixed_key : key;
    const uint8_t eff_vel = r.fixed_vel >= 0 ? (uint8_t)r.fixed_vel : vel;
    const double root = r.root_key_override >= 0 ? (double)r.root_key_override : (double)s.root_key;
    const double cents = ((double)eff_key - root) * (double)r.scale_tuning +
                         (double)r.coarse_tune * 100.0 + (double)r.fine_tune +
                         (double)s.correction_cents;
    double ratio = exp2(cents / 1200.0);
    ratio *= (pool_rate == 0 ? (double)s.rate : (double)pool_rate) / (double)set.sample_rate;
    if (!(std::isfinite(ratio) && ratio > 0.0)) return false;

    const float atten = r.attenuation_cb + velocity_atten_cb(eff_vel, r.amp_veltrack);
    const float gain = cb_to_gain(atten) * set.master_volume;
    const float theta = (clampf(r.pan, -1.0f, 1.0f) + 1.0f) * 0.5f * (PI_F32 * 0.5f);
    o.gain_l = gain * cosf(theta);
    o.gain_r = gain * sinf(theta);

    auto scale = [&](int32_t v) { return sat_i64(roundf((float)v * s.resample_ratio)); };
    const int64_t len = s.len;
    const uint32_t start_offset = (uint32_t)clamp_i64(scale(r.addr_start), 0, len - 1);
    const uint32_t smp_len = (uint32_t)clamp_i64(len + scale(r.addr_end), 1, len);
    const uint32_t loop_start =
        (uint32_t)clamp_i64((int64_t)s.loop_start + scale(r.addr_loop_start), 0, (int64_t)smp_len - 1);
    const uint32_t loop_end =
        (uint32_t)clamp_i64((int64_t)s.loop_end + scale(r.addr_loop_end), 0, (int64_t)smp_len);

    const uint32_t keys = params_stride[ri] ? 128u : 1u;
    const uint32_t ki = params_stride[ri] ? (uint32_t)eff_key : 0u;
    const uint32_t span = params_vel_span[ri];
    const uint32_t vi = span > 1 ? std::min<uint32_t>(vel > r.vel_lo ? vel - r.vel_lo : 0, span - 1) : 0;

    uint32_t flags = loop_flags(r.loop_mode);
    if (r.loop_mode != KST_LOOP_NONE && loop_end <= loop_start + 1) flags = 0;

    o.phase = (uint64_t)start_offset << 32;
    o.step = fixed_from_f64(ratio);
    o.smp_base = s.start;
    o.smp_len = smp_len;
    o.loop_start = loop_start;
    o.loop_end = loop_end;
    o.flags = flags;
    o.params = params_base[ri] + vi * keys + ki;
    o.region = ri;
    o.delay_frames = sat_u32(r.delay * (float)set.sample_rate);
    o.exclusive_class = r.exclusive_class;
    return true;
}

void Bank::build_variant(const ParamMod &m, std::vector<kst_region_params> &out) const {
    out.clear();  // src/bank.rs:836-853, the layout of variant 0
    for (size_t i = 0; i < regions.size(); ++i) {
        const kst_region &r = regions[i];
        for (uint32_t v = 0; v < params_vel_span[i]; ++v) {
            const uint8_t vel = (uint8_t)std::min<uint32_t>(r.vel_lo + v, 127);
            if (params_stride[i])
                for (uint32_t k = 0; k < 128; ++k) out.push_back(make_params(r, (uint8_t)k, vel, m));
            else
                out.push_back(make_params(r, 60, vel, m));
        }
    }
}

kst_region_params Bank::make_params(const kst_region &r, uint8_t key, uint8_t vel,
                                    const ParamMod &m) const {  // src/bank.rs:876-955
    const float sr = (float)set.sample_rate;
    const float key_delta = 60.0f - (float)key;
    const float hold = r.hold * exp2f((float)r.keynum_to_hold * key_delta / 1200.0f);
    const float decay = r.decay * exp2f((float)r.keynum_to_decay * key_delta / 1200.0f);
    auto floor_t = [](float t, float base) { return std::fmax(t, std::fmin(CC_MIN_FADE_SECS, base)); };

    const float attack_frames = std::fmax((r.attack * m.attack_scale + m.attack_add) * sr, 0.0f);
    const float hold_frames = std::fmax(hold * sr, 0.0f);
    const float decay_frames =
        std::fmax(floor_t(decay * m.decay_scale + m.decay_add, decay) * sr, 1.0f);
    const float release_frames =
        std::fmax(floor_t(r.release * m.release_scale + m.release_add, r.release) * sr, 1.0f);

    kst_region_params p{};
    p.attack_rate = attack_frames < 1.0f ? 1.0f : 1.0f / attack_frames;
    p.attack_end = 1.0f + hold_frames * p.attack_rate;
    const float sustain = clampf(r.sustain, 0.0f, 1.0f);
    auto coef = [](uint32_t curve, float frames) {
        return curve == 0 ? powf(10.0f, -100.0f / (20.0f * frames)) : 1.0f / frames;
    };
    p.decay_coef = coef(set.decay_curve, decay_frames);
    p.release_coef = coef(set.release_curve, release_frames);

    const float fc_cents =
        r.filter_fc_cents + r.filter_veltrack_cents * ((float)vel / 127.0f) + m.cutoff_cents;
    const float fc = cents_to_hz(fc_cents);
    const float nyq_guard = sr * 0.49f;
    const bool use_filter = set.filter_enabled && fc_cents < 13500.0f && fc < nyq_guard && fc > 20.0f;
    if (use_filter) {
        biquad_lowpass(fc, r.filter_q_cb + m.q_cb, sr, p.b0, p.b1, p.a1, p.a2);
    } else {
        p.b0 = 1.0f;
        p.b1 = p.a1 = p.a2 = 0.0f;
    }
    p.decay_target = std::fmax(sustain, set.env_floor);
    p.sustain = sustain;
    p.flags = use_filter ? 1u : 0u;
    p.pad = 0;
    return p;
}

}  // namespace kst
