// kst_post.cpp -- the post-render step on the host: the reference's limiters
// (src/limiter.rs), the hard clamp, a streaming WAV writer (src/wav.rs) and
// the `kestrel null` comparison (src/bin/kestrel.rs:597-646).
//
// These run on the mixed stereo block, a sequential recurrence over 2 * F
// samples per block, so they belong on the host next to the caller, as in the
// reference. Every update is the reference's f64 arithmetic in the same order
// (built without contraction), so a limited block is bit-for-bit what the
// reference's limiter makes of the same input.
#include "kestrel_hip.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <new>
#include <utility>
#include <vector>

namespace {

// src/limiter.rs:14-118 -- the follower ported from OmniConverter.
struct Omni {
    double loudness_l = 1.0, loudness_r = 1.0, velocity_l = 0.0, velocity_r = 0.0;
    double attack = 100.0, falloff, min_thresh = 0.4, strength = 1.0;
    bool reduce_high_pitch = false;
    double velocity_thresh = 1.0;
    bool first_sample = true;

    explicit Omni(uint32_t sr) : falloff((double)sr / 3.0) {}

    static double follow(double state, double x, double up, double down) {
        return state > x ? (state * down + x) / (down + 1.0) : (state * up + x) / (up + 1.0);
    }

    void process(float *buf, size_t n) {
        for (size_t i = 0; i + 1 < n; i += 2) {
            const double in_l = buf[i], in_r = buf[i + 1];
            loudness_l = follow(loudness_l, std::fabs(in_l), attack, falloff);
            loudness_r = follow(loudness_r, std::fabs(in_r), attack, falloff);
            if (loudness_l < min_thresh) loudness_l = min_thresh;
            if (loudness_r < min_thresh) loudness_r = min_thresh;
            double l = in_l / (loudness_l * strength + 2.0 * (1.0 - strength)) / 2.0;
            double r = in_r / (loudness_r * strength + 2.0 * (1.0 - strength)) / 2.0;
            if (!first_sample) {
                velocity_l = follow(velocity_l, std::fabs(in_l - l), attack, falloff);
                velocity_r = follow(velocity_r, std::fabs(in_r - r), attack, falloff);
            }
            first_sample = false;
            if (reduce_high_pitch) {
                if (velocity_l > velocity_thresh) l = l / velocity_l * velocity_thresh;
                if (velocity_r > velocity_thresh) r = r / velocity_r * velocity_thresh;
            }
            buf[i] = (float)l;
            buf[i + 1] = (float)r;
        }
    }
};

constexpr int TP_PHASES = 4, TP_TAPS = 8;

// src/limiter.rs:448-471: a 4x polyphase Blackman-windowed sinc, each phase
// normalised to unity gain.
void design_polyphase(double out[TP_PHASES][TP_TAPS]) {
    const double n = (double)(TP_TAPS * TP_PHASES);
    const double pi = 3.14159265358979323846;
    for (int p = 0; p < TP_PHASES; ++p) {
        for (int t = 0; t < TP_TAPS; ++t) {
            const double x = (double)t - ((double)TP_TAPS / 2.0 - 1.0) - (double)p / (double)TP_PHASES;
            const double sinc = std::fabs(x) < 1e-12 ? 1.0 : std::sin(pi * x) / (pi * x);
            const double k = (double)(t * TP_PHASES + p);
            const double w = 0.42 - 0.5 * std::cos(2.0 * pi * k / (n - 1.0)) +
                             0.08 * std::cos(4.0 * pi * k / (n - 1.0));
            out[p][t] = sinc * w;
        }
        double sum = 0.0;
        for (int t = 0; t < TP_TAPS; ++t) sum += out[p][t];
        if (std::fabs(sum) > 1e-12)
            for (int t = 0; t < TP_TAPS; ++t) out[p][t] /= sum;
    }
}

// src/limiter.rs:203-446 -- the lookahead true-peak brickwall.
struct Brickwall {
    double ceiling;
    size_t look;
    double release_coef, sustain_atk = 0.0, sustain_rel = 0.0;
    double g_slow = 1.0, g_fast = 1.0, gain = 1.0;
    std::vector<float> delay;
    size_t dpos = 0;
    std::deque<std::pair<double, uint64_t>> dq;  // decreasing (value, index)
    std::vector<double> gr_ring;
    size_t gr_pos = 0;
    double gr_sum;
    double hist[2][TP_TAPS] = {};
    double poly[TP_PHASES][TP_TAPS];
    size_t delay_frames;
    bool true_peak;
    uint64_t idx = 0;

    Brickwall(uint32_t sample_rate, double ceil, double lookahead_ms, double release_ms,
              double sustain_ms, bool tp)
        : ceiling(ceil), true_peak(tp) {
        const double sr = (double)sample_rate;
        const double lk = std::round(lookahead_ms * 1e-3 * sr);
        look = lk >= 1.0 ? (size_t)lk : 1;  // `as usize` saturates at 0
        const double rel = std::fmax(release_ms * 1e-3 * sr, 1.0);
        if (sustain_ms > 0.0) {
            const double a = std::fmax(sustain_ms * 1e-3 * sr, 1.0);
            sustain_atk = 1.0 - std::exp(-1.0 / a);
            sustain_rel = 1.0 - std::exp(-1.0 / (a * 4.0));
        }
        release_coef = 1.0 - std::exp(-1.0 / rel);
        const size_t group = true_peak ? TP_TAPS / 2 - 1 : 0;
        delay_frames = look + group;
        delay.assign(delay_frames * 2, 0.0f);
        gr_ring.assign(look + 1, 1.0);
        gr_sum = (double)(look + 1);
        design_polyphase(poly);
    }

    double detect(double l, double r) {
        if (!true_peak) return std::fmax(std::fabs(l), std::fabs(r));
        double peak = 0.0;
        const double v2[2] = {l, r};
        for (int ch = 0; ch < 2; ++ch) {
            double *h = hist[ch];
            for (int t = TP_TAPS - 1; t > 0; --t) h[t] = h[t - 1];
            h[0] = v2[ch];
            for (int p = 0; p < TP_PHASES; ++p) {
                double acc = 0.0;
                for (int t = 0; t < TP_TAPS; ++t) acc += poly[p][t] * h[t];
                const double a = std::fabs(acc);
                if (a > peak) peak = a;
            }
        }
        return peak;
    }

    void process(float *buf, size_t n) {
        for (size_t i = 0; i + 1 < n; i += 2) {
            const double in_l = buf[i], in_r = buf[i + 1];
            const size_t d = dpos * 2;
            const double out_l = delay[d], out_r = delay[d + 1];
            delay[d] = buf[i];
            delay[d + 1] = buf[i + 1];
            dpos = (dpos + 1) % delay_frames;

            const double det = detect(in_l, in_r);
            while (!dq.empty() && dq.back().first <= det) dq.pop_back();
            dq.emplace_back(det, idx);
            const uint64_t oldest = idx > look ? idx - look : 0;
            while (!dq.empty() && dq.front().second < oldest) dq.pop_front();
            const double env = dq.empty() ? 0.0 : dq.front().first;

            const double gr = env > ceiling ? ceiling / env : 1.0;
            gr_sum += gr - gr_ring[gr_pos];
            gr_ring[gr_pos] = gr;
            gr_pos = (gr_pos + 1) % gr_ring.size();
            const double ma = gr_sum / (double)gr_ring.size();

            if (sustain_atk > 0.0) {
                const double c = ma < g_slow ? sustain_atk : sustain_rel;
                g_slow += (ma - g_slow) * c;
                g_slow = g_slow < 1e-9 ? 1e-9 : (g_slow > 1.0 ? 1.0 : g_slow);
            }
            const double q = ma / g_slow;
            const double req = q < 1.0 ? q : 1.0;  // f64::min; ma and g_slow are finite
            if (req < g_fast) {
                g_fast = req;
            } else {
                g_fast += (req - g_fast) * release_coef;
                if (g_fast > req) g_fast = req;
            }
            gain = g_slow * g_fast;
            if (gain > ma) gain = ma;

            buf[i] = (float)(out_l * gain);
            buf[i + 1] = (float)(out_r * gain);
            ++idx;
        }
    }
};

thread_local char g_post_error[256];

int post_fail(int code, const char *msg) {
    std::snprintf(g_post_error, sizeof g_post_error, "%s", msg);
    return code;
}

}  // namespace

struct kst_limiter {
    uint32_t mode;
    Omni *omni = nullptr;
    Brickwall *bw = nullptr;
    ~kst_limiter() {
        delete omni;
        delete bw;
    }
};

struct kst_wav {
    FILE *f = nullptr;
    uint32_t channels = 2, float32 = 1;
    uint64_t data_bytes = 0;
};

extern "C" {

void kst_default_limiter_config(kst_limiter_config *c) {
    if (!c) return;
    // src/config.rs:388-395
    c->sample_rate = 48000;
    c->mode = KST_LIMITER_BRICKWALL;
    c->ceiling_db = 0.0;
    c->lookahead_ms = 2.0;
    c->release_ms = 60.0;
    c->sustain_ms = 0.0;
    c->true_peak = 1;
    c->pad = 0;
}

int kst_limiter_create(const kst_limiter_config *c, kst_limiter **out) {
    if (!c || !out) return post_fail(KST_EINVAL, "null argument");
    *out = nullptr;
    if (c->mode > KST_LIMITER_BRICKWALL) return post_fail(KST_EINVAL, "unknown limiter mode");
    if (c->sample_rate == 0) return post_fail(KST_EINVAL, "sample_rate must be non-zero");
    auto *l = new (std::nothrow) kst_limiter();
    if (!l) return post_fail(KST_EDEVICE, "out of host memory");
    l->mode = c->mode;
    if (c->mode == KST_LIMITER_OMNI) l->omni = new Omni(c->sample_rate);
    if (c->mode == KST_LIMITER_BRICKWALL)
        // src/config.rs:528-530: the ceiling in dBFS as a gain.
        l->bw = new Brickwall(c->sample_rate, std::pow(10.0, c->ceiling_db / 20.0),
                              c->lookahead_ms, c->release_ms, c->sustain_ms, c->true_peak != 0);
    *out = l;
    return KST_OK;
}

int kst_limiter_process(kst_limiter *l, float *buf, size_t n) {
    if (!l || (!buf && n)) return post_fail(KST_EINVAL, "null argument");
    if (n % 2) return post_fail(KST_EINVAL, "the block is interleaved stereo: even length");
    if (l->omni) l->omni->process(buf, n);
    if (l->bw) l->bw->process(buf, n);
    return KST_OK;
}

size_t kst_limiter_latency(const kst_limiter *l) {
    return l && l->bw ? l->bw->delay_frames : 0;
}

void kst_limiter_destroy(kst_limiter *l) { delete l; }

uint64_t kst_clamp_block(float *buf, size_t n) {
    // src/limiter.rs:120-129
    uint64_t moved = 0;
    for (size_t i = 0; buf && i < n; ++i) {
        const float v = buf[i];
        if (v > 1.0f || v < -1.0f) {
            ++moved;
            buf[i] = v > 1.0f ? 1.0f : -1.0f;
        }
    }
    return moved;
}

int kst_wav_create(const char *path, uint32_t sample_rate, uint32_t channels, uint32_t float32,
                   kst_wav **out) {
    if (!path || !out) return post_fail(KST_EINVAL, "null argument");
    *out = nullptr;
    if (channels == 0 || channels > 65535) return post_fail(KST_EINVAL, "channel count out of range");
    FILE *f = std::fopen(path, "wb");
    if (!f) return post_fail(KST_EDEVICE, "cannot create the WAV file");
    // src/wav.rs:67-82: RIFF, a 16-byte fmt chunk, then data; sizes patched at the end.
    const uint16_t bits = float32 ? 32 : 16, tag = float32 ? 3 : 1;
    const uint32_t block_align = channels * (bits / 8u);
    unsigned char h[44];
    auto put32 = [&](int at, uint32_t v) { for (int k = 0; k < 4; ++k) h[at + k] = (unsigned char)(v >> (8 * k)); };
    auto put16 = [&](int at, uint16_t v) { h[at] = (unsigned char)v; h[at + 1] = (unsigned char)(v >> 8); };
    std::memcpy(h, "RIFF", 4);
    put32(4, 0);
    std::memcpy(h + 8, "WAVEfmt ", 8);
    put32(16, 16);
    put16(20, tag);
    put16(22, (uint16_t)channels);
    put32(24, sample_rate);
    put32(28, sample_rate * block_align);
    put16(32, (uint16_t)block_align);
    put16(34, bits);
    std::memcpy(h + 36, "data", 4);
    put32(40, 0);
    if (std::fwrite(h, 1, 44, f) != 44) {
        std::fclose(f);
        return post_fail(KST_EDEVICE, "writing the WAV header");
    }
    auto *w = new kst_wav();
    w->f = f;
    w->channels = channels;
    w->float32 = float32 != 0;
    *out = w;
    return KST_OK;
}

int kst_wav_write(kst_wav *w, const float *samples, size_t n) {
    if (!w || (!samples && n)) return post_fail(KST_EINVAL, "null argument");
    if (w->float32) {
        if (std::fwrite(samples, 4, n, w->f) != n) return post_fail(KST_EDEVICE, "writing WAV data");
        w->data_bytes += 4ull * n;
    } else {
        // src/wav.rs:104-111: clamp, scale by 32767, round half away from zero.
        std::vector<unsigned char> b(2 * n);
        for (size_t i = 0; i < n; ++i) {
            float s = samples[i];
            s = s < -1.0f ? -1.0f : (s > 1.0f ? 1.0f : s);
            const int16_t v = s == s ? (int16_t)std::round(s * 32767.0f) : 0;  // NaN -> 0
            b[2 * i] = (unsigned char)(uint16_t)v;
            b[2 * i + 1] = (unsigned char)((uint16_t)v >> 8);
        }
        if (std::fwrite(b.data(), 1, b.size(), w->f) != b.size())
            return post_fail(KST_EDEVICE, "writing WAV data");
        w->data_bytes += b.size();
    }
    return KST_OK;
}

int kst_wav_finish(kst_wav *w) {
    if (!w) return post_fail(KST_EINVAL, "null argument");
    // src/wav.rs:132-155: sizes saturate at 4 GiB.
    const uint64_t riff = 36ull + w->data_bytes;
    const uint32_t r32 = riff > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)riff;
    const uint32_t d32 = w->data_bytes > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)w->data_bytes;
    unsigned char b[4];
    int rc = KST_OK;
    for (int k = 0; k < 4; ++k) b[k] = (unsigned char)(r32 >> (8 * k));
    if (std::fseek(w->f, 4, SEEK_SET) || std::fwrite(b, 1, 4, w->f) != 4) rc = KST_EDEVICE;
    for (int k = 0; k < 4; ++k) b[k] = (unsigned char)(d32 >> (8 * k));
    if (std::fseek(w->f, 40, SEEK_SET) || std::fwrite(b, 1, 4, w->f) != 4) rc = KST_EDEVICE;
    if (std::fclose(w->f)) rc = KST_EDEVICE;
    delete w;
    return rc == KST_OK ? KST_OK : post_fail(rc, "finishing the WAV file");
}

int kst_null(const float *a, size_t na, const float *b, size_t nb, kst_null_result *out) {
    if (!out || (!a && na) || (!b && nb)) return post_fail(KST_EINVAL, "null argument");
    // src/bin/kestrel.rs:597-646: the first min(na, nb) samples; `a` is the reference.
    const size_t n = na < nb ? na : nb;
    double peak_diff = 0.0, sum_sq = 0.0, peak_ref = 0.0;
    uint64_t worst = 0;
    bool identical = true;
    for (size_t i = 0; i < n; ++i) {
        const double d = std::fabs((double)(a[i] - b[i]));
        if (d > peak_diff) {
            peak_diff = d;
            worst = i;
        }
        sum_sq += d * d;
        peak_ref = std::fmax(peak_ref, (double)std::fabs(a[i]));
        identical = identical && a[i] == b[i];
    }
    out->compared = n;
    out->worst = worst;
    out->peak_ref = peak_ref;
    out->peak_diff = peak_diff;
    out->rms_diff = std::sqrt(sum_sq / (double)(n > 0 ? n : 1));
    out->identical = identical ? 1u : 0u;
    out->pad = 0;
    return KST_OK;
}

const char *kst_post_last_error(void) { return g_post_error; }

}  // extern "C"
