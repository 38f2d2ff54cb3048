// kst_sfz.cpp -- the SFZ loader and the WAV reader it needs (src/sfz.rs,
// src/wav.rs:166-365).
//
// The text is split into (header, opcodes) sections in file order, #include
// files spliced in where they stand; aliases are collapsed at the level that
// wrote them; one walk then carries <control> default_path, <global>,
// <master> and <group> down to every <region>. Each referenced WAV is decoded
// once per channel (a stereo file becomes two hard-panned regions), converted
// to the pool rate, and laid out with guard frames; every region lands in one
// preset, bank 0 program 0, built by the same Bank constructor.
#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

#include "kst_host.h"

namespace kst {
namespace {

constexpr size_t POOL_GUARD = 8;

// ---------------------------------------------------------------------------
// WAV (src/wav.rs:203-365): PCM 8/16/24/32, float 32/64, extensible; the smpl
// chunk's root key, fine tune and first loop.
// ---------------------------------------------------------------------------
struct Wav {
    uint32_t sample_rate = 0;
    uint16_t channels = 0;
    std::vector<float> interleaved;
    bool has_loop = false;
    uint32_t loop_start = 0, loop_end = 0;
    int root_key = -1;
    float fine_tune_cents = 0.0f;

    std::vector<int16_t> channel_i16(size_t ch) const {
        const size_t nch = channels;
        ch = std::min(ch, nch > 0 ? nch - 1 : 0);
        std::vector<int16_t> out;
        for (size_t i = ch; i < interleaved.size(); i += std::max<size_t>(nch, 1)) {
            float s = interleaved[i];
            s = s < -1.0f ? -1.0f : (s > 1.0f ? 1.0f : s);  // NaN passes, then rounds to 0
            const float v = std::round(s * 32767.0f);
            out.push_back(v != v ? 0 : (int16_t)v);
        }
        return out;
    }
};

uint32_t le32(const uint8_t *p) { return (uint32_t)p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); }
uint16_t le16(const uint8_t *p) { return (uint16_t)(p[0] | (p[1] << 8)); }

Wav read_wav(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("opening " + path + ": " + std::strerror(errno));
    std::vector<uint8_t> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    const uint64_t n = b.size();
    size_t pos = 0;
    auto need = [&](size_t k) {
        if (pos + k > n) throw std::runtime_error(path + ": truncated");
    };
    need(12);
    if (std::memcmp(&b[0], "RIFF", 4) != 0) throw std::runtime_error(path + ": not a RIFF file");
    if (std::memcmp(&b[8], "WAVE", 4) != 0) throw std::runtime_error(path + ": not a WAVE file");
    pos = 12;
    Wav w;
    uint16_t tag = 0, bits = 0;
    bool have_data = false;
    std::vector<uint8_t> data;
    for (;;) {
        if (pos + 4 > n) break;  // no further tag
        const size_t tag_at = pos;
        pos += 4;
        need(4);
        const uint64_t size = le32(&b[pos]);
        pos += 4;
        const uint64_t padded = size + (size & 1);
        if (size > n)
            throw std::runtime_error(path + ": chunk " + std::string((char *)&b[tag_at], 4) + " claims " +
                                     std::to_string(size) + " bytes");
        if (std::memcmp(&b[tag_at], "fmt ", 4) == 0) {
            need(16);
            tag = le16(&b[pos]);
            w.channels = le16(&b[pos + 2]);
            w.sample_rate = le32(&b[pos + 4]);
            bits = le16(&b[pos + 14]);
            if (tag == 0xFFFE && size >= 40) {
                need(26);
                tag = le16(&b[pos + 24]);  // the sub-format GUID's first two bytes
            }
            pos += padded;
        } else if (std::memcmp(&b[tag_at], "data", 4) == 0) {
            need(size);
            data.assign(b.begin() + pos, b.begin() + pos + size);
            have_data = true;
            pos += padded;
        } else if (std::memcmp(&b[tag_at], "smpl", 4) == 0) {
            need(padded);
            const uint8_t *c = &b[pos];
            if (padded >= 36) {
                const uint32_t note = le32(c + 20), frac = le32(c + 24), loops = le32(c + 28);
                if (note < 128) w.root_key = (int)note;
                w.fine_tune_cents = (float)((double)frac / 4294967296.0 * 100.0);
                if (loops > 0 && padded >= 60) {
                    w.has_loop = true;
                    w.loop_start = le32(c + 44);
                    w.loop_end = le32(c + 48);
                }
            }
            pos += padded;
        } else {
            pos += padded;
        }
    }
    if (!have_data) throw std::runtime_error(path + ": no data chunk");
    if (w.channels == 0) throw std::runtime_error(path + ": no fmt chunk");
    auto &o = w.interleaved;
    const size_t m = data.size();
    if (tag == 1 && bits == 8) {
        for (size_t i = 0; i < m; ++i) o.push_back(((float)data[i] - 128.0f) / 128.0f);
    } else if (tag == 1 && bits == 16) {
        for (size_t i = 0; i + 2 <= m; i += 2) o.push_back((float)(int16_t)le16(&data[i]) / 32768.0f);
    } else if (tag == 1 && bits == 24) {
        for (size_t i = 0; i + 3 <= m; i += 3) {
            const int32_t v = (int32_t)(((uint32_t)data[i] << 8) | ((uint32_t)data[i + 1] << 16) |
                                        ((uint32_t)data[i + 2] << 24)) >> 8;
            o.push_back((float)v / 8388608.0f);
        }
    } else if (tag == 1 && bits == 32) {
        for (size_t i = 0; i + 4 <= m; i += 4) o.push_back((float)(int32_t)le32(&data[i]) / 2147483648.0f);
    } else if (tag == 3 && bits == 32) {
        for (size_t i = 0; i + 4 <= m; i += 4) {
            float v;
            std::memcpy(&v, &data[i], 4);
            o.push_back(v);
        }
    } else if (tag == 3 && bits == 64) {
        for (size_t i = 0; i + 8 <= m; i += 8) {
            double v;
            std::memcpy(&v, &data[i], 8);
            o.push_back((float)v);
        }
    } else {
        throw std::runtime_error(path + ": unsupported format tag " + std::to_string(tag) + " / " +
                                 std::to_string(bits) + " bits");
    }
    return w;
}

// ---------------------------------------------------------------------------
// Opcode values, with the reference's parsing rules (decimal only, as Rust's
// str::parse reads them).
// ---------------------------------------------------------------------------
bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }

std::string trim(const std::string &s) {
    size_t a = 0, e = s.size();
    while (a < e && is_ws(s[a])) ++a;
    while (e > a && is_ws(s[e - 1])) --e;
    return s.substr(a, e - a);
}

std::string lower(std::string s) {
    for (char &c : s)
        if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
    return s;
}

bool parse_i32(const std::string &s, int32_t &out) {
    if (s.empty()) return false;
    size_t i = 0;
    bool neg = false;
    if (s[0] == '+' || s[0] == '-') {
        neg = s[0] == '-';
        i = 1;
    }
    if (i == s.size()) return false;
    int64_t v = 0;
    for (; i < s.size(); ++i) {
        if (s[i] < '0' || s[i] > '9') return false;
        v = v * 10 + (s[i] - '0');
        if (v > 2147483648ll) return false;
    }
    v = neg ? -v : v;
    if (v > INT32_MAX || v < INT32_MIN) return false;
    out = (int32_t)v;
    return true;
}

bool parse_f32(const std::string &s, float &out) {
    if (s.empty() || is_ws(s[0])) return false;
    for (char c : s)
        if (c == 'x' || c == 'X' || c == 'p' || c == 'P') return false;  // no hex floats
    const std::string l = lower(s);
    const std::string body = (l[0] == '+' || l[0] == '-') ? l.substr(1) : l;
    if (body == "inf" || body == "infinity" || body == "nan") {
        out = body == "nan" ? NAN : (l[0] == '-' ? -INFINITY : INFINITY);
        return true;
    }
    for (char c : body)
        if (!((c >= '0' && c <= '9') || c == '.' || c == 'e' || c == '+' || c == '-')) return false;
    char *end = nullptr;
    errno = 0;
    const float v = std::strtof(s.c_str(), &end);
    if (end != s.c_str() + s.size()) return false;
    out = v;
    return true;
}

int32_t f32_to_i32(float v) {  // Rust's saturating `as`
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}

// c4 = 60 (src/sfz.rs:88-116)
bool parse_note_name(const std::string &s, int32_t &out) {
    if (s.empty()) return false;
    int step;
    switch (s[0] | 0x20) {
    case 'c': step = 0; break;
    case 'd': step = 2; break;
    case 'e': step = 4; break;
    case 'f': step = 5; break;
    case 'g': step = 7; break;
    case 'a': step = 9; break;
    case 'b': step = 11; break;
    default: return false;
    }
    size_t i = 1;
    int acc = 0;
    while (i < s.size() && (s[i] == '#' || s[i] == 'b' || (s[i] == '-' && i == 1))) {
        if (s[i] == '#') acc += 1;
        else if (s[i] == 'b') acc -= 1;
        else break;
        ++i;
    }
    int32_t octave;
    if (!parse_i32(trim(s.substr(i)), octave)) return false;
    out = (octave + 1) * 12 + step + acc;
    return true;
}

struct Ops {
    std::map<std::string, std::string> m;
    const std::string *get(const char *k) const {
        auto it = m.find(k);
        return it == m.end() ? nullptr : &it->second;
    }
    bool f32(const char *k, float &v) const {
        const std::string *s = get(k);
        return s && parse_f32(trim(*s), v);
    }
    bool i32(const char *k, int32_t &v) const {
        const std::string *s = get(k);
        if (!s) return false;
        const std::string t = trim(*s);
        if (parse_i32(t, v)) return true;
        float f;
        if (parse_f32(t, f)) {
            v = f32_to_i32(f);
            return true;
        }
        return false;
    }
    bool key(const char *k, int32_t &v) const {
        const std::string *s = get(k);
        if (!s) return false;
        const std::string t = trim(*s);
        return parse_i32(t, v) || parse_note_name(t, v);
    }
    // Aliases collapse to one name at the level that wrote them (src/sfz.rs:51-68).
    void canonicalise() {
        auto it = m.find("key");
        if (it != m.end()) {
            const std::string k = it->second;
            m.erase(it);
            for (const char *alias : {"lokey", "hikey", "pitch_keycenter"}) m.emplace(alias, k);
        }
        for (auto [from, to] : {std::pair<const char *, const char *>{"loopmode", "loop_mode"},
                                {"loopstart", "loop_start"}, {"loopend", "loop_end"},
                                {"pitch", "tune"}}) {
            auto f = m.find(from);
            if (f != m.end()) {
                const std::string v = f->second;
                m.erase(f);
                m.emplace(to, v);
            }
        }
    }
    Ops merged(const Ops &o) const {
        Ops r = *this;
        for (const auto &kv : o.m) r.m[kv.first] = kv.second;
        return r;
    }
};

// Where a value ends: before the last whitespace-separated token that itself
// holds an '=' (or before a '<'), so file names with spaces survive
// (src/sfz.rs:226-252).
bool is_ascii_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\f' || c == '\r'; }

size_t find_value_end(const std::string &s) {
    size_t i = 0;
    bool have_ws = false;
    size_t last_ws = 0;
    while (i < s.size()) {
        if (s[i] == '=') return have_ws ? last_ws : i;
        if (s[i] == '<') return have_ws ? last_ws : i;
        if (is_ascii_ws(s[i])) {
            size_t j = i;
            while (j < s.size() && is_ascii_ws(s[j])) ++j;
            have_ws = true;
            last_ws = i;
            i = j;
            continue;
        }
        ++i;
    }
    return s.size();
}

bool valid_utf8(const std::string &s) {
    size_t i = 0;
    while (i < s.size()) {
        const unsigned char c = (unsigned char)s[i];
        size_t n = c < 0x80 ? 1 : (c >> 5) == 6 ? 2 : (c >> 4) == 14 ? 3 : (c >> 3) == 30 ? 4 : 0;
        if (!n || i + n > s.size()) return false;
        for (size_t k = 1; k < n; ++k)
            if (((unsigned char)s[i + k] >> 6) != 2) return false;
        i += n;
    }
    return true;
}

bool exists(const std::string &p) {
    struct stat st {};
    return stat(p.c_str(), &st) == 0;
}

std::string join(const std::string &a, const std::string &b) {
    if (!b.empty() && b[0] == '/') return b;
    if (a.empty()) return b;
    if (b.empty()) return a.back() == '/' ? a : a + "/";
    return a.back() == '/' ? a + b : a + "/" + b;
}

std::string replace_backslashes(std::string s) {
    std::replace(s.begin(), s.end(), '\\', '/');
    return s;
}

struct Parser {
    std::string root;
    uint32_t depth = 0;

    // src/sfz.rs:124-146: root / default_path / rel, then a case-insensitive
    // match of the file name in its directory.
    std::string resolve(const std::string &default_path, const std::string &rel) const {
        const std::string cand = join(root, join(default_path, replace_backslashes(rel)));
        if (exists(cand)) return cand;
        const size_t slash = cand.rfind('/');
        const std::string parent = slash == std::string::npos ? root : (slash == 0 ? "/" : cand.substr(0, slash));
        const std::string name = lower(slash == std::string::npos ? cand : cand.substr(slash + 1));
        if (DIR *d = opendir(parent.c_str())) {
            while (dirent *e = readdir(d)) {
                if (lower(e->d_name) == name) {
                    const std::string hit = join(parent, e->d_name);
                    closedir(d);
                    return hit;
                }
            }
            closedir(d);
        }
        return cand;
    }

    void parse_file(const std::string &path, std::vector<std::pair<std::string, Ops>> &out) {
        if (depth > 16) throw std::runtime_error("#include nesting is too deep at " + path);
        std::ifstream f(path, std::ios::binary);
        if (!f) throw std::runtime_error("reading " + path + ": " + std::strerror(errno));
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string text = ss.str();
        if (!valid_utf8(text)) throw std::runtime_error("reading " + path + ": stream did not contain valid UTF-8");
        std::string header = "global";
        Ops cur;
        bool started = false;
        size_t at = 0;
        while (at <= text.size()) {
            size_t nl = text.find('\n', at);
            if (nl == std::string::npos) {
                if (at == text.size()) break;
                nl = text.size();
            }
            std::string raw = text.substr(at, nl - at);
            at = nl + 1;
            if (!raw.empty() && raw.back() == '\r') raw.pop_back();
            const size_t cm = raw.find("//");
            const std::string line = trim(cm == std::string::npos ? raw : raw.substr(0, cm));
            if (line.empty()) continue;
            if (line.compare(0, 8, "#include") == 0) {
                std::string inc = trim(line.substr(8));
                size_t a = 0, e = inc.size();
                while (a < e && inc[a] == '"') ++a;
                while (e > a && inc[e - 1] == '"') --e;
                inc = trim(inc.substr(a, e - a));
                const std::string p = resolve("", inc);
                if (started) {
                    out.emplace_back(header, std::move(cur));
                    cur = Ops();
                    started = false;
                }
                ++depth;
                try {
                    parse_file(p, out);
                } catch (const std::exception &) {
                    // an unreadable include is skipped, as the reference warns and goes on
                }
                --depth;
                continue;
            }
            if (line.compare(0, 7, "#define") == 0) continue;
            std::string rest = line;
            while (!rest.empty()) {
                size_t a = 0;
                while (a < rest.size() && is_ws(rest[a])) ++a;
                rest = rest.substr(a);
                if (rest.empty()) break;
                if (rest[0] == '<') {
                    const size_t end = rest.find('>', 1);
                    if (end == std::string::npos) break;
                    if (started) {
                        out.emplace_back(header, std::move(cur));
                    }
                    header = lower(trim(rest.substr(1, end - 1)));
                    cur = Ops();
                    started = true;
                    rest = rest.substr(end + 1);
                    continue;
                }
                const size_t eq = rest.find('=');
                if (eq == std::string::npos) break;
                const std::string key = lower(trim(rest.substr(0, eq)));
                const std::string after = rest.substr(eq + 1);
                const size_t ve = find_value_end(after);
                const std::string value = trim(after.substr(0, ve));
                rest = after.substr(ve);
                if (key.empty()) continue;
                cur.m[key] = value;
                started = true;
            }
        }
        if (started) out.emplace_back(header, std::move(cur));
    }
};

// region_from_opcodes, src/sfz.rs:547-676.
kst_region region_from(const Ops &ops, uint32_t sample, const kst_sample_info &info) {
    kst_region r{};
    r.sample = sample;
    r.key_lo = 0;
    r.key_hi = 127;
    r.vel_lo = 0;
    r.vel_hi = 127;
    r.root_key_override = -1;
    r.fixed_key = r.fixed_vel = -1;
    r.scale_tuning = 100;
    r.amp_veltrack = 100.0f;
    r.loop_mode = KST_LOOP_NONE;
    r.attack = 0.001f;
    r.decay = 0.001f;
    r.sustain = 1.0f;
    r.release = 0.001f;
    r.filter_fc_cents = 13500.0f;
    auto clampi = [](int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); };
    auto clampf = [](float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); };
    int32_t k;
    float f;
    if (ops.key("lokey", k)) r.key_lo = (uint8_t)clampi(k, 0, 127);
    if (ops.key("hikey", k)) r.key_hi = (uint8_t)clampi(k, 0, 127);
    if (ops.key("pitch_keycenter", k)) r.root_key_override = (int16_t)clampi(k, 0, 127);
    if (r.root_key_override < 0) r.root_key_override = (int16_t)info.root_key;
    if (ops.i32("lovel", k)) r.vel_lo = (uint8_t)clampi(k, 0, 127);
    if (ops.i32("hivel", k)) r.vel_hi = (uint8_t)clampi(k, 0, 127);
    float tune = 0.0f;
    ops.f32("tune", tune);
    {
        const float t = clampf(std::round(tune), -32768.0f, 32767.0f);
        r.fine_tune = t != t ? 0 : (int16_t)t;
    }
    if (ops.i32("transpose", k)) r.coarse_tune = (int16_t)clampi(k, -127, 127);
    if (ops.f32("pitch_keytrack", f)) {
        const float t = clampf(std::round(f), -1200.0f, 1200.0f);
        r.scale_tuning = t != t ? 0 : (int16_t)t;
    }
    if (ops.f32("volume", f)) r.attenuation_cb = -f * 10.0f;
    if (ops.f32("pan", f)) r.pan = clampf(f / 100.0f, -1.0f, 1.0f);
    if (ops.f32("amp_veltrack", f)) r.amp_veltrack = clampf(f, -100.0f, 100.0f);
    const std::string *lm = ops.get("loop_mode");
    const std::string mode = lm ? *lm : "";
    if (mode == "loop_continuous") r.loop_mode = KST_LOOP_CONTINUOUS;
    else if (mode == "loop_sustain") r.loop_mode = KST_LOOP_UNTIL_RELEASE;
    else if (mode == "one_shot" || mode == "no_loop") r.loop_mode = KST_LOOP_NONE;
    else r.loop_mode = info.loop_end > info.loop_start + 1 ? KST_LOOP_CONTINUOUS : KST_LOOP_NONE;
    if (ops.i32("offset", k)) r.addr_start = std::max(k, 0);
    // Absolute positions in source frames become offsets from the sample's
    // own points, taken back to source frames first.
    auto to_source = [&](uint32_t resampled) -> int32_t {
        if (info.resample_ratio > 0.0f) return f32_to_i32(std::round((float)resampled / info.resample_ratio));
        return (int32_t)resampled;
    };
    if (ops.i32("end", k) && k > 0) r.addr_end = (int32_t)((uint32_t)k - (uint32_t)to_source(info.len));
    if (ops.i32("loop_start", k)) r.addr_loop_start = (int32_t)((uint32_t)k - (uint32_t)to_source(info.loop_start));
    if (ops.i32("loop_end", k)) r.addr_loop_end = (int32_t)((uint32_t)k - (uint32_t)to_source(info.loop_end));
    auto env = [&](const char *name, float dflt) {
        float v = dflt;
        ops.f32(name, v);
        return std::fmax(v, 0.0f);
    };
    r.delay = env("ampeg_delay", 0.0f);
    r.attack = env("ampeg_attack", 0.001f);
    r.hold = env("ampeg_hold", 0.0f);
    r.decay = env("ampeg_decay", 100.0f);
    {
        float s = 100.0f;
        ops.f32("ampeg_sustain", s);
        r.sustain = clampf(s / 100.0f, 0.0f, 1.0f);
    }
    r.release = env("ampeg_release", 0.05f);
    if (ops.f32("cutoff", f) && f > 0.0f) r.filter_fc_cents = 1200.0f * log2f(f / 8.176f);
    if (ops.f32("resonance", f)) r.filter_q_cb = f * 10.0f;
    if (ops.f32("fil_veltrack", f)) r.filter_veltrack_cents = clampf(f, -9600.0f, 9600.0f);
    if (const std::string *ft = ops.get("fil_type")) {
        if (ft->compare(0, 3, "lpf") != 0) {  // only lowpasses exist here
            r.filter_fc_cents = 13500.0f;
            r.filter_veltrack_cents = 0.0f;
        }
    }
    int32_t group = 0;
    ops.i32("group", group);
    group = clampi(group, 0, 255);
    int32_t off;
    if (ops.i32("off_by", off) && clampi(off, 0, 255) == group && group > 0) r.exclusive_class = (uint32_t)group;
    return r;
}

}  // namespace

// src/sfz.rs:254-425
Bank *load_sfz(const BankSettings &set, const std::string &path, bool resample_pool) {
    const size_t slash = path.rfind('/');
    Parser parser;
    parser.root = slash == std::string::npos ? "." : (slash == 0 ? "/" : path.substr(0, slash));
    std::vector<std::pair<std::string, Ops>> sections;
    parser.parse_file(path, sections);
    for (auto &s : sections) s.second.canonicalise();

    std::string default_path;
    Ops global, master, group;
    std::vector<std::pair<std::string, Ops>> region_sets;
    for (const auto &[h, ops] : sections) {
        if (h == "control") {
            if (const std::string *dp = ops.get("default_path")) default_path = replace_backslashes(*dp);
        } else if (h == "global") {
            global = ops;
            master = Ops();
            group = Ops();
        } else if (h == "master") {
            master = ops;
            group = Ops();
        } else if (h == "group") {
            group = ops;
        } else if (h == "region") {
            region_sets.emplace_back(default_path, global.merged(master).merged(group).merged(ops));
        }
    }
    if (region_sets.empty()) throw std::runtime_error(path + ": no <region> sections");

    const uint32_t pool_rate = resample_pool ? set.sample_rate : 0;
    std::vector<int16_t> pool;
    std::vector<kst_sample_info> samples;
    std::map<std::pair<std::string, int>, uint32_t> cache;
    std::vector<kst_region> regions;
    std::vector<uint32_t> ids;
    for (const auto &[dp, ops] : region_sets) {
        const std::string *rel = ops.get("sample");
        if (!rel) continue;
        int32_t end;
        if (ops.i32("end", end) && end == -1) continue;  // a disabled region
        const std::string spath = parser.resolve(dp, *rel);
        std::vector<uint32_t> chans;
        auto c0 = cache.find({spath, 0});
        if (c0 != cache.end()) {
            chans.push_back(c0->second);
            auto c1 = cache.find({spath, 1});
            if (c1 != cache.end()) chans.push_back(c1->second);
        } else {
            Wav w;
            try {
                w = read_wav(spath);
            } catch (const std::exception &) {
                continue;  // an unreadable sample drops its region, as the reference warns
            }
            const size_t nch = std::min<size_t>(std::max<size_t>(w.channels, 1), 2);
            for (size_t ch = 0; ch < nch; ++ch) {
                std::vector<int16_t> raw = w.channel_i16(ch);
                const uint32_t src_rate = std::max<uint32_t>(w.sample_rate, 1);
                const double ratio = pool_rate != 0 && pool_rate != src_rate
                                         ? (double)pool_rate / (double)src_rate : 1.0;
                std::vector<int16_t> data = std::fabs(ratio - 1.0) > 1e-12
                                                ? resample_i16(raw.data(), raw.size(), ratio) : raw;
                const uint32_t len = (uint32_t)data.size();
                uint32_t ls = w.has_loop ? w.loop_start : 0;
                uint32_t le = w.has_loop ? w.loop_end : (len > 0 ? len - 1 : 0);
                ls = (uint32_t)sat_u64(std::round((double)ls * ratio));
                le = (uint32_t)sat_u64(std::round((double)le * ratio));
                kst_sample_info si{};
                si.start = (uint32_t)pool.size();
                si.len = len;
                si.loop_start = std::min(ls, len > 0 ? len - 1 : 0);
                si.loop_end = std::min(le, len);
                si.rate = pool_rate ? pool_rate : src_rate;
                si.root_key = w.root_key >= 0 ? (uint32_t)w.root_key : 60u;
                si.correction_cents = w.fine_tune_cents;
                si.resample_ratio = (float)ratio;
                pool.insert(pool.end(), data.begin(), data.end());
                pool.insert(pool.end(), POOL_GUARD, 0);
                cache[{spath, (int)ch}] = (uint32_t)samples.size();
                chans.push_back((uint32_t)samples.size());
                samples.push_back(si);
            }
        }
        for (size_t ch = 0; ch < chans.size(); ++ch) {
            kst_region r = region_from(ops, chans[ch], samples[chans[ch]]);
            if (chans.size() == 2) r.pan = ch == 0 ? -1.0f : 1.0f;  // a stereo file's halves
            ids.push_back((uint32_t)regions.size());
            regions.push_back(r);
        }
    }
    if (regions.empty()) throw std::runtime_error(path + ": every region was unusable");
    const kst_preset preset{0, 0, 0, (uint32_t)ids.size()};
    return new Bank(set, pool.data(), pool.size(), pool_rate, samples.data(), samples.size(),
                    regions.data(), regions.size(), &preset, 1, ids.data(), ids.size());
}

}  // namespace kst
