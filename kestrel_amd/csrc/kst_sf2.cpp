// kst_sf2.cpp -- the SoundFont 2 loader and the pool's sample-rate converter
// (src/sf2.rs, src/resample.rs).
//
// The RIFF walk reads chunk headers only; the preset data hunk is read whole
// (it is small) and flattened preset zone by instrument zone into resolved
// regions; then only the samples some region uses are read from `smpl`,
// converted to the pool rate when they are at another (a polyphase
// windowed-sinc bank, table in f64 rounded to f32, accumulation in f32), and
// laid out with guard frames. The result goes through the same Bank
// constructor kst_bank_create uses.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstring>
#include <set>
#include <stdexcept>
#include <string>

#include "kst_host.h"

namespace kst {
namespace {

// Generator operators, SoundFont 2.01 section 8.1.2.
enum : uint16_t {
    G_START_ADDRS = 0, G_END_ADDRS = 1, G_STARTLOOP_ADDRS = 2, G_ENDLOOP_ADDRS = 3,
    G_START_ADDRS_COARSE = 4, G_INITIAL_FILTER_FC = 8, G_INITIAL_FILTER_Q = 9,
    G_END_ADDRS_COARSE = 12, G_PAN = 17, G_DELAY_VOL_ENV = 33, G_ATTACK_VOL_ENV = 34,
    G_HOLD_VOL_ENV = 35, G_DECAY_VOL_ENV = 36, G_SUSTAIN_VOL_ENV = 37, G_RELEASE_VOL_ENV = 38,
    G_KEYNUM_TO_HOLD = 39, G_KEYNUM_TO_DECAY = 40, G_INSTRUMENT = 41, G_KEY_RANGE = 43,
    G_VEL_RANGE = 44, G_STARTLOOP_ADDRS_COARSE = 45, G_KEYNUM = 46, G_VELOCITY = 47,
    G_INITIAL_ATTENUATION = 48, G_ENDLOOP_ADDRS_COARSE = 50, G_COARSE_TUNE = 51,
    G_FINE_TUNE = 52, G_SAMPLE_ID = 53, G_SAMPLE_MODES = 54, G_SCALE_TUNING = 56,
    G_EXCLUSIVE_CLASS = 57, G_OVERRIDING_ROOT_KEY = 58,
};
constexpr size_t GEN_COUNT = 60;
constexpr uint32_t POOL_GUARD = 8;  // src/sf2.rs:583
typedef std::array<int16_t, GEN_COUNT> Gens;
typedef std::array<bool, GEN_COUNT> GenSet;

// Instrument-level only: a preset zone may not offset these (src/sf2.rs:56-77).
bool absolute_only(size_t op) {
    switch (op) {
    case G_START_ADDRS: case G_END_ADDRS: case G_STARTLOOP_ADDRS: case G_ENDLOOP_ADDRS:
    case G_START_ADDRS_COARSE: case G_END_ADDRS_COARSE: case G_STARTLOOP_ADDRS_COARSE:
    case G_ENDLOOP_ADDRS_COARSE: case G_SAMPLE_ID: case G_SAMPLE_MODES: case G_EXCLUSIVE_CLASS:
    case G_OVERRIDING_ROOT_KEY: case G_KEYNUM: case G_VELOCITY: case G_INSTRUMENT:
    case G_KEY_RANGE: case G_VEL_RANGE:
        return true;
    default:
        return false;
    }
}

Gens default_gens() {  // the spec's defaults, src/sf2.rs:79-101
    Gens g{};
    g[G_INITIAL_FILTER_FC] = 13500;
    for (int op : {21, 23, 25, 26, 27, 28, 30}) g[op] = -12000;
    for (int op : {G_DELAY_VOL_ENV, G_ATTACK_VOL_ENV, G_HOLD_VOL_ENV, G_DECAY_VOL_ENV,
                   G_RELEASE_VOL_ENV})
        g[op] = -12000;
    g[G_KEY_RANGE] = (int16_t)0x7F00;
    g[G_VEL_RANGE] = (int16_t)0x7F00;
    g[G_KEYNUM] = -1;
    g[G_VELOCITY] = -1;
    g[G_SCALE_TUNING] = 100;
    g[G_OVERRIDING_ROOT_KEY] = -1;
    return g;
}

float timecents_to_secs(float tc) { return tc <= -12000.0f ? 0.0f : exp2f(tc / 1200.0f); }
float cb_to_gain(float cb) { return powf(10.0f, -cb / 200.0f); }

std::pair<uint8_t, uint8_t> range_of(const Gens &g, bool set, uint16_t op) {
    if (!set) return {0, 127};
    const uint16_t raw = (uint16_t)g[op];
    const uint8_t lo = raw & 0xFF, hi = raw >> 8;
    return lo > hi ? std::make_pair(hi, lo) : std::make_pair(lo, hi);
}

int16_t sat_add16(int16_t a, int16_t b) {
    const int v = (int)a + (int)b;
    return (int16_t)std::min(32767, std::max(-32768, v));
}

// One flattened zone: the generators of src/sf2.rs:532-581 turned into a region.
kst_region make_region(const Gens &g, uint32_t sample, std::pair<uint8_t, uint8_t> key,
                       std::pair<uint8_t, uint8_t> vel) {
    kst_region r{};
    r.sample = sample;
    r.key_lo = key.first;
    r.key_hi = key.second;
    r.vel_lo = vel.first;
    r.vel_hi = vel.second;
    r.root_key_override = g[G_OVERRIDING_ROOT_KEY];
    r.fixed_key = g[G_KEYNUM];
    r.fixed_vel = g[G_VELOCITY];
    r.coarse_tune = g[G_COARSE_TUNE];
    r.fine_tune = g[G_FINE_TUNE];
    r.scale_tuning = g[G_SCALE_TUNING];
    r.attenuation_cb = (float)g[G_INITIAL_ATTENUATION];
    r.amp_veltrack = 100.0f;  // SF2's velocity-to-attenuation modulator is full tracking
    const float pan = (float)g[G_PAN] / 500.0f;
    r.pan = pan < -1.0f ? -1.0f : (pan > 1.0f ? 1.0f : pan);
    const int mode = g[G_SAMPLE_MODES] & 3;
    r.loop_mode = mode == 1 ? KST_LOOP_CONTINUOUS : (mode == 3 ? KST_LOOP_UNTIL_RELEASE : KST_LOOP_NONE);
    r.addr_start = (int32_t)g[G_START_ADDRS] + (int32_t)g[G_START_ADDRS_COARSE] * 32768;
    r.addr_end = (int32_t)g[G_END_ADDRS] + (int32_t)g[G_END_ADDRS_COARSE] * 32768;
    r.addr_loop_start = (int32_t)g[G_STARTLOOP_ADDRS] + (int32_t)g[G_STARTLOOP_ADDRS_COARSE] * 32768;
    r.addr_loop_end = (int32_t)g[G_ENDLOOP_ADDRS] + (int32_t)g[G_ENDLOOP_ADDRS_COARSE] * 32768;
    r.exclusive_class = (uint32_t)std::min<int>(255, std::max<int>(0, g[G_EXCLUSIVE_CLASS]));
    r.delay = timecents_to_secs((float)g[G_DELAY_VOL_ENV]);
    r.attack = timecents_to_secs((float)g[G_ATTACK_VOL_ENV]);
    r.hold = timecents_to_secs((float)g[G_HOLD_VOL_ENV]);
    r.decay = timecents_to_secs((float)g[G_DECAY_VOL_ENV]);
    r.sustain = cb_to_gain(std::fmax((float)g[G_SUSTAIN_VOL_ENV], 0.0f));
    r.release = timecents_to_secs((float)g[G_RELEASE_VOL_ENV]);
    r.keynum_to_hold = g[G_KEYNUM_TO_HOLD];
    r.keynum_to_decay = g[G_KEYNUM_TO_DECAY];
    r.filter_fc_cents = (float)g[G_INITIAL_FILTER_FC];
    r.filter_q_cb = (float)g[G_INITIAL_FILTER_Q];
    r.filter_veltrack_cents = 0.0f;
    return r;
}

struct Chunk {
    char id[4];
    uint64_t offset, size;
};

class File {
public:
    explicit File(const std::string &path) : path_(path) {
        fd_ = open(path.c_str(), O_RDONLY | O_CLOEXEC);
        if (fd_ < 0) throw std::runtime_error("opening " + path + ": " + std::strerror(errno));
        struct stat st {};
        if (fstat(fd_, &st) != 0) throw std::runtime_error("opening " + path + ": " + std::strerror(errno));
        len = (uint64_t)st.st_size;
    }
    ~File() { close(fd_); }
    bool read(void *dst, size_t n, uint64_t off) const {
        size_t got = 0;
        while (got < n) {
            const ssize_t r = pread(fd_, (char *)dst + got, n - got, (off_t)(off + got));
            if (r < 0 && errno == EINTR) continue;
            if (r <= 0) return false;
            got += (size_t)r;
        }
        return true;
    }
    std::vector<uint8_t> chunk(const Chunk &c) const {
        std::vector<uint8_t> b((size_t)c.size);
        if (!read(b.data(), b.size(), c.offset))
            throw std::runtime_error(path_ + ": chunk runs past the end of the file");
        return b;
    }
    uint64_t len = 0;

private:
    std::string path_;
    int fd_ = -1;
};

uint16_t u16(const uint8_t *p) { return (uint16_t)(p[0] | (p[1] << 8)); }
uint32_t u32(const uint8_t *p) { return (uint32_t)p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); }

// The RIFF tree, headers only (src/sf2.rs:153-207): the sub-chunks of the
// sdta and pdta lists.
void scan_chunks(const File &f, std::vector<Chunk> &sdta, std::vector<Chunk> &pdta) {
    uint8_t h[12];
    if (!f.read(h, 8, 0) || std::memcmp(h, "RIFF", 4) != 0) throw std::runtime_error("not a RIFF file");
    if (!f.read(h + 8, 4, 8)) throw std::runtime_error("not a RIFF file");
    if (std::memcmp(h + 8, "sfbk", 4) != 0)
        throw std::runtime_error("RIFF form is " + std::string((char *)h + 8, 4) + ", expected sfbk");
    uint64_t pos = 12;
    while (pos + 8 <= f.len) {
        uint8_t c[12];
        if (!f.read(c, 8, pos)) break;
        const uint64_t size = std::min<uint64_t>(u32(c + 4), f.len - pos - 8);
        if (std::memcmp(c, "LIST", 4) == 0) {
            if (!f.read(c + 8, 4, pos + 8)) throw std::runtime_error("truncated LIST chunk");
            std::vector<Chunk> *target = std::memcmp(c + 8, "sdta", 4) == 0   ? &sdta
                                         : std::memcmp(c + 8, "pdta", 4) == 0 ? &pdta
                                                                              : nullptr;
            if (target) {
                const uint64_t list_end = pos + 8 + size;
                uint64_t inner = pos + 12;
                while (inner + 8 <= list_end) {
                    uint8_t s[8];
                    if (!f.read(s, 8, inner)) break;
                    const uint64_t csz = std::min<uint64_t>(u32(s + 4), list_end - inner - 8);
                    Chunk k;
                    std::memcpy(k.id, s, 4);
                    k.offset = inner + 8;
                    k.size = csz;
                    target->push_back(k);
                    inner += 8 + csz + (csz & 1);
                }
            }
        }
        pos += 8 + size + (size & 1);
    }
}

const Chunk *find(const std::vector<Chunk> &v, const char *id) {
    for (const auto &c : v)
        if (std::memcmp(c.id, id, 4) == 0) return &c;
    return nullptr;
}

struct Shdr {
    uint32_t start, end, start_loop, end_loop, rate;
    uint8_t pitch;
    int8_t correction;
    uint16_t type;
};

}  // namespace

// ---------------------------------------------------------------------------
// src/resample.rs: 32 taps x 512 phases, Blackman-windowed sinc with the
// cutoff lowered by the ratio when downsampling, every phase normalised to
// unity gain at DC.
// ---------------------------------------------------------------------------
std::vector<int16_t> resample_i16(const int16_t *src, size_t n, double ratio) {
    if (std::fabs(ratio - 1.0) < 1e-12) return std::vector<int16_t>(src, src + n);
    constexpr int TAPS = 32, PHASES = 512;
    const double PI = 3.14159265358979323846;
    const double cutoff = (ratio < 1.0 ? ratio : 1.0) * 0.95;
    std::vector<float> table((size_t)PHASES * TAPS);
    const double half = TAPS / 2.0;
    for (int p = 0; p < PHASES; ++p) {
        const double frac = (double)p / PHASES;
        double sum = 0.0;
        for (int k = 0; k < TAPS; ++k) {
            const double x = (double)k - half + 1.0 - frac;
            const double xs = x * cutoff;
            double s;
            if (std::fabs(xs) < 1e-9) {
                s = 1.0;
            } else {
                const double px = PI * xs;
                s = std::sin(px) / px;
            }
            s *= cutoff;
            double t = ((double)k - frac + 1.0) / TAPS;
            t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
            const double w = 0.42 - 0.5 * std::cos(2.0 * PI * t) + 0.08 * std::cos(4.0 * PI * t);
            const double v = s * w;
            table[(size_t)p * TAPS + k] = (float)v;
            sum += v;
        }
        if (std::fabs(sum) > 1e-9) {
            const float inv = (float)(1.0 / sum);
            for (int k = 0; k < TAPS; ++k) table[(size_t)p * TAPS + k] *= inv;
        }
    }
    std::vector<int16_t> out;
    if (n == 0) return out;
    const size_t out_len = std::max<size_t>((size_t)std::round((double)n * ratio), 1);
    out.resize(out_len);
    const int64_t N = (int64_t)n, H = TAPS / 2;
    const double inv_ratio = 1.0 / ratio;
    for (size_t j = 0; j < out_len; ++j) {
        const double t = (double)j * inv_ratio;
        const int64_t i = (int64_t)std::floor(t);
        const double frac = t - (double)i;
        const size_t p = std::min<size_t>((size_t)(frac * PHASES), PHASES - 1);
        const float *c = &table[p * TAPS];
        const int64_t base = i - H + 1;
        float acc = 0.0f;
        if (base >= 0 && base + TAPS <= N) {
            for (int k = 0; k < TAPS; ++k) acc += c[k] * (float)src[base + k];
        } else {
            for (int k = 0; k < TAPS; ++k) {
                const int64_t idx = std::min<int64_t>(std::max<int64_t>(base + k, 0), N - 1);
                acc += c[k] * (float)src[idx];
            }
        }
        acc = acc < -32768.0f ? -32768.0f : (acc > 32767.0f ? 32767.0f : acc);
        out[j] = (int16_t)acc;
    }
    return out;
}

// src/sf2.rs:221-506 and 585-703.
Bank *load_sf2(const BankSettings &set, const std::string &path, bool resample_pool,
               uint64_t budget) {
    File f(path);
    std::vector<Chunk> sdta, pdta;
    try {
        scan_chunks(f, sdta, pdta);
    } catch (const std::exception &e) {
        throw std::runtime_error(path + ": scanning RIFF structure: " + e.what());
    }
    const Chunk *smpl = find(sdta, "smpl");
    if (!smpl) throw std::runtime_error("soundfont has no smpl chunk");
    const uint64_t smpl_offset = smpl->offset;
    const uint32_t smpl_frames = (uint32_t)(smpl->size / 2);
    auto need = [&](const char *id) {
        const Chunk *c = find(pdta, id);
        if (!c) throw std::runtime_error(std::string("pdta is missing ") + id);
        return f.chunk(*c);
    };
    const auto phdr = need("phdr"), pbag = need("pbag"), pgen = need("pgen"), inst = need("inst"),
               ibag = need("ibag"), igen = need("igen"), shdr_raw = need("shdr");
    // Whole records only, as chunks_exact reads them.
    const size_t n_phdr = phdr.size() / 38, n_pbag = pbag.size() / 4, n_pgen = pgen.size() / 4,
                 n_inst = inst.size() / 22, n_ibag = ibag.size() / 4, n_igen = igen.size() / 4,
                 n_shdr = shdr_raw.size() / 46;
    auto phdr_bank = [&](size_t i) { return u16(&phdr[i * 38 + 22]); };
    auto phdr_prog = [&](size_t i) { return u16(&phdr[i * 38 + 20]); };
    auto phdr_bag = [&](size_t i) { return (size_t)u16(&phdr[i * 38 + 24]); };
    auto pbag_gen = [&](size_t i) { return (size_t)u16(&pbag[i * 4]); };
    auto ibag_gen = [&](size_t i) { return (size_t)u16(&ibag[i * 4]); };
    auto inst_bag = [&](size_t i) { return (size_t)u16(&inst[i * 22 + 20]); };
    std::vector<Shdr> shdrs(n_shdr);
    for (size_t i = 0; i < n_shdr; ++i) {
        const uint8_t *c = &shdr_raw[i * 46];
        shdrs[i] = Shdr{u32(c + 20), u32(c + 24), u32(c + 28), u32(c + 32), u32(c + 36),
                        c[40], (int8_t)c[41], u16(c + 44)};
    }
    if (n_phdr < 2) throw std::runtime_error("soundfont has no presets");

    // ---- flatten every preset into regions ----
    std::vector<kst_region> regions;
    std::vector<kst_preset> presets;
    std::vector<uint32_t> preset_regions;
    std::set<uint32_t> used;
    for (size_t pi = 0; pi + 1 < n_phdr; ++pi) {
        const size_t bag_lo = phdr_bag(pi), bag_hi = phdr_bag(pi + 1);
        if (bag_lo >= bag_hi || bag_hi > n_pbag) continue;
        Gens pglobal{};
        GenSet pglobal_set{};
        bool first_zone = true;
        const size_t first_region = preset_regions.size();
        for (size_t bi = bag_lo; bi < bag_hi; ++bi) {
            const size_t g_lo = pbag_gen(bi), g_hi = bi + 1 < n_pbag ? pbag_gen(bi + 1) : n_pgen;
            if (g_lo > g_hi || g_hi > n_pgen) continue;
            Gens zone{};
            GenSet zone_set{};
            int32_t instrument = -1;
            for (size_t gi = g_lo; gi < g_hi; ++gi) {
                const uint16_t op = u16(&pgen[gi * 4]);
                const int16_t amt = (int16_t)u16(&pgen[gi * 4 + 2]);
                if (op == G_INSTRUMENT) {
                    instrument = (uint16_t)amt;
                } else if (op < GEN_COUNT) {
                    zone[op] = amt;
                    zone_set[op] = true;
                }
            }
            if (instrument < 0) {  // the global zone, if first
                if (first_zone) {
                    pglobal = zone;
                    pglobal_set = zone_set;
                }
                first_zone = false;
                continue;
            }
            first_zone = false;
            Gens pg = pglobal;
            GenSet ps = pglobal_set;
            for (size_t i = 0; i < GEN_COUNT; ++i)
                if (zone_set[i]) {
                    pg[i] = zone[i];
                    ps[i] = true;
                }
            const auto p_key = range_of(pg, ps[G_KEY_RANGE], G_KEY_RANGE);
            const auto p_vel = range_of(pg, ps[G_VEL_RANGE], G_VEL_RANGE);
            if ((size_t)instrument >= n_inst) continue;
            const size_t ib_lo = inst_bag((size_t)instrument);
            const size_t ib_hi = (size_t)instrument + 1 < n_inst ? inst_bag((size_t)instrument + 1) : n_ibag;
            if (ib_lo >= ib_hi || ib_hi > n_ibag) continue;
            Gens iglobal = default_gens();
            bool first_izone = true;
            for (size_t ibi = ib_lo; ibi < ib_hi; ++ibi) {
                const size_t ig_lo = ibag_gen(ibi), ig_hi = ibi + 1 < n_ibag ? ibag_gen(ibi + 1) : n_igen;
                if (ig_lo > ig_hi || ig_hi > n_igen) continue;
                Gens ig = iglobal;
                int32_t sid = -1;
                for (size_t gi = ig_lo; gi < ig_hi; ++gi) {
                    const uint16_t op = u16(&igen[gi * 4]);
                    const int16_t amt = (int16_t)u16(&igen[gi * 4 + 2]);
                    if (op == G_SAMPLE_ID)
                        sid = (uint16_t)amt;
                    else if (op < GEN_COUNT)
                        ig[op] = amt;
                }
                if (sid < 0) {
                    if (first_izone) iglobal = ig;
                    first_izone = false;
                    continue;
                }
                first_izone = false;
                if ((size_t)sid >= n_shdr) continue;
                const Shdr &sh = shdrs[(size_t)sid];
                if (sh.type & 0x8000) continue;  // ROM sample
                if (sh.end <= sh.start || sh.end > smpl_frames) continue;
                const auto i_key = range_of(ig, true, G_KEY_RANGE);
                const auto i_vel = range_of(ig, true, G_VEL_RANGE);
                const uint8_t klo = std::max(p_key.first, i_key.first), khi = std::min(p_key.second, i_key.second);
                const uint8_t vlo = std::max(p_vel.first, i_vel.first), vhi = std::min(p_vel.second, i_vel.second);
                if (klo > khi || vlo > vhi) continue;
                Gens g = ig;  // preset generators offset the instrument's
                for (size_t op = 0; op < GEN_COUNT; ++op)
                    if (ps[op] && !absolute_only(op)) g[op] = sat_add16(g[op], pg[op]);
                used.insert((uint32_t)sid);
                preset_regions.push_back((uint32_t)regions.size());
                regions.push_back(make_region(g, (uint32_t)sid, {klo, khi}, {vlo, vhi}));
            }
        }
        if (preset_regions.size() > first_region)
            presets.push_back(kst_preset{phdr_bank(pi), phdr_prog(pi), (uint32_t)first_region,
                                         (uint32_t)(preset_regions.size() - first_region)});
    }
    if (presets.empty()) throw std::runtime_error(path + ": no usable presets");

    // ---- the pool: the used samples, at one rate when resampling ----
    uint32_t pool_rate = resample_pool ? set.sample_rate : 0;
    uint64_t raw_frames = 0;
    for (uint32_t sid : used) raw_frames += (uint64_t)(shdrs[sid].end - shdrs[sid].start) + POOL_GUARD;
    if (pool_rate) {
        for (;;) {  // halve the rate until the estimate fits the budget
            uint64_t est = 0;
            for (uint32_t sid : used) {
                const Shdr &s = shdrs[sid];
                const double nf = (double)(s.end - s.start) * (double)pool_rate /
                                  (double)std::max<uint32_t>(s.rate, 1);
                est += sat_u64(nf) + POOL_GUARD;
            }
            if (est * 2 <= budget || pool_rate <= 8000) break;
            pool_rate /= 2;
        }
    }
    std::vector<int16_t> pool;
    pool.reserve((size_t)raw_frames + 64);
    std::vector<kst_sample_info> samples;
    std::vector<uint32_t> remap(n_shdr, UINT32_MAX);
    std::vector<int16_t> raw;
    for (uint32_t sid : used) {
        const Shdr &sh = shdrs[sid];
        const size_t n = sh.end - sh.start;
        raw.resize(n);
        if (!f.read(raw.data(), n * 2, smpl_offset + (uint64_t)sh.start * 2))
            throw std::runtime_error(path + ": sample data runs past the end of the file");
        const uint32_t src_rate = std::max<uint32_t>(sh.rate, 1);
        uint32_t loop_start = sh.start_loop > sh.start ? sh.start_loop - sh.start : 0;
        uint32_t loop_end = sh.end_loop > sh.start ? sh.end_loop - sh.start : 0;
        std::vector<int16_t> data;
        if (pool_rate != 0 && pool_rate != src_rate) {
            const double ratio = (double)pool_rate / (double)src_rate;
            data = resample_i16(raw.data(), n, ratio);
            loop_start = (uint32_t)sat_u64(std::round((double)loop_start * ratio));
            loop_end = (uint32_t)sat_u64(std::round((double)loop_end * ratio));
        } else {
            data = raw;
        }
        const uint32_t len = (uint32_t)data.size();
        loop_start = std::min(loop_start, len > 0 ? len - 1 : 0);
        loop_end = std::min(loop_end, len);
        kst_sample_info si{};
        si.start = (uint32_t)pool.size();
        si.len = len;
        si.loop_start = loop_start;
        si.loop_end = loop_end;
        si.rate = pool_rate ? pool_rate : src_rate;
        si.root_key = std::min<uint32_t>(sh.pitch, 127);
        si.correction_cents = (float)sh.correction;
        si.resample_ratio = pool_rate ? (float)pool_rate / (float)src_rate : 1.0f;
        pool.insert(pool.end(), data.begin(), data.end());
        pool.insert(pool.end(), POOL_GUARD, 0);
        remap[sid] = (uint32_t)samples.size();
        samples.push_back(si);
    }
    for (auto &r : regions) r.sample = r.sample < remap.size() ? remap[r.sample] : UINT32_MAX;
    return new Bank(set, pool.data(), pool.size(), pool_rate, samples.data(), samples.size(),
                    regions.data(), regions.size(), presets.data(), presets.size(),
                    preset_regions.data(), preset_regions.size());
}

}  // namespace kst
