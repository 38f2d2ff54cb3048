// kst_midi.cpp -- the streaming SMF reader and tempo clock (src/midi.rs).
//
// Every track chunk is read through its own window with pread() on one shared
// file descriptor, and a min-heap on (tick, track) merges the per-track
// cursors, so a file of any size costs a bounded amount of memory and ties
// break on track order. Decoding follows src/midi.rs:128-247 byte for byte,
// including how it treats running status, truncated events and the
// end-of-track meta event.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cstring>
#include <stdexcept>

#include "kst_host.h"

namespace kst {
namespace {

constexpr size_t TRACK_WINDOW = 256 * 1024;   // src/midi.rs:15
constexpr size_t MIN_WINDOW = 4096;
constexpr size_t WINDOW_BUDGET = 256u << 20;  // across all tracks of one file

uint32_t be32(const uint8_t *p) { return ((uint32_t)p[0] << 24) | (p[1] << 16) | (p[2] << 8) | p[3]; }
uint16_t be16(const uint8_t *p) { return (uint16_t)((p[0] << 8) | p[1]); }

bool pread_all(int fd, uint8_t *dst, size_t n, uint64_t off) {
    size_t got = 0;
    while (got < n) {
        const ssize_t r = pread(fd, dst + got, n - got, (off_t)(off + got));
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        got += (size_t)r;
    }
    return true;
}

}  // namespace

MidiFile::MidiFile(const std::string &path) {
    const int fd = open(path.c_str(), O_RDONLY | O_CLOEXEC);
    if (fd < 0) throw std::runtime_error("opening " + path + ": " + std::strerror(errno));
    struct stat st {};
    if (fstat(fd, &st) != 0) {
        close(fd);
        throw std::runtime_error("opening " + path + ": " + std::strerror(errno));
    }
    const uint64_t file_len = (uint64_t)st.st_size;
    uint8_t hdr[14];
    if (file_len < 14 || !pread_all(fd, hdr, 14, 0)) {
        close(fd);
        throw std::runtime_error(path + ": too short for an SMF header");
    }
    if (std::memcmp(hdr, "MThd", 4) != 0) {
        close(fd);
        throw std::runtime_error(path + ": not a standard MIDI file");
    }
    const uint64_t hdr_len = be32(hdr + 4);
    format = be16(hdr + 8);
    const int16_t div_raw = (int16_t)be16(hdr + 12);
    division = kst_division{};
    if (div_raw > 0) {
        division.ppq = (uint32_t)div_raw;
    } else {  // SMPTE: negative frame rate in the high byte (src/midi.rs:275-281)
        division.smpte = 1;
        division.fps = (uint8_t)(-(div_raw >> 8));
        division.ticks_per_frame = (uint8_t)(div_raw & 0xFF);
    }

    // Chunk walk over the 8-byte headers only (src/midi.rs:283-322).
    std::vector<std::pair<uint64_t, uint64_t>> locs;
    uint64_t offset = 8 + hdr_len;
    while (offset + 8 <= file_len) {
        uint8_t ch[8];
        if (!pread_all(fd, ch, 8, offset)) break;
        const uint64_t data_start = offset + 8;
        const uint64_t avail = file_len > data_start ? file_len - data_start : 0;
        const uint64_t len = std::min<uint64_t>(be32(ch + 4), avail);
        if (std::memcmp(ch, "MTrk", 4) == 0) {
            locs.emplace_back(data_start, len);
        } else if (len == 0) {
            break;  // zero padding after the last chunk is not a chunk list
        }
        offset = data_start + len;
    }
    if (locs.empty()) {
        close(fd);
        throw std::runtime_error(path + ": no MTrk chunks");
    }
    track_count = (uint32_t)locs.size();
    const size_t window =
        std::max(MIN_WINDOW, std::min(TRACK_WINDOW, WINDOW_BUDGET / locs.size()));
    fd_ = fd;
    const uint32_t T = (uint32_t)locs.size();
    wide_ = T >= (1u << 20);
    tracks_.resize(T);
    for (uint32_t i = 0; i < T; ++i) {
        Track &t = tracks_[i];
        t.fd = fd;
        t.off = locs[i].first;
        t.remaining = locs[i].second;
        t.buf.resize((size_t)std::min<uint64_t>(window, std::max<uint64_t>(locs[i].second, 1)));
        t.refill();
    }
    while (leaves_ < T) leaves_ <<= 1;
    node_.assign(2 * (size_t)leaves_, UINT64_MAX);
    for (uint32_t i = 0; i < T; ++i) node_[leaves_ + i] = leaf_key(i);
    for (uint32_t i = leaves_ - 1; i >= 1; --i) node_[i] = std::min(node_[2 * i], node_[2 * i + 1]);
}

MidiFile::~MidiFile() {
    if (fd_ >= 0) close(fd_);
}

uint64_t MidiFile::leaf_key(uint32_t t) const {
    const Track &k = tracks_[t];
    if (k.at == k.n) return UINT64_MAX;
    const uint64_t tick = k.ev[k.at].tick;
    if (tick >= (1ull << 44)) return UINT64_MAX - 1;  // marks the wide case
    return (tick << 20) | t;
}

// The next event by (tick, track) over every track, for ticks too wide for
// the packed key (src/midi.rs:246-249).
uint32_t MidiFile::scan_min() const {
    uint32_t best = UINT32_MAX;
    for (uint32_t t = 0; t < tracks_.size(); ++t) {
        const Track &k = tracks_[t];
        if (k.at == k.n) continue;
        if (best == UINT32_MAX || k.ev[k.at].tick < tracks_[best].ev[tracks_[best].at].tick) best = t;
    }
    return best;
}

// Decode the next batch of a track's events (its tick carried in each).
bool MidiFile::Track::refill() {
    if (ev.empty()) ev.resize(256);
    at = n = 0;
    while (n < ev.size() && next_event(ev[n])) ev[n++].tick = tick;
    return n > 0;
}

void MidiFile::update(uint32_t t) {
    uint32_t i = leaves_ + t;
    node_[i] = leaf_key(t);
    for (i >>= 1; i >= 1; i >>= 1) node_[i] = std::min(node_[2 * i], node_[2 * i + 1]);
}

bool MidiFile::next(kst_midi_event &ev) {  // src/midi.rs:351-361
    if (run_) {
        Track &t = tracks_[run_w_];
        if (t.at < t.n || t.refill()) {
            const uint64_t tick = t.ev[t.at].tick;
            if (tick < (1ull << 44) && ((tick << 20) | run_w_) < run_limit_) {
                ev = t.ev[t.at++];
                return true;
            }
        }
        run_ = false;
        update(run_w_);
    }
    if (!wide_ && node_[1] == UINT64_MAX - 1) wide_ = true;
    if (wide_) {
        const uint32_t w = scan_min();
        if (w == UINT32_MAX) return false;
        Track &t = tracks_[w];
        ev = t.ev[t.at++];
        if (t.at == t.n) t.refill();
        return true;
    }
    if (node_[1] == UINT64_MAX) return false;
    const uint32_t w = (uint32_t)(node_[1] & ((1u << 20) - 1));
    uint64_t limit = UINT64_MAX;
    for (uint32_t i = leaves_ + w; i > 1; i >>= 1) limit = std::min(limit, node_[i ^ 1]);
    Track &t = tracks_[w];
    ev = t.ev[t.at++];
    run_ = true;
    run_w_ = w;
    run_limit_ = limit;
    return true;
}

bool MidiFile::Track::fill() {
    if (remaining == 0) return false;
    const size_t want = (size_t)std::min<uint64_t>(remaining, buf.size());
    size_t got = 0;
    while (got < want) {
        const ssize_t r = pread(fd, buf.data() + got, want - got, (off_t)(off + got));
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) break;
        got += (size_t)r;
    }
    remaining -= got;
    off += got;
    pos = 0;
    filled = got;
    return got > 0;
}

bool MidiFile::Track::byte(uint8_t &b) {
    if (pos == filled && !fill()) return false;
    b = buf[pos++];
    return true;
}

bool MidiFile::Track::varlen(uint64_t &v) {  // at most 8 bytes, src/midi.rs:101-112
    v = 0;
    for (int i = 0; i < 8; ++i) {
        uint8_t b;
        if (!byte(b)) return false;
        v = (v << 7) | (b & 0x7F);
        if (!(b & 0x80)) return true;
    }
    return false;
}

bool MidiFile::Track::skip(uint64_t n) {
    while (n > 0) {
        if (pos == filled && !fill()) return false;
        const uint64_t take = std::min<uint64_t>(filled - pos, n);
        pos += (size_t)take;
        n -= take;
    }
    return true;
}

// One event, advancing `tick`; false at the end of the track (or where the
// track is cut short, which ends it the same way).
bool MidiFile::Track::next_event(kst_midi_event &ev) {  // src/midi.rs:128-247
    if (ended) return false;
    uint64_t delta;
    if (!varlen(delta)) {
        ended = true;
        return false;
    }
    tick += delta;
    uint8_t status;
    if (!byte(status)) {
        ended = true;
        return false;
    }
    uint8_t first = 0;
    if (status < 0x80) {  // running status: this byte is the first data byte
        first = status;
        status = running;
        if (status < 0x80) {
            ended = true;
            return false;
        }
    } else {
        if (status < 0xF0)
            running = status;
        else if (status == 0xFF)
            running = 0;
        if (status != 0xFF && status != 0xF0 && status != 0xF7 && !byte(first)) {
            ended = true;
            return false;
        }
    }
    ev = kst_midi_event{};
    ev.kind = KST_EV_OTHER;
    ev.ch = status & 0x0F;
    uint8_t b;
    switch (status & 0xF0) {
    case 0x80:
        if (!byte(b)) return ended = true, false;
        ev.kind = KST_EV_NOTE_OFF;
        ev.a = first & 0x7F;
        return true;
    case 0x90:
        if (!byte(b)) return ended = true, false;
        ev.kind = b == 0 ? KST_EV_NOTE_OFF : KST_EV_NOTE_ON;
        ev.a = first & 0x7F;
        ev.b = b == 0 ? 0 : (b & 0x7F);
        return true;
    case 0xA0:
        if (!byte(b)) return ended = true, false;
        ev.ch = 0;
        return true;
    case 0xB0:
        if (!byte(b)) return ended = true, false;
        ev.kind = KST_EV_CC;
        ev.a = first & 0x7F;
        ev.b = b & 0x7F;
        return true;
    case 0xC0:
        ev.kind = KST_EV_PROGRAM;
        ev.a = first & 0x7F;
        return true;
    case 0xD0:
        ev.ch = 0;
        return true;
    case 0xE0:
        if (!byte(b)) return ended = true, false;
        ev.kind = KST_EV_PITCH_BEND;
        ev.value = (int32_t)((((int)b & 0x7F) << 7) | (first & 0x7F)) - 8192;
        return true;
    default:
        break;
    }
    ev.ch = 0;
    if (status == 0xFF) {
        uint8_t meta;
        uint64_t len;
        if (!byte(meta) || !varlen(len)) return ended = true, false;
        if (meta == 0x2F) {  // end of track: reported, then nothing more
            if (!skip(len)) return ended = true, false;
            ended = true;
            return true;
        }
        if (meta == 0x51 && len == 3) {
            uint8_t x, y, z;
            if (!byte(x) || !byte(y) || !byte(z)) return ended = true, false;
            ev.kind = KST_EV_TEMPO;
            ev.value = (int32_t)(((uint32_t)x << 16) | ((uint32_t)y << 8) | z);
            return true;
        }
        if (!skip(len)) return ended = true, false;
        return true;
    }
    if (status == 0xF0 || status == 0xF7) {
        uint64_t len;
        if (!varlen(len) || !skip(len)) return ended = true, false;
        return true;
    }
    return true;  // other system messages: one data byte already taken
}

Prefetch::Prefetch(std::unique_ptr<EventSource> inner, size_t chunk, size_t depth)
    : inner_(std::move(inner)), chunk_(chunk), depth_(depth) {
    division = inner_->division;
    th_ = std::thread([this] { run(); });
}

Prefetch::~Prefetch() {
    {
        std::lock_guard<std::mutex> g(m_);
        stop_ = true;
    }
    cv_.notify_all();
    th_.join();
}

void Prefetch::run() {
    for (;;) {
        std::vector<kst_midi_event> c;
        {
            std::lock_guard<std::mutex> g(m_);
            if (!spare_.empty()) {
                c = std::move(spare_.back());
                spare_.pop_back();
            }
        }
        c.clear();
        c.reserve(chunk_);
        kst_midi_event e;
        while (c.size() < chunk_ && inner_->next(e)) c.push_back(e);
        const bool last = c.size() < chunk_;
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [&] { return stop_ || q_.size() < depth_; });
        if (stop_) return;
        if (!c.empty()) q_.push_back(std::move(c));
        if (last) done_ = true;
        g.unlock();
        cv_.notify_all();
        if (last) return;
    }
}

bool Prefetch::refill() {
    std::unique_lock<std::mutex> g(m_);
    cv_.wait(g, [&] { return !q_.empty() || done_; });
    if (q_.empty()) return false;
    if (cur_.capacity()) spare_.push_back(std::move(cur_));
    cur_ = std::move(q_.front());
    q_.pop_front();
    g.unlock();
    cv_.notify_all();
    at_ = 0;
    return true;
}

bool Prefetch::next(kst_midi_event &ev) {
    if (at_ >= cur_.size() && !refill()) return false;
    ev = cur_[at_++];
    return true;
}

size_t Prefetch::span(const kst_midi_event **first) {
    if (at_ >= cur_.size() && !refill()) return 0;
    *first = cur_.data() + at_;
    const size_t n = cur_.size() - at_;
    at_ = cur_.size();
    return n;
}

TempoClock::TempoClock(const kst_division &d, uint32_t sample_rate)
    : div_(d), sample_rate_((double)sample_rate) {
    set_raw(500000);  // 120 bpm until told otherwise
}

void TempoClock::set_raw(uint32_t us_per_qn) {  // src/midi.rs:390-401
    if (!div_.smpte) {
        const double ppq = (double)std::max<uint32_t>(div_.ppq, 1);
        frames_per_tick_ = ((double)us_per_qn * 1e-6) * sample_rate_ / ppq;
    } else {
        const double tps = (double)std::max<uint32_t>(div_.fps, 1) *
                           (double)std::max<uint32_t>(div_.ticks_per_frame, 1);
        frames_per_tick_ = sample_rate_ / tps;
    }
}

void TempoClock::set_tempo(uint64_t tick, uint32_t us_per_qn) {  // src/midi.rs:404-411
    if (div_.smpte) return;  // SMPTE division ignores tempo meta events
    base_frame_ = frame_at(tick);
    base_tick_ = tick;
    set_raw(us_per_qn);
}

}  // namespace kst
