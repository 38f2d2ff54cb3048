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
    tracks_.resize(locs.size());
    pending_.resize(locs.size());
    for (size_t i = 0; i < locs.size(); ++i) {
        Track &t = tracks_[i];
        t.fd = fd;
        t.off = locs[i].first;
        t.remaining = locs[i].second;
        t.buf.resize((size_t)std::min<uint64_t>(window, std::max<uint64_t>(locs[i].second, 1)));
        if (t.next_event(pending_[i])) heap_.emplace(t.tick, (uint32_t)i);
    }
}

MidiFile::~MidiFile() {
    if (fd_ >= 0) close(fd_);
}

bool MidiFile::next(kst_midi_event &ev) {  // src/midi.rs:351-361
    if (heap_.empty()) return false;
    const auto [tick, idx] = heap_.top();
    heap_.pop();
    ev = pending_[idx];
    ev.tick = tick;
    Track &t = tracks_[idx];
    if (t.next_event(pending_[idx])) heap_.emplace(t.tick, idx);
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
