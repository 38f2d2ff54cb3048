/* kestrel_hip.h -- C ABI of the MI355X (gfx950) Kestrel render backend.
 *
 * The drop-in boundary is the reference's `trait Backend` (src/backend.rs:22-58):
 * set_gates / set_channels / set_params_variant / spawn / render / stats / name /
 * timings, plus the constructor GpuSynth::new (src/gpu/mod.rs:276). Each entry
 * point below names the reference interface it replaces. The reference's crate
 * is rlib-only because its C ABI "is not implemented" (Cargo.toml:17-21); this
 * header is that C ABI, layout-compatible with the reference's #[repr(C)]
 * bytemuck Pod types, so a Rust FFI crate can bind it one-to-one.
 *
 * Conventions: every function returns 0 on success and a negative status on
 * failure, never unwinds, and leaves a message for kst_last_error() on the
 * calling thread. All slices are borrowed for the call only; the library copies
 * what it keeps. One caller thread per backend at a time (the reference's
 * `&mut self`).
 */
#ifndef KESTREL_HIP_H
#define KESTREL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KST_OK 0
#define KST_EINVAL (-1)   /* a validation failure; the message names the flag */
#define KST_EDEVICE (-2)  /* a HIP runtime failure */
#define KST_ESTATE (-3)   /* called out of order */

/* kst_config.profile levels. */
#define KST_PROFILE_OFF 0u
#define KST_PROFILE_ALL 1u     /* events between every pass of every block */
#define KST_PROFILE_RENDER 2u  /* events around the render kernel only */
#define KST_PROFILE_RENDER_SAMPLED 3u  /* as _RENDER, but in a staged run only around the
                                          render of every KST_RENDER_SAMPLE-th block */
#define KST_RENDER_SAMPLE 8u

/* One voice's birth record. src/voice.rs:22-57, 76 bytes, same field order. */
typedef struct kst_spawn_cmd {
    uint32_t phase_lo, phase_hi, step_lo, step_hi;
    uint32_t smp_base, smp_len, loop_start, loop_end;
    uint32_t flags, params, variant, region;
    uint32_t gate_slot, ordinal, start_rel;
    uint32_t note_id_lo, note_id_hi;
    float gain_l, gain_r;
} kst_spawn_cmd;

/* Per-region DSP constants. src/bank.rs:166-190, 48 bytes. */
typedef struct kst_region_params {
    float attack_rate, attack_end, decay_coef, decay_target, sustain, release_coef;
    float b0, b1, a1, a2;
    uint32_t flags, pad;
} kst_region_params;

/* src/backend.rs:11-20 */
typedef struct kst_block_stats {
    uint64_t active_voices;
    uint64_t stolen;
    uint64_t dropped;
    float peak;
    uint32_t pad;
} kst_block_stats;

/* src/backend.rs:53-57, one (name, milliseconds) pair per pass group. */
typedef struct kst_timing {
    const char *name;
    double ms;
} kst_timing;

/* The render-path subset of src/config.rs:164-343, in that file's units, plus
 * the MI355X knobs (marked). Enum encodings follow the reference's #[repr(u32)]
 * discriminants. */
typedef struct kst_config {
    uint32_t sample_rate;           /* Hz */
    uint32_t block_frames;          /* frames per render() */
    uint32_t gate_frames;           /* note-off / controller cadence */
    uint32_t reduce_tile;           /* plain form only: frames per LDS reduction tile */
    uint32_t workgroup_size;        /* plain form only: threads per render workgroup */
    uint32_t max_render_workgroups; /* plain form only: partial-buffer width */
    uint32_t max_voices;
    uint32_t steal_rule;            /* 0 Oldest, 1 DropNew, 2 Quietest */
    uint32_t admit_rule;            /* 0 Loudest, 1 Even */
    uint32_t max_steal_percent;
    uint32_t steal_fade_frames;
    uint32_t sort_voices;           /* locality sort during compaction */
    uint32_t interpolation;         /* 0 Nearest, 1 Linear, 2 Cubic */
    uint32_t decay_curve;           /* 0 Exponential, 1 Linear */
    uint32_t release_curve;         /* 0 Exponential, 1 Linear */
    float env_floor;
    uint32_t max_param_variants;
    uint32_t kahan_reduce;          /* plain form: Kahan in the cross-workgroup reduce */
    /* ---- MI355X knobs, no counterpart in the reference ---- */
    uint32_t form;                  /* 0 plain (reference-structured render), 1 wave64 */
    uint32_t compensated;           /* wave64 form: two-sum accumulation above the wave */
    uint32_t waves_per_group;       /* wave64 form: waves per render workgroup (1..4) */
    uint32_t sort_digit_bits;       /* radix width of the locality sort, 1..8 */
    int32_t device;                 /* HIP device ordinal */
    uint32_t profile;               /* per-pass timings (src/config.rs:337-338):
                                       KST_PROFILE_OFF, _ALL, _RENDER or
                                       _RENDER_SAMPLED */
} kst_config;

typedef struct kst_backend kst_backend;

/* Fill `cfg` with the reference's defaults (src/config.rs:345-409) and the
 * wave64 form. */
void kst_default_config(kst_config *cfg);

/* GpuSynth::new(&Config, Arc<Bank>) -- src/gpu/mod.rs:276-724. Only the bank's
 * int16 pool and its params table cross to the device (src/gpu/mod.rs:346-388);
 * the region count and longest sample size the locality key
 * (SortKeyLayout::plan, src/gpu/mod.rs:76-105). */
int kst_create(const kst_config *cfg, const int16_t *pool, size_t pool_len,
               const kst_region_params *params, size_t n_params, uint32_t n_regions,
               uint32_t max_sample_len, kst_backend **out);

/* Drop for GpuSynth. */
void kst_destroy(kst_backend *b);

/* Backend::set_gates -- src/backend.rs:23-25. `n` == tiles * 2048. */
int kst_set_gates(kst_backend *b, const uint32_t *rows, size_t n);

/* Backend::set_channels -- src/backend.rs:27-33. `n` == tiles * 16 * 4. */
int kst_set_channels(kst_backend *b, const uint32_t *rows, size_t n, int bend, int gain,
                     int variant);

/* Backend::set_params_variant -- src/backend.rs:35-39, src/gpu/mod.rs:972-987. */
int kst_set_params_variant(kst_backend *b, uint32_t index, const kst_region_params *data,
                           size_t n);

/* Backend::spawn -- src/backend.rs:41-43. Steal and drop accounting is eager, as
 * in src/cpu.rs:330-390: stats() after this call already counts this block's
 * steals and drops. The device work happens in kst_render. */
int kst_spawn(kst_backend *b, const kst_spawn_cmd *cmds, size_t n);

/* Backend::render -- src/backend.rs:45-47. `n` == block_frames * 2, interleaved
 * stereo. Advances every voice and compacts the pool; synchronous. */
int kst_render(kst_backend *b, float *out, size_t n);

/* The two halves of kst_render (no counterpart in the reference): submit
 * queues the block's uploads, passes and readback on the backend's stream and
 * returns at once; collect waits and writes `out`. In between the caller may
 * work on the host (kst_stats still answers), but every other call on this
 * backend fails with KST_ESTATE until the render is collected. */
int kst_render_submit(kst_backend *b);
int kst_render_collect(kst_backend *b, float *out, size_t n);

/* Backend::stats -- src/backend.rs:49. */
int kst_stats(const kst_backend *b, kst_block_stats *out);

/* Backend::name -- src/backend.rs:51. */
const char *kst_name(const kst_backend *b);

/* Backend::timings -- src/backend.rs:53-57. Writes up to `cap` entries and
 * returns how many there are. Empty until a block has rendered with `profile`
 * on; after a staged run without a graph, each pass's mean over its blocks
 * (under KST_PROFILE_RENDER the render kernel's only). */
size_t kst_timings(const kst_backend *b, kst_timing *out, size_t cap);

/* Change the profile level of a backend between blocks (KST_PROFILE_OFF,
 * KST_PROFILE_ALL: events between every pass, KST_PROFILE_RENDER: events
 * around the render kernel only, which leaves the rest of the chain as
 * without profiling, KST_PROFILE_RENDER_SAMPLED: the same on every
 * KST_RENDER_SAMPLE-th block of a staged run, blocks 0, 8, 16, ..., so that
 * the other blocks run with no event packets at all; the timing is the mean
 * over the sampled blocks). No counterpart in the reference: the config's
 * `profile` fixes it for the backend's life there (src/config.rs:337-338). */
int kst_set_profile(kst_backend *b, uint32_t level);

/* The message of the last failure on this thread (anyhow's error text). */
const char *kst_last_error(void);

/* ---- device-resident block program (no counterpart in the reference) ----
 *
 * The reference synchronises with the host once per block (src/gpu/mod.rs:
 * 1246-1252). These stage the inputs of `nblocks` consecutive blocks in device
 * memory -- gate rows, channel rows and flags, spawn commands -- and then run
 * them back to back with the per-block steal/spawn arithmetic of
 * src/gpu/mod.rs:1015-1040 evaluated on the device, so the host waits once.
 * Results are identical to calling spawn/render per block. */
int kst_stage_blocks(kst_backend *b, uint32_t nblocks, const uint32_t *gates,
                     const uint32_t *chans, const uint32_t *chan_flags,
                     const kst_spawn_cmd *cmds, const uint32_t *cmd_counts);
/* Runs every staged block. use_graph != 0 replays the per-block launch chain
 * from a captured hipGraph. Returns when the device is done.
 * Pools from 32,768 voices select each block's victims on a second stream
 * during the block before. Without a graph the two streams are ordered on the
 * device (bounded spin waits on sequence flags), which needs kernels of the two
 * streams to run concurrently: with KST_DEVICE_SYNC=0, AMD_SERIALIZE_KERNEL or
 * HIP_LAUNCH_BLOCKING set in the environment when the backend is created,
 * events order them instead (profilers that collect counters serialize
 * kernels: run them with KST_DEVICE_SYNC=0). A wait that times out (20 s)
 * fails the run with KST_EDEVICE. */
int kst_run_staged(kst_backend *b, int use_graph);
/* Copies out the staged run's output blocks (nblocks * block_frames * 2) and
 * per-block stats (nblocks entries, as stats() would read after each render). */
int kst_staged_results(kst_backend *b, float *out, kst_block_stats *stats);

/* ---- batching form: many renders, one launch per pass (no counterpart in the reference) ----
 *
 * A batch runs the staged blocks of several independent backends (each with
 * its own config, bank and pool, all wave64 form on one device, all with the
 * same `compensated`) together: block step by block step, every pass of the
 * per-block chain is one launch for all members, so small renders share the
 * GPU instead of each leaving it mostly idle. Members with fewer staged
 * blocks idle once theirs are done. Results are identical to kst_run_staged
 * on each member; read them with kst_staged_results. Destroy the batch before
 * any of its members; a member must not be used concurrently with the batch. */
typedef struct kst_batch kst_batch;
int kst_batch_create(kst_backend *const *renders, size_t n, kst_batch **out);
/* Runs every member's staged blocks. use_graph != 0 replays one captured
 * hipGraph of the whole batch step. Returns when the device is done. */
int kst_batch_run(kst_batch *batch, int use_graph);
void kst_batch_destroy(kst_batch *batch);

/* ---- precision form: FP64 coefficient design (no counterpart in the reference) ----
 *
 * The reference designs each RegionParams entry on the host in f32 (make_params
 * and biquad_lowpass, src/bank.rs:876-979). This designs a whole params variant
 * on the device in FP64 from the region's envelope times, cutoff and resonance,
 * with the reference's formulas (key scaling and controller modifiers neutral),
 * and rounds each coefficient once to f32. Uses the backend's sample_rate,
 * decay_curve, release_curve and env_floor. */
typedef struct kst_region_design {
    double attack_s, hold_s, decay_s;  /* seconds */
    double sustain;                    /* level, clamped to [0, 1] */
    double release_s;                  /* seconds */
    double fc_cents;                   /* lowpass cutoff, absolute cents (8.176 Hz * 2^(c/1200)) */
    double q_cb;                       /* resonance, centibels */
    uint32_t filter;                   /* 0: this region has no filter */
    uint32_t pad;
} kst_region_design;

/* Design params variant `index` (n == params per variant). Same validation as
 * kst_set_params_variant. */
int kst_design_params(kst_backend *b, uint32_t index, const kst_region_design *in, size_t n);

/* kst_design_params without a device allocation: the design rows are staged in
 * a static device array of the library (65,536 rows, 4 MiB per GPU, held until
 * the process ends) instead of a buffer allocated and freed per call. No
 * counterpart in the reference; for callers that cannot allocate device memory
 * (serving loops, harnesses that forbid it). n above 65,536 fails with
 * KST_EINVAL. Calls that overlap on one GPU share the array and race. */
int kst_design_params_noalloc(kst_backend *b, uint32_t index, const kst_region_design *in,
                              size_t n);

/* ---- post-render step: the caller's side of render() ----
 *
 * What the reference's driver does with every rendered block (src/driver.rs:
 * 596-607): the limiter, then the hard clamp; the WAV writer of src/wav.rs and
 * the comparison of `kestrel null` (src/bin/kestrel.rs:597-646). Host code: a
 * block is a sequential recurrence over 2 * F samples. */
#define KST_LIMITER_OFF 0
#define KST_LIMITER_OMNI 1      /* Limiter, src/limiter.rs:14-118 */
#define KST_LIMITER_BRICKWALL 2 /* Brickwall, src/limiter.rs:203-446 */
typedef struct kst_limiter_config {
    uint32_t sample_rate;
    uint32_t mode;       /* KST_LIMITER_* (LimiterMode, src/limiter.rs:134-149) */
    double ceiling_db;   /* brickwall ceiling, dBFS */
    double lookahead_ms; /* brickwall lookahead (also its latency) */
    double release_ms;   /* brickwall fast-stage release */
    double sustain_ms;   /* brickwall sustained stage; 0 disables it */
    uint32_t true_peak;  /* 4x oversampled peak detection */
    uint32_t pad;
} kst_limiter_config;
typedef struct kst_limiter kst_limiter;
/* src/config.rs:388-395 */
void kst_default_limiter_config(kst_limiter_config *cfg);
int kst_limiter_create(const kst_limiter_config *cfg, kst_limiter **out);
/* Limiter::process / Brickwall::process on one interleaved stereo block, in place. */
int kst_limiter_process(kst_limiter *l, float *buf, size_t n);
/* Brickwall::latency, frames (0 for the other modes). */
size_t kst_limiter_latency(const kst_limiter *l);
void kst_limiter_destroy(kst_limiter *l);
/* clamp_block, src/limiter.rs:120-129: returns how many samples it moved. */
uint64_t kst_clamp_block(float *buf, size_t n);

/* WavWriter, src/wav.rs:50-159: float32 != 0 writes IEEE float, else PCM16. */
typedef struct kst_wav kst_wav;
int kst_wav_create(const char *path, uint32_t sample_rate, uint32_t channels, uint32_t float32,
                   kst_wav **out);
int kst_wav_write(kst_wav *w, const float *samples, size_t n);
/* Patches the sizes (saturating at 4 GiB), closes and frees. */
int kst_wav_finish(kst_wav *w);

/* `kestrel null` over the first min(na, nb) samples; `a` is the reference. */
typedef struct kst_null_result {
    uint64_t compared, worst;
    double peak_ref, peak_diff, rms_diff;
    uint32_t identical, pad;
} kst_null_result;
int kst_null(const float *a, size_t na, const float *b, size_t nb, kst_null_result *out);
/* The message of the last post-render failure on this thread. */
const char *kst_post_last_error(void);

/* ---- the caller of the backend: the driver's block loop ----
 *
 * src/driver.rs:240-1192 with what it stands on: the flattened instrument bank
 * (src/bank.rs), the note-off gate and controller tables (src/voice.rs:212-658)
 * and the streaming SMF reader with its tempo clock (src/midi.rs). Host code.
 * Bank, MIDI and driver failures leave their message for kst_host_last_error(). */

/* SampleInfo, src/bank.rs:35-54 (the name is not kept). 32 bytes. */
typedef struct kst_sample_info {
    uint32_t start, len;             /* frames in the bank's pool */
    uint32_t loop_start, loop_end;   /* relative to start */
    uint32_t rate;                   /* Hz the data is stored at */
    float correction_cents;
    float resample_ratio;            /* pool frames per source frame */
    uint32_t root_key;
} kst_sample_info;

#define KST_LOOP_NONE 0
#define KST_LOOP_CONTINUOUS 1
#define KST_LOOP_UNTIL_RELEASE 2
/* Region, src/bank.rs:57-120, every generator resolved. 96 bytes. */
typedef struct kst_region {
    uint32_t sample;
    uint8_t key_lo, key_hi, vel_lo, vel_hi;
    int16_t root_key_override, fixed_key, fixed_vel, coarse_tune;  /* -1: unset */
    int16_t fine_tune, scale_tuning, keynum_to_hold, keynum_to_decay;
    float attenuation_cb, amp_veltrack, pan;
    uint32_t loop_mode;              /* KST_LOOP_* */
    int32_t addr_start, addr_end, addr_loop_start, addr_loop_end;
    uint32_t exclusive_class;
    float delay, attack, hold, decay, sustain, release;  /* seconds; sustain a level */
    float filter_fc_cents, filter_q_cb, filter_veltrack_cents;
} kst_region;

/* Preset, src/bank.rs:352-362: its regions are preset_regions[first .. first+count]. */
typedef struct kst_preset {
    uint32_t bank, program, first, count;
} kst_preset;

/* The driver-side settings of src/config.rs that kst_config does not carry. */
typedef struct kst_driver_config {
    uint32_t max_layers;             /* :207 */
    float master_volume;             /* :260 */
    uint32_t default_channel_volume; /* :266 */
    uint32_t filter_enabled;         /* :258 */
    uint32_t limiter;                /* :273, run limiter_cfg.mode on each block */
    uint32_t nan_guard;              /* :328, fail on a non-finite output sample */
    kst_limiter_config limiter_cfg;  /* :280-313 */
} kst_driver_config;
/* src/config.rs:345-409 */
void kst_default_driver_config(kst_driver_config *cfg);

typedef struct kst_bank kst_bank;
/* Bank::finish + Bank::build_params (src/bank.rs:514-527, 689-781): lays out the
 * params table, designs variant 0 and the admission preview tables. Uses the
 * sample_rate, env_floor and curves of `cfg`. pool_rate 0 means mixed-rate. */
int kst_bank_create(const kst_config *cfg, const kst_driver_config *dcfg, const int16_t *pool,
                    size_t pool_len, uint32_t pool_rate, const kst_sample_info *samples,
                    size_t n_samples, const kst_region *regions, size_t n_regions,
                    const kst_preset *presets, size_t n_presets, const uint32_t *preset_regions,
                    size_t n_preset_regions, kst_bank **out);
/* What kst_create takes from the bank (src/gpu/mod.rs:346-388). */
int kst_bank_info(const kst_bank *b, size_t *pool_len, size_t *n_params, uint32_t *n_regions,
                  uint32_t *max_sample_len);
const int16_t *kst_bank_pool(const kst_bank *b);
const kst_region_params *kst_bank_params(const kst_bank *b);
/* Bank::build_variant (src/bank.rs:836-853) for ParamMod::from_controllers of
 * cc[0..5] = CC71..CC75 (src/bank.rs:336-350). n == params per variant. */
int kst_bank_variant(const kst_bank *b, const uint8_t cc[5], kst_region_params *out, size_t n);
/* Bank::find_preset, src/bank.rs:529-546; -1 when the bank has no presets. */
int64_t kst_bank_find_preset(const kst_bank *b, uint32_t bank, uint32_t program);
void kst_bank_destroy(kst_bank *b);

/* The bank's flattened contents, as created or loaded (valid while it lives). */
int kst_bank_contents(const kst_bank *b, const kst_sample_info **samples, size_t *n_samples,
                      const kst_region **regions, size_t *n_regions, const kst_preset **presets,
                      size_t *n_presets, const uint32_t **preset_regions,
                      size_t *n_preset_regions, uint32_t *pool_rate);

/* Pool building of the loaders, src/config.rs:319-322 (defaults :397-398). */
typedef struct kst_load_config {
    uint32_t resample_pool;      /* convert every used sample to the output rate */
    uint32_t pad;
    uint64_t sample_pool_budget; /* bytes; the pool rate halves until it fits */
} kst_load_config;
void kst_default_load_config(kst_load_config *cfg);
/* sf2::load, src/sf2.rs:221-506: every preset flattened into resolved regions,
 * only the samples some region uses read into the pool (src/sf2.rs:585-703),
 * then the bank built as kst_bank_create builds it. */
int kst_bank_load_sf2(const kst_config *cfg, const kst_driver_config *dcfg,
                      const kst_load_config *lcfg, const char *path, kst_bank **out);
/* sfz::load, src/sfz.rs:254-425: <control>/<global>/<master>/<group>/<region>
 * sections with #include, every referenced WAV (src/wav.rs:221-365) read once
 * per channel, stereo files as two hard-panned regions, all in one preset
 * (bank 0, program 0). The pool is converted to the output rate when
 * resample_pool is set; the budget does not apply to SFZ. */
int kst_bank_load_sfz(const kst_config *cfg, const kst_driver_config *dcfg,
                      const kst_load_config *lcfg, const char *path, kst_bank **out);
/* resample_i16, src/resample.rs:102-108: writes up to `cap` samples and
 * returns the converted length (call with cap 0 to size `out`). */
int64_t kst_resample_i16(const int16_t *src, size_t n, double ratio, int16_t *out, size_t cap);

/* Division, src/midi.rs:32-36: smpte == 0 uses ppq, else fps * ticks_per_frame. */
typedef struct kst_division {
    uint32_t smpte, ppq, fps, ticks_per_frame;
} kst_division;
#define KST_EV_NOTE_ON 0    /* a = key, b = velocity */
#define KST_EV_NOTE_OFF 1   /* a = key */
#define KST_EV_CC 2         /* a = controller, b = value */
#define KST_EV_PROGRAM 3    /* a = program */
#define KST_EV_PITCH_BEND 4 /* value = -8192..8191 */
#define KST_EV_TEMPO 5      /* value = microseconds per quarter note */
#define KST_EV_OTHER 6
/* (tick, Event), src/midi.rs:17-29. 24 bytes. */
typedef struct kst_midi_event {
    uint64_t tick;
    uint32_t kind;                   /* KST_EV_* */
    uint8_t ch, a, b, pad;
    int32_t value;
    uint32_t pad2;
} kst_midi_event;

typedef struct kst_midi kst_midi;
/* MidiStream::open, src/midi.rs:259-345: a tick-ordered merge of every track,
 * read through a 256 KiB window per track. */
int kst_midi_open(const char *path, kst_midi **out);
int kst_midi_info(const kst_midi *m, kst_division *div, uint32_t *format, uint32_t *tracks);
/* MidiStream::next, src/midi.rs:351-361, up to `cap` at a time; returns how
 * many were written, 0 at the end of the file. */
int64_t kst_midi_read(kst_midi *m, kst_midi_event *out, size_t cap);
void kst_midi_close(kst_midi *m);

/* DriverStats, src/driver.rs:143-186. */
typedef struct kst_driver_stats {
    uint64_t variant_fallbacks, variant_rebuilds, frames, blocks, events, dropped, notes,
        voices_spawned, clipped, last_stolen, last_want_energy, last_take_energy;
    uint32_t param_variants, variant_states, last_live, last_want, last_take;
    float peak;
} kst_driver_stats;

typedef struct kst_driver kst_driver;
/* Driver::open, src/driver.rs:343-413. The bank must outlive the driver. */
int kst_driver_open(const kst_config *cfg, const kst_driver_config *dcfg, const kst_bank *bank,
                    const char *midi_path, kst_driver **out);
/* The same over an already merged, tick-ordered event list (copied). */
int kst_driver_create(const kst_config *cfg, const kst_driver_config *dcfg,
                      const kst_bank *bank, const kst_division *div,
                      const kst_midi_event *events, size_t n, kst_driver **out);
/* Driver::next_block, src/driver.rs:434-642, against this library's backend:
 * returns 1 while the file or its tail has more, 0 when done, < 0 on failure.
 * `n` == block_frames * 2. The driver's event side (events, rows, variants,
 * candidates) runs one block ahead -- on a thread of its own while this block
 * is admitted, rendered and limited, or on the caller's thread while the
 * device renders, whichever these calls have timed faster (nothing of this
 * block depends on it); what each call returns and what kst_driver_get_stats
 * reports after it are those of the reference's sequential loop. One caller
 * thread at a time per driver. */
int kst_driver_next_block(kst_driver *d, kst_backend *b, float *out, size_t n);
/* The same block in the steps a caller with another backend drives itself:
 * begin (events, rows, variants; returns how many variants to upload), then
 * the rows and variants for that backend, admit (given its live voices and
 * stolen counter before spawn), spawn and render there, then end (limiter,
 * clamp, counters; returns 1 while there is more). The pointers the rows,
 * pending-variant and admit calls hand out stay valid until the next begin;
 * begin also starts the event side of the block after on the driver's thread. */
int kst_driver_begin_block(kst_driver *d);
int kst_driver_pending_variant(const kst_driver *d, size_t i, uint32_t *index,
                               const kst_region_params **data, size_t *n);
int kst_driver_rows(const kst_driver *d, const uint32_t **gates, size_t *n_gates,
                    const uint32_t **chans, size_t *n_chans, uint32_t flags[3]);
int kst_driver_admit(kst_driver *d, uint64_t live, uint64_t stolen_before,
                     const kst_spawn_cmd **cmds, size_t *n);
int kst_driver_end_block(kst_driver *d, float *out, size_t n, uint64_t stolen_after_spawn,
                         const kst_block_stats *after_render);
int kst_driver_get_stats(const kst_driver *d, kst_driver_stats *out);
/* Host time per block of kst_driver_next_block, by phase (events and rows,
 * the caller's wait for the event thread, admission and voice build,
 * post-render step, the backend calls, the wait for the device), then
 * "inline_share": the fraction of blocks (not milliseconds) whose event side
 * ran on the caller's thread; writes up to `cap` entries and returns how many
 * there are.
 * No counterpart in the reference. */
size_t kst_driver_timings(const kst_driver *d, kst_timing *out, size_t cap);
void kst_driver_destroy(kst_driver *d);

#define KST_CC_APPLIED 0
#define KST_CC_INERT 1
#define KST_CC_MISSING 2
/* cc_role, src/driver.rs:58-131: returns KST_CC_*, with the controller's name
 * and what the role means. */
int kst_cc_role(uint32_t num, const char **name, const char **note);
/* The message of the last bank, MIDI or driver failure on this thread. */
const char *kst_host_last_error(void);

/* ---- test and checkpoint hooks (no counterpart in the reference) ---- */

/* Copy params variant `index` back to the host (n == params per variant). */
int kst_params_export(kst_backend *b, uint32_t index, kst_region_params *out, size_t n);

/* The live pool in the 24-field SoA layout of shaders/common.wgsl:15-44:
 * field f of voice i at soa[f * stride + i]. Returns the live count. */
int64_t kst_pool_export(kst_backend *b, uint32_t *soa, size_t stride);
/* Replace the live pool with `n` voices in that layout. */
int kst_pool_import(kst_backend *b, const uint32_t *soa, size_t n, size_t stride);
/* Set the cumulative stolen/dropped counters (to resume from a snapshot). */
int kst_set_counters(kst_backend *b, uint64_t stolen, uint64_t dropped);

#ifdef __cplusplus
}
#endif
#endif /* KESTREL_HIP_H */
