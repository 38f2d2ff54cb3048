// kst_design.hip -- precision form: region coefficient design on the device in FP64.
//
// The reference designs every RegionParams entry on the host in f32
// (make_params / biquad_lowpass, src/bank.rs:876-979). At high resonance and low
// cutoff the f32 design loses most of its bits: 1 - cos(w0) cancels, and the
// pole radius sits next to 1. Here the same formulas run in FP64 on the device
// and each coefficient is rounded once to f32 for the f32 voice arithmetic.
//
// The transcendentals are written out from IEEE double operations only (rint,
// fma, ldexp, +, *, /), so the design is reproducible bit for bit on any
// device that implements those correctly rounded:
//  * 2^x: x = n + f with n = rint(x), |f| <= 1/2; 2^f by its Taylor series
//    sum (ln 2)^k f^k / k!, k <= 15 (truncation < 2^-68 relative), Horner with
//    fma; times 2^n exactly;
//  * sin, cos on [0, pi]: folded to |t| <= pi/4 around 0, pi/2 or pi (split
//    in two doubles), Taylor series through t^25 and t^26 (truncation < 2^-90).
// Powers of 10 are powers of 2 with exponent x * log2(10); the square root of
// the resonance gain is the power 2^(-1/4 - e/2) of the same exponent e.
#include "kst_device.h"

#include <mutex>

namespace kst {

// Staging for the allocation-free design entry point: the caller's design rows
// are copied here instead of into a buffer allocated per call. One copy per
// device (every GPU the library runs on holds it until the process ends); host
// code uses the address hipGetSymbolAddress returns for the current device,
// never the array's name. Calls that overlap on one device share it and race.
static __device__ __align__(256) kst_region_design scratch[DESIGN_SCRATCH_ROWS];

namespace {
std::mutex scratch_mu;
constexpr int SCRATCH_DEVICES = 64;
kst_region_design *scratch_ptr[SCRATCH_DEVICES] = {};
}  // namespace

// The scratch array's device address on `device`, which must be current.
hipError_t design_scratch(int device, kst_region_design **out) {
    *out = nullptr;
    if (device < 0 || device >= SCRATCH_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_ptr[device] = static_cast<kst_region_design *>(p);
    }
    *out = scratch_ptr[device];
    return hipSuccess;
}

namespace {

constexpr double LOG2_10 = 3.321928094887362;
constexpr double LOG2_MIN_Q = -9.965784284662087;  // log2(0.001), the design's Q floor
constexpr double PI_HI = 3.141592653589793;
constexpr double PI_LO = 1.2246467991473532e-16;
constexpr double HALF_PI = 1.5707963267948966;
constexpr double HALF_PI_LO = 6.123233995736766e-17;
constexpr double QUARTER_PI = 0.7853981633974483;
constexpr double TWO_PI = 6.283185307179586;

__device__ double exp2_det(double x) {
    const double n = __builtin_rint(x);
    const double f = x - n;
    double p = 3.1324367070884287e-15;
    p = __builtin_fma(p, f, 6.778726354822545e-14);
    p = __builtin_fma(p, f, 1.3691488853904128e-12);
    p = __builtin_fma(p, f, 2.5678435993488206e-11);
    p = __builtin_fma(p, f, 4.4455382718708116e-10);
    p = __builtin_fma(p, f, 7.054911620801123e-09);
    p = __builtin_fma(p, f, 1.01780860092397e-07);
    p = __builtin_fma(p, f, 1.321548679014431e-06);
    p = __builtin_fma(p, f, 1.5252733804059841e-05);
    p = __builtin_fma(p, f, 0.0001540353039338161);
    p = __builtin_fma(p, f, 0.0013333558146428443);
    p = __builtin_fma(p, f, 0.009618129107628477);
    p = __builtin_fma(p, f, 0.05550410866482158);
    p = __builtin_fma(p, f, 0.24022650695910072);
    p = __builtin_fma(p, f, 0.6931471805599453);
    p = __builtin_fma(p, f, 1.0);
    return __builtin_ldexp(p, (int)n);
}

// sin t and cos t for |t| <= pi/4: Taylor series through t^25 and t^26.
__device__ void sincos_poly(double t, double &s, double &c) {
    const double t2 = t * t;
    // sin t = t + t^3 * P(t^2), P = sum_{k=1..12} (-1)^k t^(2k-2) / (2k+1)!
    double ps = 6.446950284384474e-26;
    ps = __builtin_fma(ps, t2, -3.868170170630684e-23);
    ps = __builtin_fma(ps, t2, 1.9572941063391263e-20);
    ps = __builtin_fma(ps, t2, -8.22063524662433e-18);
    ps = __builtin_fma(ps, t2, 2.8114572543455206e-15);
    ps = __builtin_fma(ps, t2, -7.647163731819816e-13);
    ps = __builtin_fma(ps, t2, 1.6059043836821613e-10);
    ps = __builtin_fma(ps, t2, -2.505210838544172e-08);
    ps = __builtin_fma(ps, t2, 2.7557319223985893e-06);
    ps = __builtin_fma(ps, t2, -0.0001984126984126984);
    ps = __builtin_fma(ps, t2, 0.008333333333333333);
    ps = __builtin_fma(ps, t2, -0.16666666666666666);
    s = __builtin_fma(t * t2, ps, t);
    // cos t = 1 + t^2 * Q(t^2), Q = sum_{k=1..13} (-1)^k t^(2k-2) / (2k)!
    double pc = -2.4795962632247976e-27;
    pc = __builtin_fma(pc, t2, 1.6117375710961184e-24);
    pc = __builtin_fma(pc, t2, -8.896791392450574e-22);
    pc = __builtin_fma(pc, t2, 4.110317623312165e-19);
    pc = __builtin_fma(pc, t2, -1.5619206968586225e-16);
    pc = __builtin_fma(pc, t2, 4.779477332387385e-14);
    pc = __builtin_fma(pc, t2, -1.1470745597729725e-11);
    pc = __builtin_fma(pc, t2, 2.08767569878681e-09);
    pc = __builtin_fma(pc, t2, -2.755731922398589e-07);
    pc = __builtin_fma(pc, t2, 2.48015873015873e-05);
    pc = __builtin_fma(pc, t2, -0.001388888888888889);
    pc = __builtin_fma(pc, t2, 0.041666666666666664);
    pc = __builtin_fma(pc, t2, -0.5);
    c = __builtin_fma(t2, pc, 1.0);
}

// sin w and cos w for 0 <= w <= pi, folded to |t| <= pi/4 around 0, pi/2 or
// pi (each split in two doubles), so neither value loses its relative
// accuracy near a zero.
__device__ void sincos_det(double w, double &s, double &c) {
    double ps, pc;
    if (w <= QUARTER_PI) {
        sincos_poly(w, s, c);
    } else if (w <= 3.0 * QUARTER_PI) {
        sincos_poly((HALF_PI - w) + HALF_PI_LO, pc, ps);  // sin(pi/2 - w), cos(pi/2 - w)
        s = ps;
        c = pc;
    } else {
        sincos_poly((PI_HI - w) + PI_LO, ps, pc);
        s = ps;
        c = -pc;
    }
}

__device__ double clampd(double x, double lo, double hi) {
    return x < lo ? lo : (x > hi ? hi : x);
}

// src/bank.rs:876-954 and 965-979 with key scaling and controller modifiers
// neutral, in FP64, rounded once per coefficient.
__device__ kst_region_params design_one(const kst_region_design &d, double sr, bool exp_decay,
                                        bool exp_release, float env_floor) {
    const double attack_frames = fmax(d.attack_s * sr, 0.0);
    const double hold_frames = fmax(d.hold_s * sr, 0.0);
    const double decay_frames = fmax(d.decay_s * sr, 1.0);
    const double release_frames = fmax(d.release_s * sr, 1.0);
    const double attack_rate = attack_frames < 1.0 ? 1.0 : 1.0 / attack_frames;
    const double attack_end = 1.0 + hold_frames * attack_rate;
    const double sustain = clampd(d.sustain, 0.0, 1.0);
    // 10^(-100 / (20 frames)) = 2^(-5 log2(10) / frames)
    const double decay_coef =
        exp_decay ? exp2_det(-5.0 * LOG2_10 / decay_frames) : 1.0 / decay_frames;
    const double release_coef =
        exp_release ? exp2_det(-5.0 * LOG2_10 / release_frames) : 1.0 / release_frames;
    const double fc = 8.176 * exp2_det(d.fc_cents / 1200.0);  // cents_to_hz
    const bool use_filter = d.filter != 0 && d.fc_cents < 13500.0 && fc < sr * 0.49 && fc > 20.0;
    double b0 = 1.0, b1 = 0.0, a1 = 0.0, a2 = 0.0;
    if (use_filter) {
        // q = max(10^(q_db / 20), 0.001) = 2^e; gain = sqrt(Q_DEFAULT / q) = 2^(-1/4 - e/2).
        const double q_db = d.q_cb / 10.0 - 3.01;
        const double e = fmax(q_db / 20.0 * LOG2_10, LOG2_MIN_Q);
        const double q = exp2_det(e);
        const double gain = exp2_det(-0.25 - 0.5 * e);
        const double w0 = TWO_PI * clampd(fc / sr, 1.0e-5, 0.49);
        double sw, cw;
        sincos_det(w0, sw, cw);
        const double alpha = sw / (2.0 * q);
        const double a0 = 1.0 + alpha;
        b0 = gain * (1.0 - cw) * 0.5 / a0;
        b1 = gain * (1.0 - cw) / a0;
        a1 = -2.0 * cw / a0;
        a2 = (1.0 - alpha) / a0;
    }
    kst_region_params p;
    p.attack_rate = (float)attack_rate;
    p.attack_end = (float)attack_end;
    p.decay_coef = (float)decay_coef;
    p.decay_target = fmaxf((float)sustain, env_floor);
    p.sustain = (float)sustain;
    p.release_coef = (float)release_coef;
    p.b0 = (float)b0;
    p.b1 = (float)b1;
    p.a1 = (float)a1;
    p.a2 = (float)a2;
    p.flags = use_filter ? RP_FILTER : 0u;
    p.pad = 0;
    return p;
}

__global__ void k_design(const kst_region_design *in, uint32_t n, kst_region_params *out,
                         double sr, uint32_t exp_decay, uint32_t exp_release, float env_floor) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = design_one(in[i], sr, exp_decay != 0, exp_release != 0, env_floor);
}

}  // namespace

void launch_design(const kst_region_design *in, uint32_t n, kst_region_params *out, double sr,
                   bool exp_decay, bool exp_release, float env_floor, hipStream_t s) {
    if (n) k_design<<<(n + 255) / 256, 256, 0, s>>>(in, n, out, sr, exp_decay, exp_release,
                                                     env_floor);
}

}  // namespace kst
