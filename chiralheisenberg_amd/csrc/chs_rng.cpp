// chs_rng.cpp -- host random numbers with the reference's exact stream
// (rhmc/cheisdwf.f:1678-1801): a 48-bit multiplicative LCG
// s <- s * 44485709377909 mod 2^48 (the multiplier 2651554*2^24 + 15184245,
// start state 2868876*2^24 + 16651885 unless set), uniforms s / 2^48 in FP64,
// a 97-entry Bays-Durham shuffle whose table is shared by the two pointer
// streams yran (heatbath, measurement) and yrun (trajectory length,
// Metropolis), and Box-Muller Gaussians that draw all radii first, then all
// angles.  Integer arithmetic makes the stream exact on every host.
#include "../../include/chs.h"

#include <cmath>
#include <cstring>

namespace {
constexpr uint64_t kMul = 2651554ULL * 16777216ULL + 15184245ULL;
constexpr uint64_t kMask = (1ULL << 48) - 1;
constexpr double kTwoPi = 6.283185307179586476925286766559;

double draw(chs_rng *r) {
    r->s = (r->s * kMul) & kMask;   // 64-bit wrap keeps the low 48 bits exact
    return (double)r->s * (1.0 / 281474976710656.0);
}
}  // namespace

extern "C" {

void chs_rng_set(chs_rng *r, double seed) {
    r->s = ((uint64_t)seed) & kMask;
    r->idum = -1;
    r->yran = 0.0;
    r->yrun = 0.0;   // the reference never initialises yrun; it starts at 0
    std::memset(r->v, 0, sizeof r->v);
}

double chs_rng_get(const chs_rng *r) { return (double)r->s; }

double chs_rano(chs_rng *r, int stream) {
    double &y = stream ? r->yrun : r->yran;
    if (r->idum < 0) {
        r->idum = 1;
        for (int j = 0; j < 97; j++) (void)draw(r);
        for (int j = 0; j < 97; j++) r->v[j] = draw(r);
        y = draw(r);
    }
    int j = 1 + (int)(97.0 * y);
    j = j > 97 ? 97 : (j < 1 ? 1 : j);
    y = r->v[j - 1];
    const double out = y;
    r->v[j - 1] = draw(r);
    return out;
}

void chs_gauss0(chs_rng *r, double *ps, int kvol) {
    for (int i = 0; i < kvol; i++) ps[i + kvol] = std::sqrt(-std::log(chs_rano(r, 0)));
    for (int i = 0; i < kvol; i++) {
        const double th = kTwoPi * chs_rano(r, 0);
        ps[i] = ps[i + kvol] * std::sin(th);
        ps[i + kvol] = ps[i + kvol] * std::cos(th);
    }
}

void chs_gaussp(chs_rng *r, double *ps, int kvol) {
    for (int i = 0; i < kvol; i++) ps[i + kvol] = std::sqrt(-2.0 * std::log(chs_rano(r, 0)));
    for (int i = 0; i < kvol; i++) {
        const double th = kTwoPi * chs_rano(r, 0);
        ps[i] = ps[i + kvol] * std::sin(th);
        ps[i + kvol] = ps[i + kvol] * std::cos(th);
    }
}

}  // extern "C"
