// Host-side helpers of the stage-2 library: error reporting and the
// host-sized pieces of the prep (histogram rule, MATLAB v5 uniform generator
// for the random-coherence histogram).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace ps2 {

void set_error(const std::string &msg);

// ---------------------------------------------------------------------------
// hist_with_centers rule (native_stage2.rs:1581-1640)
struct HistRule {
    bool equal = true;
    double cutoff0 = 0.0, d = 1.0;
    std::vector<double> mids;
};

inline HistRule make_hist_rule(const double *c, int64_t n) {
    HistRule h;
    if (n < 2) return h;
    const double ref = c[1] - c[0];
    double max_center = 0.0;
    for (int64_t k = 0; k < n; ++k) max_center = std::fmax(max_center, std::fabs(c[k]));
    const double tol = 2.220446049250313e-16 * (max_center > 1.0 ? max_center : 1.0);
    for (int64_t k = 0; k + 1 < n; ++k)
        if (!(std::fabs((c[k + 1] - c[k]) - ref) <= tol)) h.equal = false;
    h.d = n < 3 ? 1.0 : (c[n - 1] - c[0]) / (double)(n - 1);
    h.cutoff0 = (c[0] + c[1]) / 2.0;
    for (int64_t k = 0; k + 1 < n; ++k) h.mids.push_back((c[k] + c[k + 1]) / 2.0);
    return h;
}

// ---------------------------------------------------------------------------
// MATLAB v5 uniform generator ('state' method; native_stage2.rs:1785-1873).
// The 32 lagged values are multiples of 2^-53 in [0, 1); they are held here as
// 53-bit integers, so the subtract-with-borrow step is integer arithmetic and
// the output double is assembled from its exponent and mantissa bits.
class MatlabV5Uniform {
  public:
    explicit MatlabV5Uniform(uint32_t seed);
    double next() {
        constexpr uint64_t kOne = 1ull << 53, m52 = (1ull << 52) - 1;
        // x_n = x_{n-12} - x_{n-27} - borrow, modulo 2^53
        const uint64_t d = state_[(index_ + 20) & 31] - state_[(index_ + 5) & 31] - borrow_;
        borrow_ = d >> 63;
        const uint64_t v = d + (borrow_ ? kOne : 0);
        state_[index_] = v;
        index_ = (index_ + 1) & 31;
        // xor the 52 low mantissa bits with a xorshift word pair
        const uint32_t jlo = j_;
        const uint32_t jhi = xorshift(jlo);
        j_ = jhi;
        const uint64_t mask = (((uint64_t)jhi << 32) & m52) ^ (uint64_t)jlo;
        if (v == 0) return (double)mask * 0x1p-53;
        // v * 2^-53 = mant * 2^(lead - 105) with mant's bit 52 set; the xor
        // keeps bit 52, so the result has biased exponent lead + 970
        const int lead = 63 - __builtin_clzll(v);
        const uint64_t mant = v << (52 - lead);
        const uint64_t bits = ((uint64_t)(lead + 970) << 52) | ((mant ^ mask) & m52);
        double out;
        std::memcpy(&out, &bits, 8);
        return out;
    }

  private:
    static uint32_t xorshift(uint32_t v) {
        v ^= v << 13;
        v ^= v >> 17;
        v ^= v << 5;
        return v;
    }
    int index_ = 0;
    uint64_t borrow_ = 0;
    uint32_t j_ = 0;
    uint64_t state_[32];
};

}  // namespace ps2
