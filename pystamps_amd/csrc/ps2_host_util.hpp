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
class MatlabV5Uniform {
  public:
    explicit MatlabV5Uniform(uint32_t seed);
    double next();

  private:
    static uint32_t xorshift(uint32_t v) {
        v ^= v << 13;
        v ^= v >> 17;
        v ^= v << 5;
        return v;
    }
    int index_ = 0;
    double borrow_ = 0.0;
    uint32_t j_ = 0;
    double state_[32];
};

}  // namespace ps2
