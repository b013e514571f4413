// Host-sized pieces of the stage-2 loop and prep (see ps2_host_util.hpp).
#include "ps2_host_util.hpp"

#include <algorithm>

namespace ps2 {

MatlabV5Uniform::MatlabV5Uniform(uint32_t seed) {
    j_ = seed == 0 ? (1u << 31) : seed;
    uint32_t jj = j_;
    const double ulp = 1.1102230246251565e-16;  // 2^-53
    for (int k = 0; k < 32; ++k) {
        uint64_t x = 0;
        for (int b = 0; b < 53; ++b) {
            jj = xorshift(jj);
            x = (x << 1) | (uint64_t)((jj >> 19) & 1u);
        }
        state_[k] = (double)x * ulp;
    }
}

double MatlabV5Uniform::next() {
    const double ulp = 1.1102230246251565e-16;
    double v = state_[(index_ + 20) & 31] - state_[(index_ + 5) & 31] - borrow_;
    if (v < 0.0) {
        v += 1.0;
        borrow_ = ulp;
    } else {
        borrow_ = 0.0;
    }
    state_[index_] = v;
    index_ = (index_ + 1) & 31;
    // xor the 52 low mantissa bits with a xorshift word pair
    const uint32_t jlo = j_;
    const uint32_t jhi = xorshift(jlo);
    j_ = jhi;
    const uint64_t m52 = (1ull << 52) - 1;
    const uint64_t mask = (((uint64_t)jhi << 32) & m52) ^ (uint64_t)jlo;
    if (v == 0.0) return std::ldexp((double)(0ull ^ mask), -53);
    uint64_t bits;
    std::memcpy(&bits, &v, 8);
    const int eb = (int)((bits >> 52) & 0x7ff);
    const uint64_t mb = bits & m52;
    uint64_t mant;
    int ex;
    if (eb == 0) {
        const int lead = 63 - __builtin_clzll(mb);
        ex = lead - 1073;
        mant = mb << (52 - lead);
    } else {
        ex = eb - 1022;
        mant = (1ull << 52) | mb;
    }
    return std::ldexp((double)(mant ^ mask), ex - 53);
}

}  // namespace ps2
