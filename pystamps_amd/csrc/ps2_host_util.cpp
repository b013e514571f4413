// Host-sized pieces of the stage-2 loop and prep (see ps2_host_util.hpp).
#include "ps2_host_util.hpp"

#include <algorithm>

namespace ps2 {

MatlabV5Uniform::MatlabV5Uniform(uint32_t seed) {
    j_ = seed == 0 ? (1u << 31) : seed;
    uint32_t jj = j_;
    for (int k = 0; k < 32; ++k) {
        uint64_t x = 0;
        for (int b = 0; b < 53; ++b) {
            jj = xorshift(jj);
            x = (x << 1) | (uint64_t)((jj >> 19) & 1u);
        }
        state_[k] = x;
    }
}

}  // namespace ps2
