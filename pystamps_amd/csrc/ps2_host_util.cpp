// Host-sized pieces of the stage-2 loop and prep (see ps2_host_util.hpp).
#include "ps2_host_util.hpp"

#include <algorithm>

namespace ps2 {

namespace {

// MATLAB interp(x, 10) FIR, the fixed 79 taps the reference audits
// (native_stage2.rs:33-113); the filter delay is 39 samples.
const double kInterp10Taps[79] = {
    -7.92925708923623813e-05, -1.73293349766172333e-04, -3.00248431516961407e-04,
    -4.48605550383722462e-04, -5.95835452295412349e-04, -7.08759204842826507e-04,
    -7.46231412934915406e-04, -6.64375214040318490e-04, -4.24121369705182660e-04,
    -3.01053961794528480e-07, 6.08939001185030554e-04,  1.37399755086099230e-03,
    2.22872001918483081e-03,  3.06999955141492101e-03,  3.76333143346069959e-03,
    4.15488007985555935e-03,  4.08970016389962471e-03,  3.43472424182160452e-03,
    2.10413493213892910e-03,  8.39421246326618742e-05,  -2.54786253685349538e-03,
    -5.60918261221970971e-03, -8.81278723412589476e-03, -1.17783753398490003e-02,
    -1.40582499688596801e-02, -1.51753768689126753e-02, -1.46708218626883501e-02,
    -1.21559904304523667e-02, -7.36396963079587880e-03, -1.93792008491816548e-04,
    9.25827682864269544e-03,  2.06852593184593814e-02,  3.35768644427044682e-02,
    4.72485589714582085e-02,  6.08905923193367796e-02,  7.36328154139928304e-02,
    8.46190871379102622e-02,  9.30836882813813049e-02,  9.84216619119140101e-02,
    1.00245459871706619e-01,  9.84216619119134828e-02,  9.30836882813811800e-02,
    8.46190871379124271e-02,  7.36328154139975766e-02,  6.08905923193391457e-02,
    4.72485589714637388e-02,  3.35768644427109145e-02,  2.06852593184574732e-02,
    9.25827682864464527e-03,  -1.93792008489998774e-04, -7.36396963079247527e-03,
    -1.21559904304541604e-02, -1.46708218626876996e-02, -1.51753768689126579e-02,
    -1.40582499688608267e-02, -1.17783753398508096e-02, -8.81278723413165405e-03,
    -5.60918261221661496e-03, -2.54786253685826761e-03, 8.39421246353382272e-05,
    2.10413493213411785e-03,  3.43472424182329154e-03,  4.08970016390024574e-03,
    4.15488007985538154e-03,  3.76333143346146634e-03,  3.06999955141576842e-03,
    2.22872001918783102e-03,  1.37399755086213202e-03,  6.08939001188346478e-04,
    -3.01053965413812172e-07, -4.24121369703597340e-04, -6.64375214039929912e-04,
    -7.46231412935480926e-04, -7.08759204842192465e-04, -5.95835452295058465e-04,
    -4.48605550383040716e-04, -3.00248431517695195e-04, -1.73293349768269614e-04,
    -7.92925708931989381e-05,
};

// y[n] = sum_k h[k] x[n-k], summed in k order (scipy/MATLAB filter(b, 1, x))
std::vector<double> fir(const double *h, size_t nh, const std::vector<double> &x) {
    std::vector<double> y(x.size(), 0.0);
    for (size_t n = 0; n < x.size(); ++n) {
        double s = 0.0;
        for (size_t k = 0; k < nh && k <= n; ++k) s += h[k] * x[n - k];
        y[n] = s;
    }
    return y;
}

}  // namespace

std::vector<double> psquare_prand_hi(const std::vector<double> &nr, const std::vector<double> &na,
                                     int64_t low_coh_thresh, double nr_max_nz_ix) {
    const size_t nb = nr.size();
    std::vector<double> prand(nb);
    for (size_t k = 0; k < nb; ++k) {
        const double q = nr[k] / (na[k] == 0.0 ? 1.0 : na[k]);
        prand[k] = (std::isnan(q) || q > 1.0) ? 1.0 : q;
    }
    const size_t lct = (size_t)std::max<int64_t>(0, std::min<int64_t>(low_coh_thresh, (int64_t)nb));
    for (size_t k = 0; k < lct; ++k) prand[k] = 1.0;
    size_t zero_from = 0;
    if (nr_max_nz_ix > 0.0) zero_from = (size_t)nr_max_nz_ix;
    for (size_t k = std::min(zero_from, nb); k < nb; ++k) prand[k] = 0.0;

    // 7-point Gaussian window (gausswin(7, 2.5)) run over [1 x7, prand]
    double g[7], gsum = 0.0;
    const double sd = 6.0 / 5.0;
    for (int k = 0; k < 7; ++k) {
        const double x = (double)k - 3.0;
        g[k] = std::exp(-0.5 * (x / sd) * (x / sd));
    }
    for (int k = 0; k < 7; ++k) gsum += g[k];
    std::vector<double> padded(7, 1.0);
    padded.insert(padded.end(), prand.begin(), prand.end());
    const std::vector<double> sm = fir(g, 7, padded);
    for (size_t k = 0; k < nb; ++k) prand[k] = sm[k + 7] / gsum;

    // interp([1, prand], 10): upsample by 10, 79-tap FIR, drop the 39-sample
    // delay, scale by 10, keep all but the last 9 samples
    const size_t nin = nb + 1, factor = 10, delay = 39;
    std::vector<double> up(nin * factor + delay, 0.0);
    up[0] = 1.0;
    for (size_t k = 0; k < nb; ++k) up[(k + 1) * factor] = prand[k];
    const std::vector<double> fy = fir(kInterp10Taps, 79, up);
    const size_t nhi = nin * factor;
    const size_t keep = nhi >= 9 ? nhi - 9 : 0;
    std::vector<double> hi(keep);
    for (size_t k = 0; k < keep; ++k) hi[k] = fy[delay + k] * (double)factor;
    return hi;
}

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
