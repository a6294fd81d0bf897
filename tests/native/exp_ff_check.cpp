// exp_ff_check.cpp -- the single-precision exponential of the device code
// (indicatrix_amd/csrc/exp_ff.h, compiled here for the host) against the
// exponential evaluated in double precision and rounded once to f32, over
// evenly spread and random-bit arguments of the f32 range; prints the
// mismatch counts and the table as hex so the test can hold it to 2^(j/64).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../indicatrix_amd/csrc/exp_ff.h"

int main() {
    uint64_t s = 88172645463325252ull;
    long n = 0, bad_normal = 0, bad_subnormal = 0;
    for (long i = 0; i < 8000000; ++i) {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        const uint32_t u = (uint32_t)s;
        float x;
        if (i & 1) {
            x = -104.0f + 192.8f * (float)(u >> 8) / 16777216.0f;
        } else {
            std::memcpy(&x, &u, 4);
            if (!(x > -104.0f && x < 88.8f)) continue;
        }
        const float a = ixg_exp::exp_ff(x), b = (float)std::exp((double)x);
        ++n;
        if (std::memcmp(&a, &b, 4) != 0) {
            if (b < 1.17549435e-38f) ++bad_subnormal;
            else ++bad_normal;
        }
    }
    const float edge[] = {0.0f, -0.0f, 1.0f, -1.0f, 88.7228f, 88.72284f, -87.3365f, -103.97f, -104.0f, 89.0f};
    long bad_edge = 0;
    for (float x : edge) {
        const float a = ixg_exp::exp_ff(x), b = (float)std::exp((double)x);
        if (std::memcmp(&a, &b, 4) != 0 && b >= 1.17549435e-38f) ++bad_edge;
    }
    const float inf = ixg_exp::exp_ff(INFINITY), nan = ixg_exp::exp_ff(NAN), ninf = ixg_exp::exp_ff(-INFINITY);
    std::printf("{\"n\": %ld, \"bad_normal\": %ld, \"bad_subnormal\": %ld, \"bad_edge\": %ld, \"inf_ok\": %d, "
                "\"nan_ok\": %d, \"ninf_ok\": %d, \"hi\": [",
                n, bad_normal, bad_subnormal, bad_edge, std::isinf(inf) && inf > 0, std::isnan(nan), ninf == 0.0f);
    for (int j = 0; j < 64; ++j) std::printf("%s\"%a\"", j ? ", " : "", (double)ixg_exp::kExp2Hi[j]);
    std::printf("], \"lo\": [");
    for (int j = 0; j < 64; ++j) std::printf("%s\"%a\"", j ? ", " : "", (double)ixg_exp::kExp2Lo[j]);
    std::printf("]}\n");
    return 0;
}
