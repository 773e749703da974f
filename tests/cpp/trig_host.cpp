// Host build of the device sine/cosine (csrc/hip/libm_sincosf.h) next to the C
// library it restates, for tests/test_trig.py.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <thread>
#include <vector>

#include "libm_sincosf.h"

extern "C" {

// The library's and the restatement's values for n floats.
void trig_host_eval(const float* x, size_t n, float* s_lib, float* c_lib, float* s_own,
                    float* c_own) {
    for (size_t i = 0; i < n; ++i) {
        sincosf(x[i], s_lib + i, c_lib + i);
        att::sincosf_libm(x[i], s_own + i, c_own + i);
    }
}

// Counts the floats (both signs of every bit pattern lo, lo + stride, ... < hi,
// magnitudes only) where the restatement differs from the library's sincosf,
// sinf or cosf in any bit; NaN results compare by being NaN.  `threads` workers.
uint64_t trig_host_mismatches(uint32_t lo, uint32_t hi, uint32_t stride, int threads) {
    if (threads < 1) threads = 1;
    std::vector<uint64_t> bad(threads, 0);
    std::vector<std::thread> pool;
    auto same = [](float a, float b) {
        uint32_t ua, ub;
        memcpy(&ua, &a, 4);
        memcpy(&ub, &b, 4);
        return ua == ub || (a != a && b != b);
    };
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            uint64_t nb = 0;
            for (uint64_t u = lo + (uint64_t)t * stride; u < hi; u += (uint64_t)threads * stride) {
                for (uint32_t sg = 0; sg < 2; ++sg) {
                    const uint32_t bits = (uint32_t)u | (sg << 31);
                    float x;
                    memcpy(&x, &bits, 4);
                    float s, c, rs, rc;
                    sincosf(x, &rs, &rc);
                    att::sincosf_libm(x, &s, &c);
                    if (!same(s, rs) || !same(c, rc) || !same(sinf(x), rs) || !same(cosf(x), rc)) ++nb;
                }
            }
            bad[t] = nb;
        });
    for (auto& th : pool) th.join();
    uint64_t tot = 0;
    for (auto b : bad) tot += b;
    return tot;
}

}  // extern "C"
