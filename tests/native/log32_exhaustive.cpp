// Host build of ath::neg_log_f32 (algotom_amd/csrc/log32.h) checked against
// (float)(-log((double)x)) from the C library for every positive finite
// float32 (bit patterns 1 .. 0x7f7fffff), split over hardware threads.
// Prints the counts; exit status 1 on any mismatch.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "log32.h"

int main() {
    const unsigned long long first = 1, end = 0x7f800000ull;
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 4;
    std::vector<unsigned long long> bad(nt, 0);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t) {
        pool.emplace_back([&, t] {
            const unsigned long long lo = first + (end - first) * t / nt;
            const unsigned long long hi = first + (end - first) * (t + 1) / nt;
            for (unsigned long long i = lo; i < hi; ++i) {
                float x;
                const unsigned u = (unsigned)i;
                std::memcpy(&x, &u, 4);
                const float a = ath::neg_log_f32(x);
                const float r = (float)(-std::log((double)x));
                if (std::memcmp(&a, &r, 4) != 0) ++bad[t];
            }
        });
    }
    unsigned long long total = 0;
    for (unsigned t = 0; t < nt; ++t) {
        pool[t].join();
        total += bad[t];
    }
    std::printf("checked %llu values, mismatches %llu\n", end - first, total);
    return total != 0;
}
