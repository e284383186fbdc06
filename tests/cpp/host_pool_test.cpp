// Host test of wenox::HostPool (csrc/wenox_hostpool.h): jobs of every size
// and thread count, run synchronously and started while the caller works,
// each range visited exactly once; built and run by test_host_pool_cpu.py
// (under ThreadSanitizer where the compiler has it).
#include <cstdio>

#include "wenox_hostpool.h"

int main()
{
    wenox::HostPool* pool = new wenox::HostPool;   // kept, as the library keeps its pool
    std::vector<int64_t> v(1 << 12), w(1 << 12);
    for (int rep = 0; rep < 1500; ++rep) {
        const int64_t M = 1 + (rep * 37) % 997;
        const int T = 1 + rep % (wenox::HostPool::MAX_THREADS + 2);
        std::fill(v.begin(), v.begin() + M, 0);
        pool->start(M, T, [&](int64_t i0, int64_t i1) {
            for (int64_t i = i0; i < i1; ++i) v[i] += i + 1;
        });
        for (int64_t i = 0; i < M; ++i) w[i] = 2 * i;   // the caller's own work meanwhile
        pool->wait();
        for (int64_t i = 0; i < M; ++i)
            if (v[i] != i + 1 || w[i] != 2 * i) {
                std::printf("FAIL start/wait rep %d index %lld\n", rep, (long long)i);
                return 1;
            }
        std::fill(v.begin(), v.begin() + M, 0);
        pool->run(M, T, [&](int64_t i0, int64_t i1) {
            for (int64_t i = i0; i < i1; ++i) v[i] += 3;
        });
        for (int64_t i = 0; i < M; ++i)
            if (v[i] != 3) {
                std::printf("FAIL run rep %d index %lld\n", rep, (long long)i);
                return 1;
            }
    }
    std::printf("ok\n");
    return 0;
}
