// CPU checks of the lease table (vapoursynth-zipcu_amd/csrc/slot_cache.h), the bookkeeping of the
// engine's source-frame and accumulator caches: deterministic cases, then a multithreaded stress
// run with fills that fail at random. Exits non-zero on the first violated property.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

#include "slot_cache.h"

#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::fprintf(stderr, "CHECK failed: %s (line %d)\n", #c, __LINE__); \
            std::exit(1);                                                    \
        }                                                                    \
    } while (0)

static void deterministic() {
    LeaseTable t;
    t.reset(4);
    // duplicates share one buffer and one fill duty
    {
        const int64_t k[5] = {7, 7, 8, 7, 9};
        auto L = t.lease(k, 5);
        CHECK(L.held());
        CHECK(L.slot(0) == L.slot(1) && L.slot(0) == L.slot(3));
        CHECK(L.slot(0) != L.slot(2) && L.slot(2) != L.slot(4) && L.slot(0) != L.slot(4));
        CHECK(L.must_fill(0) && !L.must_fill(1) && L.must_fill(2) && !L.must_fill(3) && L.must_fill(4));
        L.filled(0);
        L.filled(2);
        // key 9 is never filled: ending the lease forgets it
    }
    {
        const int64_t k[3] = {7, 8, 9};
        auto L = t.lease(k, 3);
        CHECK(!L.must_fill(0) && !L.must_fill(1) && L.must_fill(2));
        L.filled(2);
    }
    // LRU: keys 7, 8, 9 resident, one empty buffer; touch 7, then two new keys take the empty
    // buffer and the least recently used one (8)
    int slot8 = -1;
    {
        const int64_t k[1] = {8};
        auto L = t.lease(k, 1);
        slot8 = L.slot(0);
    }
    {
        const int64_t k[2] = {7, 9};
        auto L = t.lease(k, 2);
        CHECK(!L.must_fill(0) && !L.must_fill(1));
    }
    {
        const int64_t k[2] = {10, 11};
        auto L = t.lease(k, 2);
        CHECK(L.must_fill(0) && L.must_fill(1));
        CHECK(L.slot(0) == slot8 || L.slot(1) == slot8);
        L.filled(0);
        L.filled(1);
    }
    {
        const int64_t k[1] = {8};
        auto L = t.lease(k, 1);
        CHECK(L.must_fill(0));  // evicted
    }
    // more distinct keys than buffers: never grantable, returned at once
    {
        const int64_t k[5] = {1, 2, 3, 4, 5};
        auto L = t.lease(k, 5);
        CHECK(!L.held());
    }
    // a lease moved into another variable ends once
    {
        const int64_t k[1] = {42};
        LeaseTable::Lease a = t.lease(k, 1);
        LeaseTable::Lease b = std::move(a);
        CHECK(!a.held() && b.held());
        b.filled(0);
        b.end();
        b.end();
        auto c = t.lease(k, 1);
        CHECK(!c.must_fill(0));
    }
}

static void stress(int threads, int window, int iters, unsigned seed) {
    const int capacity = window + threads;
    LeaseTable t;
    t.reset(capacity);
    std::vector<std::atomic<int64_t>> content(capacity);
    for (auto &c : content) c = -1;
    std::vector<std::atomic<int>> fillers(1024);
    for (auto &f : fillers) f = 0;
    std::atomic<long> fills{0}, failures{0}, hits{0};
    auto worker = [&](int id) {
        std::mt19937 rng(seed + 977u * unsigned(id));
        for (int it = 0; it < iters; ++it) {
            // a clamped temporal window around a random frame: duplicates at the clip edges
            const int nf = 48, n = int(rng() % nf), r = window / 2;
            std::vector<int64_t> keys(window);
            for (int j = 0; j < window; ++j) keys[j] = std::min(std::max(n - r + j, 0), nf - 1);
            auto L = t.lease(keys.data(), window);
            CHECK(L.held());
            bool failed = false;
            for (int j = 0; j < window; ++j) {
                if (!L.must_fill(j)) continue;
                CHECK(fillers[keys[j]].fetch_add(1) == 0);  // one filler per key at a time
                content[L.slot(j)] = -2;                     // being written
                std::this_thread::sleep_for(std::chrono::microseconds(rng() % 50));
                if (rng() % 10 == 0) {                       // the fill fails: abandon
                    fillers[keys[j]].fetch_sub(1);
                    failures++;
                    failed = true;
                    break;
                }
                content[L.slot(j)] = keys[j];
                fillers[keys[j]].fetch_sub(1);
                L.filled(j);
                fills++;
            }
            if (!failed)
                for (int j = 0; j < window; ++j) {
                    CHECK(content[L.slot(j)] == keys[j]);  // resident buffers hold their key
                    if (!L.must_fill(j)) hits++;
                }
            std::this_thread::sleep_for(std::chrono::microseconds(rng() % 30));
            L.end();
        }
    };
    std::vector<std::thread> pool;
    for (int i = 0; i < threads; ++i) pool.emplace_back(worker, i);
    for (auto &th : pool) th.join();
    std::printf("stress threads=%d window=%d: fills %ld failures %ld shared %ld\n", threads, window, fills.load(),
                failures.load(), hits.load());
    CHECK(fills > 0 && failures > 0 && hits > 0);
}

int main() {
    deterministic();
    stress(4, 5, 400, 1);
    stress(8, 9, 300, 2);
    stress(16, 3, 500, 3);
    std::printf("lease table: ok\n");
    return 0;
}
