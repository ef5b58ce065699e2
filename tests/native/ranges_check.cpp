// Randomised check of dtb::RangeSet against a per-unit occupancy model:
// every take returns a free, in-bounds range of the requested length that
// is the smallest free run able to hold it; gives merge back so that after
// everything is returned the set is one range again.  Prints "ok".
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "dtb_ranges.hpp"

static int fail(const char *m)
{
    std::printf("FAIL %s\n", m);
    return 1;
}

int main()
{
    const size_t N = 4096;
    std::mt19937_64 rng(12345);
    for (int round = 0; round < 50; ++round) {
        dtb::RangeSet rs(N);
        std::vector<char> used(N, 0);
        std::vector<std::pair<size_t, size_t>> held;
        for (int op = 0; op < 2000; ++op) {
            if (held.empty() || rng() % 3 != 0) {
                const size_t len = 1 + rng() % (rng() % 4 == 0 ? 512 : 32);
                // smallest free run >= len in the model
                size_t best = 0;
                for (size_t i = 0; i < N;) {
                    if (used[i]) { ++i; continue; }
                    size_t j = i;
                    while (j < N && !used[j]) ++j;
                    if (j - i >= len && (best == 0 || j - i < best)) best = j - i;
                    i = j;
                }
                size_t off = 0;
                const bool ok = rs.take(len, &off);
                if (ok != (best != 0)) return fail("take success differs from model");
                if (rs.best(1) == 0 && rs.free_bytes() != 0) return fail("free bytes without a range");
                if (!ok) continue;
                if (off + len > N) return fail("out of bounds");
                size_t run = 0;
                for (size_t i = off; i < N && !used[i]; ++i) ++run;
                if (off > 0 && !used[off - 1]) return fail("not at the front of a free run");
                if (run != best) return fail("not the best fit");
                for (size_t i = off; i < off + len; ++i) used[i] = 1;
                held.emplace_back(off, len);
            } else {
                const size_t k = rng() % held.size();
                for (size_t i = held[k].first; i < held[k].first + held[k].second; ++i) used[i] = 0;
                rs.give(held[k].first, held[k].second);
                held[k] = held.back();
                held.pop_back();
            }
            size_t nfree = 0;
            for (char u : used) nfree += !u;
            if (nfree != rs.free_bytes()) return fail("free byte count");
        }
        for (auto &h : held) rs.give(h.first, h.second);
        if (!rs.all_free() || rs.best(N) != N) return fail("not merged back into one range");
    }
    std::printf("ok\n");
    return 0;
}
