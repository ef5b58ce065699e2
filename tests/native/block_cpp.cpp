// A C++ node driving the step through dtb::Block (include/dtb_block.hpp):
// generates the seeded 80/20 network of dtb_gen_uniform, sets the OU
// background of every population, runs `steps` steps with per-population
// counts, and prints "counts <step> <events> <spikes> <sum of freq> <xor of
// V bits>".  Without a GPU it prints "no device <status>".  Arguments: n, k,
// steps, propagation layout (DTB_LAYOUT_*, default auto).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dtb_block.hpp"

int main(int argc, char **argv)
{
    const int64_t n = argc > 1 ? std::atoll(argv[1]) : 3000;
    const uint32_t k = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 100;
    const int32_t steps = argc > 3 ? std::atoi(argv[3]) : 200;
    const int32_t layout = argc > 4 ? std::atoi(argv[4]) : DTB_LAYOUT_AUTO;
    const float g[4] = {0.01f, 0.0005f, 0.1f, 0.005f};
    std::vector<float> props((size_t)n * 22), w((size_t)n * k * 2);
    std::vector<int64_t> rowptr((size_t)n + 1);
    std::vector<uint32_t> col((size_t)n * k), kinds((size_t)(n * k + 31) / 32 + 1);
    dtb::check(dtb_gen_uniform(n, k, 1, g, 0, props.data(), rowptr.data(), col.data(), w.data(), kinds.data()));
    try {
        dtb::Block blk(0, 0.1f, 11);
        blk.set_schedule(2, layout);
        const dtb_network net{n, props.data(), n * (int64_t)k, rowptr.data(), col.data(), w.data(), kinds.data()};
        blk.from_network(net);
        const unsigned int npop = blk.get_total_subblocks();
        std::vector<unsigned int> ids(blk.get_sub_bids_cpu(), blk.get_sub_bids_cpu() + npop);
        blk.set_I_ou_current_param(ids, std::vector<float>(npop, 1.5f), std::vector<float>(npop, 0.1f),
                                   std::vector<float>(npop, 4.0f));
        std::vector<uint32_t> freq((size_t)steps * npop);
        dtb_run_outputs out;
        std::memset(&out, 0, sizeof(out));
        out.freq = freq.data();
        blk.run(steps, 1, DTB_RUN_USE_OU | DTB_RUN_FREQ, out);
        int64_t step = 0;
        uint64_t events = 0, spikes = 0;
        blk.counters(step, events, spikes);
        std::vector<float> v;
        std::vector<int32_t> t;
        std::vector<uint8_t> f;
        blk.get_state(v, t, f);
        uint32_t h = 0;
        for (float x : v) {
            uint32_t b;
            std::memcpy(&b, &x, 4);
            h ^= b;
        }
        unsigned long long total = 0;
        for (uint32_t c : freq)
            total += c;
        std::printf("counts %lld %llu %llu %llu %u\n", (long long)step, (unsigned long long)events,
                    (unsigned long long)spikes, total, h);
    } catch (const dtb::BlockError &e) {
        if (e.status == DTB_NO_DEVICE) {
            std::printf("no device %d\n", e.status);
            return 0;
        }
        std::printf("error %d %s\n", e.status, e.what());
        return 1;
    }
    return 0;
}
