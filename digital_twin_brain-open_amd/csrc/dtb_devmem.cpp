// dtb_devmem.cpp -- device memory for the engine's buffers.
//
// hipMalloc and hipFree cost time that grows with the block size (page
// mapping: seconds for tens of GB), which dominates when instances of that
// size come and go every few hundred milliseconds (serving).  So freed memory
// stays with the library:
//   * blocks of 1 MB and more are ranges of segments (device allocations),
//     taken best-fit and merged with free neighbours when released; a
//     segment is either reserved up front (reserve) or made for a request
//     that no free range holds;
//   * smaller blocks are kept by size and handed out again.
// A released block may still be read by work in flight, so it is parked
// first; parked blocks become reusable after one device synchronisation,
// taken only when a request finds no free memory.  When the device runs out
// of memory, segments with nothing allocated and cached small blocks are
// returned to the runtime and the allocation is retried.
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "dtb_engine.hpp"
#include "dtb_ranges.hpp"

namespace dtb {
namespace devmem {
namespace {

constexpr size_t kLarge = (size_t)1 << 20;
constexpr size_t kGrain = (size_t)2 << 20;    // large blocks: multiples of 2 MB
constexpr size_t kSmallGrain = 512;

struct Segment {
    char *base = nullptr;
    RangeSet ranges;
    bool reserved = false;
};

struct Block {
    int device;
    size_t bytes;
    int seg;           // segment index, -1 for a small block
    size_t off;
};

struct DeviceCache {
    std::vector<std::unique_ptr<Segment>> segs;   // null once returned
    std::multimap<size_t, void *> small;           // reusable small blocks by size
    std::vector<std::pair<void *, Block>> parked;  // released, maybe still in use
    size_t cached = 0;                              // bytes free or parked
};

struct State {
    std::mutex mu;
    std::unordered_map<int, DeviceCache> dev;
    std::unordered_map<void *, Block> live;
};

// Never destroyed: buffers may be released from static destructors.
State &state()
{
    static State *s = new State;
    return *s;
}

size_t round_size(size_t bytes)
{
    if (bytes < kLarge)
        return (bytes + kSmallGrain - 1) / kSmallGrain * kSmallGrain;
    return (bytes + kGrain - 1) / kGrain * kGrain;
}

void *take_large(State &st, DeviceCache &c, int device, size_t want)
{
    int pick = -1;
    size_t best = 0;
    for (size_t i = 0; i < c.segs.size(); ++i) {
        if (!c.segs[i])
            continue;
        const size_t b = c.segs[i]->ranges.best(want);
        if (b != 0 && (pick < 0 || b < best)) {
            pick = (int)i;
            best = b;
        }
    }
    if (pick < 0)
        return nullptr;
    size_t off = 0;
    c.segs[(size_t)pick]->ranges.take(want, &off);
    void *p = c.segs[(size_t)pick]->base + off;
    st.live[p] = Block{device, want, pick, off};
    c.cached -= want;
    return p;
}

void *take_small(State &st, DeviceCache &c, int device, size_t want)
{
    auto it = c.small.find(want);
    if (it == c.small.end())
        return nullptr;
    void *p = it->second;
    c.small.erase(it);
    st.live[p] = Block{device, want, -1, 0};
    c.cached -= want;
    return p;
}

void unpark(DeviceCache &c)
{
    for (auto &pb : c.parked) {
        const Block &b = pb.second;
        if (b.seg < 0)
            c.small.emplace(b.bytes, pb.first);
        else
            c.segs[(size_t)b.seg]->ranges.give(b.off, b.bytes);
    }
    c.parked.clear();
}

// Returns unused memory to the runtime (all of it with `reserved`, else the
// segments that were not reserved).  The device is current.
void drop_unused(DeviceCache &c, bool reserved)
{
    if (!c.parked.empty()) {
        (void)hipDeviceSynchronize();
        unpark(c);
    }
    for (auto &kv : c.small) {
        (void)hipFree(kv.second);
        c.cached -= kv.first;
    }
    c.small.clear();
    for (auto &s : c.segs) {
        if (s && s->ranges.all_free() && (reserved || !s->reserved)) {
            (void)hipFree(s->base);
            c.cached -= s->ranges.size();
            s.reset();
        }
    }
}

void *new_segment(DeviceCache &c, size_t bytes, bool reserved)
{
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
        (void)hipGetLastError();
        drop_unused(c, false);
        e = hipMalloc(&p, bytes);
    }
    DTB_HIP(e);
    auto s = std::make_unique<Segment>();
    s->base = static_cast<char *>(p);
    s->ranges = RangeSet(bytes);
    s->reserved = reserved;
    c.segs.push_back(std::move(s));
    c.cached += bytes;
    return p;
}

}  // namespace

void *alloc(size_t bytes)
{
    int device = 0;
    DTB_HIP(hipGetDevice(&device));
    const size_t want = round_size(bytes);
    const bool large = want >= kLarge;
    State &st = state();
    std::lock_guard<std::mutex> lock(st.mu);
    DeviceCache &c = st.dev[device];
    auto take = [&]() { return large ? take_large(st, c, device, want) : take_small(st, c, device, want); };
    void *p = take();
    if (p == nullptr && !c.parked.empty()) {
        DTB_HIP(hipDeviceSynchronize());
        unpark(c);
        p = take();
    }
    if (p != nullptr)
        return p;
    if (large) {
        new_segment(c, want, false);
        return take_large(st, c, device, want);
    }
    hipError_t e = hipMalloc(&p, want);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
        (void)hipGetLastError();
        drop_unused(c, false);
        e = hipMalloc(&p, want);
    }
    DTB_HIP(e);
    st.live[p] = Block{device, want, -1, 0};
    return p;
}

void release(void *p)
{
    if (p == nullptr)
        return;
    State &st = state();
    std::lock_guard<std::mutex> lock(st.mu);
    auto it = st.live.find(p);
    if (it == st.live.end())
        return;
    DeviceCache &c = st.dev[it->second.device];
    c.parked.emplace_back(p, it->second);
    c.cached += it->second.bytes;
    st.live.erase(it);
}

void reserve(size_t bytes)
{
    int device = 0;
    DTB_HIP(hipGetDevice(&device));
    State &st = state();
    std::lock_guard<std::mutex> lock(st.mu);
    DeviceCache &c = st.dev[device];
    size_t have = 0;
    for (auto &s : c.segs)
        if (s && s->reserved)
            have += s->ranges.size();
    const size_t want = round_size(std::max(bytes, kLarge));
    if (have < want)
        new_segment(c, want - have < kLarge ? kLarge : round_size(want - have), true);
}

void trim()
{
    State &st = state();
    std::lock_guard<std::mutex> lock(st.mu);
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (auto &kv : st.dev) {
        (void)hipSetDevice(kv.first);
        drop_unused(kv.second, true);
    }
    (void)hipSetDevice(cur);
}

size_t cached_bytes(int device)
{
    State &st = state();
    std::lock_guard<std::mutex> lock(st.mu);
    auto it = st.dev.find(device);
    return it == st.dev.end() ? 0 : it->second.cached;
}

}  // namespace devmem
}  // namespace dtb
