// dtb_devmem.cpp -- device memory for the engine's buffers.
//
// hipMalloc and hipFree cost time that grows with the block size (page
// mapping: seconds for tens of GB), which dominates when instances of that
// size come and go every few hundred milliseconds (serving).  So device
// memory stays with the library: every block is a range of a segment (a
// device allocation), taken best-fit in 4 KB units and merged with free
// neighbours when released.  A segment is either reserved up front
// (reserve: a server's whole budget in one allocation) or made when no free
// range holds a request (at least kMinSegment, so small buffers share one;
// with a reservation only small blocks may go outside it).
// A released block may still be read by work in flight, so it is parked
// first; parked blocks become reusable after one device synchronisation,
// taken only when a request finds no free range.  When the device runs out
// of memory, segments with nothing allocated that were not reserved are
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

constexpr size_t kGrain = (size_t)4 << 10;         // allocation unit
constexpr size_t kSegGrain = (size_t)2 << 20;      // segment sizes: multiples of 2 MB
constexpr size_t kMinSegment = (size_t)256 << 20;  // segments made on demand

struct Segment {
    char *base = nullptr;
    RangeSet ranges;
    bool reserved = false;
};

struct Block {
    int device;
    size_t bytes;
    int seg;
    size_t off;
};

struct DeviceCache {
    std::vector<std::unique_ptr<Segment>> segs;   // null once returned
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

// NoAlloc scopes alive on this thread.
thread_local int t_no_alloc = 0;

size_t round_up(size_t bytes, size_t grain)
{
    return (bytes + grain - 1) / grain * grain;
}

void *take(State &st, DeviceCache &c, int device, size_t want)
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

void unpark(DeviceCache &c)
{
    for (auto &pb : c.parked)
        c.segs[(size_t)pb.second.seg]->ranges.give(pb.second.off, pb.second.bytes);
    c.parked.clear();
}

// Returns segments with nothing allocated to the runtime (all of them with
// `reserved`, else those made on demand).  The device is current.
void drop_unused(DeviceCache &c, bool reserved)
{
    if (!c.parked.empty()) {
        (void)hipDeviceSynchronize();
        unpark(c);
    }
    for (auto &s : c.segs) {
        if (s && s->ranges.all_free() && (reserved || !s->reserved)) {
            (void)hipFree(s->base);
            c.cached -= s->ranges.size();
            s.reset();
        }
    }
}

void new_segment(DeviceCache &c, size_t bytes, bool reserved)
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
}

}  // namespace

NoAlloc::NoAlloc()
{
    ++t_no_alloc;
}

NoAlloc::~NoAlloc()
{
    --t_no_alloc;
}

void *alloc(size_t bytes)
{
    if (t_no_alloc > 0)
        throw Error(kInvalidParameter, "this call allocates no device memory, but the engine needs " +
                                           std::to_string(bytes) + " bytes first (new instances, samples or "
                                           "schedule: build the tables with one dtb_run or dtb_engine_run)");
    int device = 0;
    DTB_HIP(hipGetDevice(&device));
    const size_t want = round_up(std::max<size_t>(bytes, 1), kGrain);
    State &st = state();
    std::lock_guard<std::mutex> lock(st.mu);
    DeviceCache &c = st.dev[device];
    void *p = take(st, c, device, want);
    if (p == nullptr && !c.parked.empty()) {
        DTB_HIP(hipDeviceSynchronize());
        unpark(c);
        p = take(st, c, device, want);
    }
    if (p != nullptr)
        return p;
    // With a reservation, large blocks come from it or not at all: the memory
    // outside it stays with the runtime (its own buffers, launches) and
    // other users; the caller (a server) waits for ranges to free up.
    bool reserved = false;
    for (auto &s : c.segs)
        reserved = reserved || (s && s->reserved);
    if (reserved && want > kMinSegment / 4)
        throw Error(kOutOfMemory, "no free range of the reserved device memory holds the block");
    new_segment(c, round_up(std::max(want, kMinSegment), kSegGrain), false);
    return take(st, c, device, want);
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
    if (have < bytes)
        new_segment(c, round_up(bytes - have, kSegGrain), true);
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
