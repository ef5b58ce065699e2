// dtb_engine_loop.cpp -- the step engine: instance tables, launch sequence,
// HIP-graph replay and kernel timing (tools/dist_simulator.cpp:2072-2315 is
// the reference per-step loop this replaces for one GPU).
#include <algorithm>
#include <vector>

#include "dtb_engine.hpp"
#include "dtb_kernels.hpp"

namespace dtb {

Engine::Engine(int device) : device_(device)
{
    int count = 0;
    DTB_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        throw Error(kNoDevice, "device index out of range");
    DTB_HIP(hipSetDevice(device));
    DTB_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    DTB_HIP(hipStreamCreateWithFlags(&admit_stream_, hipStreamNonBlocking));
    d_ctl_.resize(1);
    d_ctl_.mutable_gpu_data();
}

Engine::~Engine()
{
    drop_graphs();
    insts_.clear();
    if (stream_)
        (void)hipStreamDestroy(stream_);
    if (admit_stream_)
        (void)hipStreamDestroy(admit_stream_);
}

int Engine::add(std::unique_ptr<Instance> inst)
{
    // The instance's queues are allocated now, so that running out of device
    // memory shows up when it is added, not in the middle of a later run.
    DTB_HIP(hipSetDevice(device_));
    inst->ensure_queue(queue_slots(*inst));
    if (use_gather(*inst))
        inst->ensure_gather(admit_stream_);
    else
        try_pull(*inst, admit_stream_);
    for (size_t i = 0; i < insts_.size(); ++i)
        if (!insts_[i]) {
            insts_[i] = std::move(inst);
            dirty_ = true;
            staged_.insert((int)i);
            return (int)i;
        }
    insts_.push_back(std::move(inst));
    dirty_ = true;
    staged_.insert((int)insts_.size() - 1);
    return (int)insts_.size() - 1;
}

std::unique_ptr<Instance> Engine::remove(int id)
{
    if (id < 0 || (size_t)id >= insts_.size() || !insts_[id])
        throw Error(kInvalidParameter, "no such instance");
    settle();
    sync();
    dirty_ = true;
    staged_.erase(id);
    return std::move(insts_[id]);
}

Instance *Engine::get(int id)
{
    if (id < 0 || (size_t)id >= insts_.size() || !insts_[id])
        throw Error(kInvalidParameter, "no such instance");
    return insts_[id].get();
}

size_t Engine::active() const
{
    size_t k = 0;
    for (auto &p : insts_)
        k += p ? 1 : 0;
    return k;
}

void Engine::set_schedule(uint32_t pass_k, int layout)
{
    if (pass_k != 1 && pass_k != 2 && pass_k != 4)
        throw Error(kInvalidParameter, "pass size must be 1, 2 or 4 synapses per lane");
    if (layout != kLayoutAuto && layout != kLayoutDense && layout != kLayoutSpread && layout != kLayoutGather)
        throw Error(kInvalidParameter, "unknown propagation layout");
    if (pass_k != pass_k_ || layout != count_layout_) {
        pass_k_ = pass_k;
        count_layout_ = layout;
        dirty_ = true;   // graphs hold the old kernel; counts move to the new layout
    }
}

bool Engine::use_gather(const Instance &p) const
{
    return p.gather_allowed() && count_layout_ == kLayoutGather;
}

bool Engine::any_gather() const
{
    for (auto &p : insts_)
        if (p && use_gather(*p))
            return true;
    return false;
}

bool Engine::gather_possible() const
{
    for (auto &p : insts_)
        if (p && p->gather_allowed())
            return true;
    return false;
}

void Engine::autotune(int64_t trial_steps, uint32_t *best_k, int *best_layout, float *us)
{
    if (trial_steps < 1)
        throw Error(kInvalidParameter, "autotune needs at least one step per trial");
    if (active() == 0)
        throw Error(kInvalidParameter, "autotune needs an instance to step");
    static const uint32_t ks[7] = {1, 2, 4, 1, 2, 4, kPassKDefault};
    static const int layouts[7] = {kLayoutDense, kLayoutDense, kLayoutDense, kLayoutSpread,
                                   kLayoutSpread, kLayoutSpread, kLayoutGather};
    // Without an instance that allows gather (or device memory for its
    // buffers) the seventh turn runs the default schedule and is not a
    // candidate, so the number of steps does not depend on the networks.
    int nc = gather_possible() ? 7 : 6;
    // The cost of a step follows the network's activity, which drifts and
    // bursts; so the schedules take turns in short windows (kAutotuneWindow
    // timed steps, after 2 steps that apply the schedule) over the whole
    // tuning period, and their totals are compared.  A first untimed turn
    // loads every kernel variant (and builds the gather offsets).
    const int64_t rounds = (trial_steps + kAutotuneWindow - 1) / kAutotuneWindow;
    double t_sum[7] = {0, 0, 0, 0, 0, 0, 0};
    // The gather buffers stay built between the gather turns (a table build
    // would otherwise return them at every other turn and build them again);
    // the first table build after the tuning returns them unless gather won.
    struct Tuning {
        bool &flag;
        explicit Tuning(bool &f) : flag(f) { flag = true; }
        ~Tuning() { flag = false; }
    } tuning(tuning_);
    for (int64_t r = -1; r < rounds; ++r) {
        for (int s = 0; s < 7; ++s) {
            if (s < nc)
                set_schedule(ks[s], layouts[s]);
            else
                set_schedule(kPassKDefault, kLayoutDense);
            try {
                run(2, 0);   // applies the schedule (rebuild) outside the timed span
            } catch (const Error &e) {
                // the gather buffers did not fit: drop the candidate (the
                // failed table build is redone by the next run)
                if (e.code != kOutOfMemory || layouts[s] != kLayoutGather || s >= nc)
                    throw;
                nc = 6;
                set_schedule(kPassKDefault, kLayoutDense);
                run(2, 0);
            }
            if (r < 0)
                continue;
            float ms = 0.0f;
            run_span(kAutotuneWindow, 0, &ms);
            t_sum[s] += ms;
        }
    }
    int bi = 0;
    for (int i = 1; i < nc; ++i)
        if (t_sum[i] < t_sum[bi])
            bi = i;
    if (us)
        for (int i = 0; i < 7; ++i)
            us[i] = i < nc ? (float)(t_sum[i] * 1e3 / (double)(rounds * kAutotuneWindow)) : 0.0f;
    const uint32_t bk = ks[bi];
    const int bl = layouts[bi];
    set_schedule(bk, bl);
    dirty_ = true;   // the next table build returns gather buffers no longer used
    if (best_k)
        *best_k = bk;
    if (best_layout)
        *best_layout = bl;
}

void Engine::set_pull(bool on, uint32_t defer_min)
{
    if (on != pull_on_ || defer_min != defer_min_) {
        pull_on_ = on;
        defer_min_ = defer_min;
        dirty_ = true;
    }
}

void Engine::try_pull(Instance &p, hipStream_t s)
{
    if (!pull_on_ || p.pull_built() || !p.pull_allowed())
        return;
    try {
        p.ensure_pull(s);
    } catch (const Error &e) {
        // without device memory for the in-edges the instance keeps queues
        // with spill accumulators (same results, slower volleys)
        if (e.code != kOutOfMemory)
            throw;
    }
}

CtlTotals Engine::propagation_stats()
{
    StepCtl c{};
    d_ctl_.download(&c, 1, stream_);
    DTB_HIP(hipStreamSynchronize(stream_));
    CtlTotals t{};
    for (uint32_t k = 0; k < kCtlSlots; ++k) {
        t.spilled += c.slot[k].spilled;
        t.deferred += c.slot[k].deferred;
        t.pulled += c.slot[k].pulled;
    }
    return t;
}

uint64_t Engine::queue_slots(const Instance &p) const
{
    uint64_t slots = 0;
    for (uint32_t b = 0, t = 0; b < p.n(); b += kBlock, ++t)
        slots += queue_capacity(p.tile_synapses(t));
    return slots;
}

void Engine::set_queue_capacity(uint32_t floor, uint32_t shift)
{
    if (shift > 31)
        throw Error(kInvalidParameter, "queue capacity shift must be at most 31");
    queue_floor_ = floor;
    queue_shift_ = shift;
    dirty_ = true;
}

uint32_t Engine::queue_capacity(uint32_t synapses) const
{
    // Every row fires at most once per step, so `synapses` entries always
    // suffice; fewer keep the queues small, and the synapses past the
    // capacity take the spill path.
    const uint64_t c = (uint64_t)queue_floor_ + (synapses >> queue_shift_);
    return (uint32_t)std::min<uint64_t>(synapses, c);
}

void Engine::drop_graphs()
{
    for (int p = 0; p < 2; ++p)
        if (graph_[p]) {
            (void)hipGraphExecDestroy(graph_[p]);
            graph_[p] = nullptr;
            graph_steps_[p] = 0;
        }
}

QueueSet Engine::queues()
{
    QueueSet q{};
    q.count = d_qcnt_.mutable_gpu_data();
    q.ntiles = ntiles_;
    q.cshift = cshift_;
    if (pass_cap_ > 0) {
        q.pass = d_pass_.mutable_gpu_data();
        q.pcount = d_pcount_.mutable_gpu_data();
        q.pcap = pass_cap_;
    }
    return q;
}

void Engine::rebuild()
{
    settle();   // the queues are reallocated below; nothing pending may be lost
    drop_graphs();
    staged_.clear();
    std::vector<InstDev> table;
    std::vector<TileDesc> tiles;
    std::vector<GroupDesc> groups;
    std::vector<SampleDev> samples;
    std::vector<SampleRecDev> recs;
    tile_base_.assign(insts_.size(), 0u);
    uint64_t total = 0;
    double coverage = 0.0;
    uint64_t pass_want = 0;   // split binning: pass-list entries for the queue instances
    uint32_t fmt = kWeightAny + 1;   // the instances' common weight format (kWeightAny: mixed)
    for (size_t slot = 0; slot < insts_.size(); ++slot) {
        Instance *p = insts_[slot].get();
        if (!p)
            continue;
        const uint32_t idx = (uint32_t)table.size();
        tile_base_[slot] = (uint32_t)tiles.size();
        fmt = fmt == kWeightAny + 1 || fmt == p->weight_format() ? p->weight_format() : kWeightAny;
        uint64_t qtotal = 0;   // this instance's queue slots per parity
        const auto &pb = p->pop_begin();
        uint32_t q = 0;
        for (uint32_t b = 0, t = 0; b < p->n(); b += kBlock, ++t) {
            while (pb[q + 1] <= b)
                ++q;
            TileDesc d{};
            d.inst = idx;
            d.begin = b;
            d.count = std::min<uint32_t>(kBlock, p->n() - b);
            d.pop0 = q;
            d.qcap = queue_capacity(p->tile_synapses(t));
            d.pcls = p->tile_class(t);
            d.qoff = qtotal;
            qtotal += d.qcap;
            tiles.push_back(d);
        }
        // Each instance owns its queues (both parities and the read-ahead
        // slack), so instances of any size come and go without one large
        // engine-wide block.
        p->ensure_queue(qtotal);
        const bool gat = use_gather(*p);
        if (gat) {
            p->ensure_gather(stream_);
            for (uint32_t g = 0; g < p->ngroups(); ++g)
                groups.push_back(GroupDesc{idx, g, std::min<uint32_t>(1u << kGroupBits, p->n() - (g << kGroupBits)), 0u});
        } else if (p->gather_built() && !tuning_) {
            // the old tables (settled above) were the last to use them;
            // device memory returns to the cache
            p->release_gather();
        }
        if (!gat)
            try_pull(*p, stream_);
        const bool pl = !gat && pull_on_ && p->pull_built();
        table.push_back(p->dev(tile_base_[slot], gat, pl, defer_min_));
        recs.push_back(p->sample_rec());
        if (recs.back().nsample > 0) {
            const auto &ids = p->samples();
            for (uint32_t k = 0; k < ids.size(); ++k)
                samples.push_back(SampleDev{idx, ids[k], k, 0});
        }
        if (!gat) {
            total += p->n();
            coverage += p->row_coverage() * p->n();
            // a step listing one pass per 1/64 of the neurons and per 4096
            // synapses fits; beyond, tiles bin their own passes
            pass_want += p->n() / 64 + (uint64_t)p->nsyn() / 4096;
        }
    }
    if (total > 0xffffffffull || tiles.size() > 0xffffffffull)
        throw Error(kInvalidParameter, "too many neurons in one engine");
    ninst_ = (uint32_t)table.size();
    wfmt_ = fmt > kWeightAny ? kWeightF32 : fmt;
    ntiles_ = (uint32_t)tiles.size();
    ngroups_ = (uint32_t)groups.size();
    nsamples_ = (uint32_t)samples.size();
    // The tables only grow (the kernels get their lengths as arguments), so
    // a rebuild after the largest set so far allocates nothing.
    auto grow = [](auto &buf, size_t n) {
        if (buf.count() < std::max<size_t>(n, 1))
            buf.resize(std::max<size_t>(n, 1));
    };
    grow(d_groups_, groups.size());
    grow(d_insts_, table.size());
    grow(d_tiles_, tiles.size());
    // sized by the samples set, recorded or not: a recording run then finds
    // the table built
    size_t sample_set = 0;
    for (auto &p : insts_)
        if (p)
            sample_set += p->samples().size();
    grow(d_samples_, std::max(samples.size(), sample_set));
    grow(d_recs_, recs.size());
    // Auto: dense, unless rows reach most of their instance's tiles (the
    // reservations of many waves then meet on the same lines).  Measured with
    // the tile's passes dealt to all its waves, dense was as fast or faster
    // up to a coverage of 0.92 (in-degree 5000 on 500k neurons), so the
    // threshold sits above that; the autotuner measures instead of guessing.
    if (count_layout_ == kLayoutDense || count_layout_ == kLayoutGather)
        cshift_ = kCountShiftDense;
    else if (count_layout_ == kLayoutSpread)
        cshift_ = kCountShiftSpread;
    else
        cshift_ = (total > 0 && coverage / (double)total > kSpreadCoverage) ? kCountShiftSpread : kCountShiftDense;
    grow(d_qcnt_, count_words(2 * std::max<size_t>(tiles.size(), 1), cshift_));
    if (split_ && total > 0) {
        pass_cap_ = split_cap_ != 0 ? split_cap_
                                    : (uint32_t)std::min<uint64_t>(pass_want + 4096, (uint64_t)4 << 20);
        grow(d_pass_, pass_cap_);
        grow(d_pcount_, 2 * kPassCountStride);
        bin_grid_ = std::min<uint32_t>(kBinGridMax, std::max<uint32_t>(64, (uint32_t)(total / kBlock)));
    } else {
        pass_cap_ = 0;
        bin_grid_ = 0;
    }
    // Allocate (zero-filled) everything the step kernels touch before any
    // capture: no allocation may happen while a graph is being recorded.
    d_qcnt_.mutable_gpu_data();
    d_samples_.mutable_gpu_data();
    d_recs_.mutable_gpu_data();
    d_groups_.mutable_gpu_data();
    DTB_HIP(hipMemsetAsync(d_qcnt_.mutable_gpu_data(), 0, d_qcnt_.size(), stream_));
    if (pass_cap_ > 0)
        DTB_HIP(hipMemsetAsync(d_pcount_.mutable_gpu_data(), 0, d_pcount_.size(), stream_));
    d_insts_.upload(table.data(), table.size(), stream_);
    d_tiles_.upload(tiles.data(), tiles.size(), stream_);
    d_groups_.upload(groups.data(), groups.size(), stream_);
    d_samples_.upload(samples.data(), samples.size(), stream_);
    d_recs_.upload(recs.data(), recs.size(), stream_);
    DTB_HIP(hipStreamSynchronize(stream_));
    dirty_ = false;
}

void Engine::launch_gather_k(uint32_t par)
{
    DTB_HIP(launch_gather(d_groups_.gpu_data(), ngroups_, d_insts_.gpu_data(), par, wfmt_, stream_));
}

void Engine::launch_step_k()
{
    launch_gather_k(par_);
    DTB_HIP(launch_step_kernel(d_tiles_.gpu_data(), ntiles_, d_insts_.gpu_data(), queues(), par_,
                               d_ctl_.mutable_gpu_data(), pass_k_, wfmt_, stream_));
    DTB_HIP(launch_bin_kernel(d_tiles_.gpu_data(), bin_grid_, d_insts_.gpu_data(), queues(), par_,
                              d_ctl_.mutable_gpu_data(), pass_k_, wfmt_, nullptr, nullptr, stream_));
}

void Engine::launch_aux()
{
    // par_ has already advanced: the step just run had parity par_ ^ 1 and
    // its spikes are queued in set par_.
    if (nsamples_ > 0) {
        launch_gather_k(par_);
        DTB_HIP(launch_materialize(d_tiles_.gpu_data(), ntiles_, d_insts_.gpu_data(), queues(), par_, stream_));
        DTB_HIP(launch_sample(d_insts_.gpu_data(), d_samples_.gpu_data(), nsamples_, d_recs_.gpu_data(), par_ ^ 1u,
                              stream_));
    }
}

void Engine::launch_step()
{
    launch_step_k();
    par_ ^= 1u;
    launch_aux();
}

void Engine::run(int64_t steps, int graph_chunk)
{
    if (steps < 0)
        throw Error(kInvalidParameter, "negative step count");
    DTB_HIP(hipSetDevice(device_));
    if (dirty_)
        rebuild();
    if (ninst_ == 0 || steps == 0)
        return;
    pending_ = true;
    if (graph_chunk > 1 && steps >= graph_chunk) {
        int64_t done = 0;
        while (steps - done >= graph_chunk) {
            const uint32_t p0 = par_;
            if (!graph_[p0] || graph_steps_[p0] != graph_chunk) {
                if (graph_[p0]) {
                    DTB_HIP(hipGraphExecDestroy(graph_[p0]));
                    graph_[p0] = nullptr;
                }
                hipGraph_t g = nullptr;
                DTB_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
                for (int k = 0; k < graph_chunk; ++k)
                    launch_step();
                DTB_HIP(hipStreamEndCapture(stream_, &g));
                DTB_HIP(hipGraphInstantiate(&graph_[p0], g, nullptr, nullptr, 0));
                DTB_HIP(hipGraphDestroy(g));
                graph_steps_[p0] = graph_chunk;
                par_ = p0;   // capture recorded the launches without running them
            }
            DTB_HIP(hipGraphLaunch(graph_[p0], stream_));
            par_ = p0 ^ (uint32_t)(graph_chunk & 1);
            done += graph_chunk;
        }
        for (; done < steps; ++done)
            launch_step();
    } else {
        for (int64_t k = 0; k < steps; ++k)
            launch_step();
    }
}

void Engine::settle()
{
    if (!pending_)
        return;
    DTB_HIP(hipSetDevice(device_));
    launch_gather_k(par_);
    DTB_HIP(launch_settle(d_tiles_.gpu_data(), ntiles_, d_insts_.gpu_data(), queues(), par_, stream_));
    // the deferred rows were folded with the rest of the pending input
    for (auto &p : insts_)
        if (p)
            p->clear_deferred(stream_);
    DTB_HIP(hipStreamSynchronize(stream_));
    pending_ = false;
}

void Engine::materialize()
{
    DTB_HIP(hipSetDevice(device_));
    if (dirty_)
        rebuild();
    launch_gather_k(par_);
    DTB_HIP(launch_materialize(d_tiles_.gpu_data(), ntiles_, d_insts_.gpu_data(), queues(), par_, stream_));
    DTB_HIP(hipStreamSynchronize(stream_));
}

void Engine::clear_queues(int id)
{
    get(id);
    DTB_HIP(hipSetDevice(device_));
    if (dirty_)
        rebuild();
    const uint32_t ntile = (get(id)->n() + kBlock - 1) / kBlock;
    for (uint32_t p = 0; p < 2; ++p) {
        // the counts of these tiles and the unused words between them
        const size_t first = count_slot((size_t)p * ntiles_ + tile_base_[(size_t)id], cshift_);
        const size_t last = count_slot((size_t)p * ntiles_ + tile_base_[(size_t)id] + ntile - 1, cshift_);
        DTB_HIP(hipMemsetAsync(d_qcnt_.mutable_gpu_data() + first, 0, (last + 1 - first) * sizeof(uint32_t), stream_));
    }
    get(id)->clear_spill(stream_);
    get(id)->clear_fired(stream_);
    get(id)->clear_deferred(stream_);
    DTB_HIP(hipStreamSynchronize(stream_));
}

void Engine::stat_tail()
{
    DTB_HIP(hipSetDevice(device_));
    if (dirty_)
        rebuild();
    if (ninst_ == 0)
        return;
    launch_gather_k(par_);
    DTB_HIP(launch_materialize(d_tiles_.gpu_data(), ntiles_, d_insts_.gpu_data(), queues(), par_, stream_));
    DTB_HIP(launch_stat_tail(d_tiles_.gpu_data(), ntiles_, d_insts_.gpu_data(), par_, stream_));
    DTB_HIP(hipStreamSynchronize(stream_));
}

void Engine::run_timed(int64_t steps, float *ms_step)
{
    DTB_HIP(hipSetDevice(device_));
    if (dirty_)
        rebuild();
    if (ninst_ == 0 || steps <= 0)
        return;
    pending_ = true;
    // Every step kernel records its own start and stop events (at the
    // dispatch boundaries, on the engine stream); they are read back once
    // per chunk so the launch pipeline is not drained every step.
    // With gather instances the step is two kernels (k_gather, k_step),
    // each timed.
    const int64_t chunk = 256;
    std::vector<hipEvent_t> ev((size_t)(6 * chunk));
    for (auto &e : ev)
        DTB_HIP(hipEventCreate(&e));
    double ts = 0.0;
    for (int64_t base = 0; base < steps; base += chunk) {
        const int64_t m = std::min<int64_t>(chunk, steps - base);
        for (int64_t k = 0; k < m; ++k) {
            hipEvent_t *e = &ev[(size_t)(6 * k)];
            if (ngroups_ > 0)
                DTB_HIP(launch_gather_timed(d_groups_.gpu_data(), ngroups_, d_insts_.gpu_data(), par_, wfmt_, e[2], e[3],
                                            stream_));
            DTB_HIP(launch_step_kernel_timed(d_tiles_.gpu_data(), ntiles_, d_insts_.gpu_data(), queues(), par_,
                                             d_ctl_.mutable_gpu_data(), pass_k_, wfmt_, e[0], e[1], stream_));
            if (pass_cap_ > 0)
                DTB_HIP(launch_bin_kernel(d_tiles_.gpu_data(), bin_grid_, d_insts_.gpu_data(), queues(), par_,
                                          d_ctl_.mutable_gpu_data(), pass_k_, wfmt_, e[4], e[5], stream_));
            par_ ^= 1u;
            launch_aux();
        }
        DTB_HIP(hipEventSynchronize(ev[(size_t)(6 * (m - 1) + (pass_cap_ > 0 ? 5 : 1))]));
        for (int64_t k = 0; k < m; ++k) {
            hipEvent_t *e = &ev[(size_t)(6 * k)];
            float a = 0.0f;
            DTB_HIP(hipEventElapsedTime(&a, e[0], e[1]));
            ts += a;
            if (ngroups_ > 0) {
                DTB_HIP(hipEventElapsedTime(&a, e[2], e[3]));
                ts += a;
            }
            if (pass_cap_ > 0) {
                DTB_HIP(hipEventElapsedTime(&a, e[4], e[5]));
                ts += a;
            }
        }
    }
    for (auto &e : ev)
        DTB_HIP(hipEventDestroy(e));
    if (ms_step)
        *ms_step = (float)ts;
}

void Engine::run_span(int64_t steps, int graph_chunk, float *ms)
{
    DTB_HIP(hipSetDevice(device_));
    if (dirty_)
        rebuild();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    DTB_HIP(hipEventCreate(&e0));
    DTB_HIP(hipEventCreate(&e1));
    DTB_HIP(hipEventRecord(e0, stream_));
    run(steps, graph_chunk);
    DTB_HIP(hipEventRecord(e1, stream_));
    DTB_HIP(hipEventSynchronize(e1));
    float t = 0.0f;
    DTB_HIP(hipEventElapsedTime(&t, e0, e1));
    DTB_HIP(hipEventDestroy(e0));
    DTB_HIP(hipEventDestroy(e1));
    if (ms)
        *ms = t;
}

void Engine::run_span_units(int64_t steps, int graph_chunk, float *ms, float *ms_steps)
{
    DTB_HIP(hipSetDevice(device_));
    if (dirty_)
        rebuild();
    if (steps <= 0) {
        if (ms)
            *ms = 0.0f;
        return;
    }
    // one event before the first step and one after every unit: a step run
    // eagerly, or a replayed graph chunk
    std::vector<hipEvent_t> ev;
    std::vector<int64_t> len;
    auto mark = [&]() {
        hipEvent_t e = nullptr;
        DTB_HIP(hipEventCreate(&e));
        ev.push_back(e);
        DTB_HIP(hipEventRecord(e, stream_));
    };
    try {
        mark();
        for (int64_t done = 0; done < steps;) {
            const bool chunk = graph_chunk > 1 && steps - done >= graph_chunk;
            const int64_t u = chunk ? graph_chunk : 1;
            run(u, chunk ? graph_chunk : 0);
            mark();
            len.push_back(u);
            done += u;
        }
        DTB_HIP(hipEventSynchronize(ev.back()));
        float total = 0.0f;
        int64_t k = 0;
        for (size_t i = 0; i < len.size(); ++i) {
            float t = 0.0f;
            DTB_HIP(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
            total += t;
            for (int64_t j = 0; j < len[i]; ++j)
                if (ms_steps)
                    ms_steps[k++] = t / (float)len[i];
        }
        if (ms)
            *ms = total;
    } catch (...) {
        for (auto e : ev)
            (void)hipEventDestroy(e);
        throw;
    }
    for (auto e : ev)
        DTB_HIP(hipEventDestroy(e));
}

void Engine::sync()
{
    DTB_HIP(hipSetDevice(device_));
    DTB_HIP(hipStreamSynchronize(stream_));
}

}  // namespace dtb
