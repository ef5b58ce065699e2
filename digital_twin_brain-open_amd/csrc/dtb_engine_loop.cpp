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
    d_ctl_.resize(1);
    d_ctl_.mutable_gpu_data();
    int cus = 0;
    DTB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    prop_blocks_ = (uint32_t)std::max(1, cus) * 4u;   // 16 binning waves per CU
}

Engine::~Engine()
{
    drop_graph();
    insts_.clear();
    if (stream_)
        (void)hipStreamDestroy(stream_);
}

int Engine::add(std::unique_ptr<Instance> inst)
{
    for (size_t i = 0; i < insts_.size(); ++i)
        if (!insts_[i]) {
            insts_[i] = std::move(inst);
            dirty_ = true;
            return (int)i;
        }
    insts_.push_back(std::move(inst));
    dirty_ = true;
    return (int)insts_.size() - 1;
}

std::unique_ptr<Instance> Engine::remove(int id)
{
    if (id < 0 || (size_t)id >= insts_.size() || !insts_[id])
        throw Error(kInvalidParameter, "no such instance");
    settle();
    sync();
    dirty_ = true;
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

bool Engine::queue_overflow()
{
    StepCtl c{};
    d_ctl_.download(&c, 1, stream_);
    DTB_HIP(hipStreamSynchronize(stream_));
    return c.overflow != 0;
}

void Engine::drop_graph()
{
    if (graph_) {
        (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
        graph_steps_ = 0;
    }
}

void Engine::rebuild()
{
    settle();   // the queues are reallocated below; nothing pending may be lost
    drop_graph();
    std::vector<InstDev> table;
    std::vector<TileDesc> tiles;
    std::vector<SampleDev> samples;
    std::vector<SampleRecDev> recs;
    tile_base_.assign(insts_.size(), 0u);
    uint64_t total = 0, qtotal = 0;
    for (size_t slot = 0; slot < insts_.size(); ++slot) {
        Instance *p = insts_[slot].get();
        if (!p)
            continue;
        const uint32_t idx = (uint32_t)table.size();
        tile_base_[slot] = (uint32_t)tiles.size();
        table.push_back(p->dev((uint32_t)tiles.size()));
        recs.push_back(p->sample_rec());
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
            d.qcap = p->tile_runs(t);
            d.qoff = qtotal;
            qtotal += d.qcap;
            tiles.push_back(d);
        }
        if (recs.back().nsample > 0) {
            const auto &ids = p->samples();
            for (uint32_t k = 0; k < ids.size(); ++k)
                samples.push_back(SampleDev{idx, ids[k], k, 0});
        }
        total += p->n();
    }
    if (total > 0xffffffffull || tiles.size() > 0xffffffffull)
        throw Error(kInvalidParameter, "too many neurons in one engine");
    ninst_ = (uint32_t)table.size();
    ntiles_ = (uint32_t)tiles.size();
    nsamples_ = (uint32_t)samples.size();
    fired_cap_ = (uint32_t)std::max<uint64_t>(total, 1);
    d_insts_.resize(std::max<size_t>(table.size(), 1));
    d_tiles_.resize(std::max<size_t>(tiles.size(), 1));
    d_samples_.resize(std::max<size_t>(samples.size(), 1));
    d_recs_.resize(std::max<size_t>(recs.size(), 1));
    d_fired_.resize((size_t)fired_cap_ * 2);
    d_qcnt_.resize(std::max<size_t>(tiles.size(), 1));
    d_queue_.resize((size_t)std::max<uint64_t>(qtotal, 1));
    // Allocate (zero-filled) everything the step kernels touch before any
    // capture: no allocation may happen while a graph is being recorded.
    d_fired_.mutable_gpu_data();
    d_qcnt_.mutable_gpu_data();
    d_queue_.mutable_gpu_data();
    d_samples_.mutable_gpu_data();
    d_recs_.mutable_gpu_data();
    DTB_HIP(hipMemsetAsync(d_qcnt_.mutable_gpu_data(), 0, d_qcnt_.size(), stream_));
    d_insts_.upload(table.data(), table.size(), stream_);
    d_tiles_.upload(tiles.data(), tiles.size(), stream_);
    d_samples_.upload(samples.data(), samples.size(), stream_);
    d_recs_.upload(recs.data(), recs.size(), stream_);
    DTB_HIP(hipStreamSynchronize(stream_));
    dirty_ = false;
}

void Engine::launch_step_k()
{
    DTB_HIP(launch_step_kernel(d_tiles_.gpu_data(), ntiles_, d_insts_.gpu_data(), d_queue_.gpu_data(),
                               d_qcnt_.mutable_gpu_data(), d_fired_.mutable_gpu_data(), fired_cap_,
                               d_ctl_.mutable_gpu_data(), stream_));
}

void Engine::launch_bin_k()
{
    DTB_HIP(launch_bin(d_insts_.gpu_data(), ninst_, d_fired_.gpu_data(), fired_cap_, d_tiles_.gpu_data(),
                       d_qcnt_.mutable_gpu_data(), d_queue_.mutable_gpu_data(), d_ctl_.mutable_gpu_data(),
                       prop_blocks_, stream_));
}

void Engine::launch_step()
{
    launch_step_k();
    launch_bin_k();
    if (nsamples_ > 0) {
        DTB_HIP(launch_materialize(d_tiles_.gpu_data(), ntiles_, d_insts_.gpu_data(), d_queue_.gpu_data(),
                                   d_qcnt_.gpu_data(), stream_));
        DTB_HIP(launch_sample(d_insts_.gpu_data(), d_samples_.gpu_data(), nsamples_, d_recs_.gpu_data(), stream_));
    }
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
        if (!graph_ || graph_steps_ != graph_chunk) {
            drop_graph();
            hipGraph_t g = nullptr;
            DTB_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < graph_chunk; ++k)
                launch_step();
            DTB_HIP(hipStreamEndCapture(stream_, &g));
            DTB_HIP(hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0));
            DTB_HIP(hipGraphDestroy(g));
            graph_steps_ = graph_chunk;
        }
        int64_t done = 0;
        while (steps - done >= graph_chunk) {
            DTB_HIP(hipGraphLaunch(graph_, stream_));
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
    DTB_HIP(launch_settle(d_tiles_.gpu_data(), ntiles_, d_insts_.gpu_data(), d_queue_.gpu_data(),
                          d_qcnt_.mutable_gpu_data(), stream_));
    DTB_HIP(hipStreamSynchronize(stream_));
    pending_ = false;
}

void Engine::materialize()
{
    DTB_HIP(hipSetDevice(device_));
    if (dirty_)
        rebuild();
    DTB_HIP(launch_materialize(d_tiles_.gpu_data(), ntiles_, d_insts_.gpu_data(), d_queue_.gpu_data(),
                               d_qcnt_.gpu_data(), stream_));
    DTB_HIP(hipStreamSynchronize(stream_));
}

void Engine::clear_queues(int id)
{
    get(id);
    DTB_HIP(hipSetDevice(device_));
    if (dirty_)
        rebuild();
    const uint32_t ntile = (get(id)->n() + kBlock - 1) / kBlock;
    DTB_HIP(hipMemsetAsync(d_qcnt_.mutable_gpu_data() + tile_base_[(size_t)id], 0, ntile * sizeof(uint32_t),
                           stream_));
    DTB_HIP(hipStreamSynchronize(stream_));
}

void Engine::run_timed(int64_t steps, float *ms_step, float *ms_bin)
{
    DTB_HIP(hipSetDevice(device_));
    if (dirty_)
        rebuild();
    if (ninst_ == 0 || steps <= 0)
        return;
    pending_ = true;
    // Events bracket every kernel on the engine stream; they are read back
    // once per chunk so the launch pipeline is not drained every step.
    const int64_t chunk = 256;
    std::vector<hipEvent_t> ev((size_t)(3 * chunk));
    for (auto &e : ev)
        DTB_HIP(hipEventCreate(&e));
    double ts = 0.0, tb = 0.0;
    for (int64_t base = 0; base < steps; base += chunk) {
        const int64_t m = std::min<int64_t>(chunk, steps - base);
        for (int64_t k = 0; k < m; ++k) {
            hipEvent_t *e = &ev[(size_t)(3 * k)];
            DTB_HIP(hipEventRecord(e[0], stream_));
            launch_step_k();
            DTB_HIP(hipEventRecord(e[1], stream_));
            launch_bin_k();
            DTB_HIP(hipEventRecord(e[2], stream_));
            if (nsamples_ > 0) {
                DTB_HIP(launch_materialize(d_tiles_.gpu_data(), ntiles_, d_insts_.gpu_data(), d_queue_.gpu_data(),
                                           d_qcnt_.gpu_data(), stream_));
                DTB_HIP(launch_sample(d_insts_.gpu_data(), d_samples_.gpu_data(), nsamples_, d_recs_.gpu_data(),
                                      stream_));
            }
        }
        DTB_HIP(hipEventSynchronize(ev[(size_t)(3 * (m - 1) + 2)]));
        for (int64_t k = 0; k < m; ++k) {
            hipEvent_t *e = &ev[(size_t)(3 * k)];
            float a = 0.0f, b = 0.0f;
            DTB_HIP(hipEventElapsedTime(&a, e[0], e[1]));
            DTB_HIP(hipEventElapsedTime(&b, e[1], e[2]));
            ts += a;
            tb += b;
        }
    }
    for (auto &e : ev)
        DTB_HIP(hipEventDestroy(e));
    if (ms_step)
        *ms_step = (float)ts;
    if (ms_bin)
        *ms_bin = (float)tb;
}

void Engine::sync()
{
    DTB_HIP(hipSetDevice(device_));
    DTB_HIP(hipStreamSynchronize(stream_));
}

}  // namespace dtb
