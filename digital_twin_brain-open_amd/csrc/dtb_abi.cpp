// dtb_abi.cpp -- extern "C" boundary of libdtb.so (declared in include/dtb_abi.h).
#include "dtb_abi.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "dtb_engine.hpp"
#include "dtb_gen.hpp"
#include "dtb_kernels.hpp"
#include "dtb_npz.hpp"

static_assert(DTB_NOALLOC_SCRATCH_BYTES == dtb::kScratchBytes, "the header states the static buffer's size");

struct dtb_engine_s {
    std::unique_ptr<dtb::Engine> eng;
};

namespace {

thread_local std::string g_last_error;

template <typename F>
int guarded(F f)
{
    try {
        g_last_error.clear();
        f();
        return DTB_OK;
    } catch (const dtb::Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return DTB_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return DTB_INTERNAL;
    } catch (...) {
        g_last_error = "unknown error";
        return DTB_INTERNAL;
    }
}

dtb::Engine &engine(dtb_engine_t h)
{
    if (h == nullptr || !h->eng)
        throw dtb::Error(dtb::kInvalidParameter, "null engine handle");
    return *h->eng;
}

void need(bool ok, const char *what)
{
    if (!ok)
        throw dtb::Error(dtb::kInvalidParameter, what);
}

}  // namespace

extern "C" {

int dtb_version(void)
{
    return DTB_ABI_VERSION;
}

const char *dtb_last_error(void)
{
    return g_last_error.c_str();
}

int dtb_device_count(int *count)
{
    return guarded([&] {
        need(count != nullptr, "count is NULL");
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess)
            c = 0;
        *count = c;
    });
}

int dtb_device_memory_cached(int device, uint64_t *bytes)
{
    return guarded([&] {
        need(bytes != nullptr, "bytes is NULL");
        *bytes = (uint64_t)dtb::devmem::cached_bytes(device);
    });
}

int dtb_device_memory_reserve(int device, uint64_t bytes)
{
    return guarded([&] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
            throw dtb::Error(DTB_NO_DEVICE, "no HIP device");
        need(device >= 0 && device < count, "device out of range");
        DTB_HIP(hipSetDevice(device));
        dtb::devmem::reserve((size_t)bytes);
    });
}

int dtb_device_memory_trim(void)
{
    return guarded([&] { dtb::devmem::trim(); });
}

int dtb_engine_create(int device, dtb_engine_t *out)
{
    return guarded([&] {
        need(out != nullptr, "out is NULL");
        auto h = std::make_unique<dtb_engine_s>();
        h->eng = std::make_unique<dtb::Engine>(device);
        *out = h.release();
    });
}

int dtb_engine_destroy(dtb_engine_t h)
{
    return guarded([&] {
        if (h == nullptr)
            return;
        if (h->eng)
            h->eng->sync();
        delete h;
    });
}

int dtb_instance_add(dtb_engine_t h, const dtb_network *net, float dt, uint64_t seed, int *inst_id)
{
    return guarded([&] {
        need(net != nullptr && inst_id != nullptr, "network and inst_id are required");
        dtb::Engine &e = engine(h);
        DTB_HIP(hipSetDevice(e.device()));
        dtb::NetworkHost nh;
        nh.n = net->n;
        nh.props = net->props;
        nh.nsyn = net->nsyn;
        nh.rowptr = net->rowptr;
        nh.col = net->colinds;
        nh.w = net->weights;
        nh.kind_bits = net->connkinds;
        need(net->weight_type >= DTB_WEIGHTS_F32 && net->weight_type <= DTB_WEIGHTS_U8, "unknown weight type");
        nh.wfmt = net->weight_type == DTB_WEIGHTS_F16 ? dtb::kWeightF16
                  : net->weight_type == DTB_WEIGHTS_U8 ? dtb::kWeightU8
                                                       : dtb::kWeightF32;
        auto inst = std::make_unique<dtb::Instance>(nh, dt, seed, e.admit_stream());
        *inst_id = e.add(std::move(inst));
    });
}

int dtb_instance_add_fanout(dtb_engine_t h, int64_t n, uint32_t k, uint64_t net_seed, const float *g_ui,
                            uint32_t sub_base, float dt, uint64_t seed, int *inst_id)
{
    return guarded([&] {
        need(g_ui != nullptr && inst_id != nullptr, "g_ui and inst_id are required");
        dtb::Engine &e = engine(h);
        DTB_HIP(hipSetDevice(e.device()));
        dtb::FanoutSpec f;
        f.n = n;
        f.K = k;
        f.net_seed = net_seed;
        for (int c = 0; c < 4; ++c)
            f.g_ui[c] = g_ui[c];
        f.sub_base = sub_base;
        auto inst = std::make_unique<dtb::Instance>(f, dt, seed, e.admit_stream());
        *inst_id = e.add(std::move(inst));
    });
}

int dtb_instance_remove(dtb_engine_t h, int inst)
{
    return guarded([&] { engine(h).remove(inst); });
}

int dtb_instance_info(dtb_engine_t h, int inst, int64_t *n, int64_t *nsyn, int32_t *npop, uint32_t *pop_ids,
                      uint32_t *pop_begin)
{
    return guarded([&] {
        dtb::Instance *I = engine(h).get(inst);
        if (n)
            *n = I->n();
        if (nsyn)
            *nsyn = I->nsyn();
        if (npop)
            *npop = (int32_t)I->npop();
        if (pop_ids)
            std::memcpy(pop_ids, I->pop_ids().data(), sizeof(uint32_t) * I->npop());
        if (pop_begin)
            std::memcpy(pop_begin, I->pop_begin().data(), sizeof(uint32_t) * (I->npop() + 1));
    });
}

int dtb_set_ou(dtb_engine_t h, int inst, const uint32_t *brain_ids, const float *mean, const float *deviation,
               const float *correlation_time, int32_t n)
{
    return guarded([&] {
        need(n >= 0, "negative count");
        need(n == 0 || (brain_ids && mean && deviation && correlation_time), "OU arrays are required");
        dtb::Engine &e = engine(h);
        e.get(inst);
        e.settle_for(inst);
        e.get(inst)->set_ou(brain_ids, mean, deviation, correlation_time, (size_t)n, e.stream_for(inst));
        e.mark_dirty();
    });
}

int dtb_set_samples(dtb_engine_t h, int inst, const uint32_t *ids, int32_t n)
{
    return guarded([&] {
        need(n >= 0 && (n == 0 || ids), "sample ids are required");
        dtb::Engine &e = engine(h);
        e.get(inst)->set_samples(ids, (size_t)n);
        e.mark_dirty();
    });
}

int dtb_update_props(dtb_engine_t h, int inst, const uint32_t *neuron_ids, const uint32_t *prop_ids,
                     const float *vals, int32_t n)
{
    return guarded([&] {
        need(n >= 0 && (n == 0 || (neuron_ids && prop_ids && vals)), "property arrays are required");
        dtb::Engine &e = engine(h);
        e.get(inst);
        e.settle_for(inst);
        e.get(inst)->update_props(neuron_ids, prop_ids, vals, (size_t)n, e.stream_for(inst));
        e.mark_dirty();   // the parameter tables may have been reallocated
    });
}

int dtb_update_prop_cols(dtb_engine_t h, int inst, const uint32_t *prop_ids, const uint32_t *brain_ids,
                         const float *vals, int32_t n, int32_t assign)
{
    return guarded([&] {
        need(n >= 0 && (n == 0 || (prop_ids && brain_ids && vals)), "property arrays are required");
        dtb::Engine &e = engine(h);
        e.get(inst);
        e.settle_for(inst);
        e.get(inst)->update_prop_cols(prop_ids, brain_ids, vals, (size_t)n, assign != 0, e.stream_for(inst));
        e.mark_dirty();   // the parameter tables may have been reallocated
    });
}

namespace {

// One recorded run of `iters` steps of instance I (the body of dtb_run).
void run_recorded(dtb::Engine &e, dtb::Instance *I, int32_t iters, int32_t t_steps, uint32_t flags,
                  const dtb_run_outputs *out)
{
    // dist_simulator.cpp:3225-3234 parameter checks.
    need(iters > 0, "iter must be positive");
    const bool use_ou = (flags & DTB_RUN_USE_OU) != 0;
    const bool want_samples =
        (flags & (DTB_RUN_SAMPLE_SPIKE | DTB_RUN_SAMPLE_VMEMB | DTB_RUN_SAMPLE_ISYN | DTB_RUN_SAMPLE_IOU)) != 0;
    if ((flags & DTB_RUN_SAMPLE_IOU) && !use_ou)
        throw dtb::Error(dtb::kInvalidParameter, "i_ou samples need the OU background");
    if (want_samples && I->samples().empty())
        throw dtb::Error(dtb::kUnsetSampleFailure, "samples are not set");
    e.settle();
    I->configure_run(t_steps, use_ou, e.stream());
    const bool freq = (flags & DTB_RUN_FREQ) != 0 && out && out->freq;
    const bool samples = want_samples && out &&
                         (out->sample_spike || out->sample_vmemb || out->sample_isynaptic || out->sample_iou);
    // Per-population means (RunRequest output_vmean / output_imean /
    // output_receptor_imeas, snn.proto:63-77).
    uint32_t stat_mask = 0;
    float *means[dtb::kStatFields] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (out) {
        if ((flags & DTB_RUN_VMEAN) && out->vmean) {
            stat_mask |= dtb::kStatV;
            means[0] = out->vmean;
        }
        if ((flags & DTB_RUN_IMEAN) && out->imean) {
            stat_mask |= dtb::kStatI;
            means[1] = out->imean;
        }
        if ((flags & DTB_RUN_RECEPTOR_IMEAN) && out->ampa_imean && out->nmda_imean && out->gabaa_imean &&
            out->gabab_imean) {
            stat_mask |= dtb::kStatReceptor;
            means[2] = out->ampa_imean;
            means[3] = out->nmda_imean;
            means[4] = out->gabaa_imean;
            means[5] = out->gabab_imean;
        }
    }
    I->begin_record(iters, freq, samples, stat_mask, e.stream());
    e.mark_dirty();
    e.run(iters, (flags & DTB_RUN_GRAPH) ? 64 : 0);
    if (stat_mask & (dtb::kStatI | dtb::kStatReceptor))
        e.stat_tail();
    e.sync();
    if (out)
        I->fetch_record(freq ? out->freq : nullptr, (flags & DTB_RUN_SAMPLE_SPIKE) ? out->sample_spike : nullptr,
                        (flags & DTB_RUN_SAMPLE_VMEMB) ? out->sample_vmemb : nullptr,
                        (flags & DTB_RUN_SAMPLE_ISYN) ? out->sample_isynaptic : nullptr,
                        (flags & DTB_RUN_SAMPLE_IOU) ? out->sample_iou : nullptr, means, e.stream());
    I->end_record();
    e.mark_dirty();
}

}  // namespace

int dtb_run(dtb_engine_t h, int inst, int32_t iters, int32_t t_steps, uint32_t flags, const dtb_run_outputs *out)
{
    return guarded([&] {
        dtb::Engine &e = engine(h);
        run_recorded(e, e.get(inst), iters, t_steps, flags, out);
    });
}

int dtb_run_noalloc(dtb_engine_t h, int inst, int32_t iters, int32_t t_steps, uint32_t flags,
                    const dtb_run_outputs *out)
{
    return guarded([&] {
        need(iters > 0, "iter must be positive");
        dtb::Engine &e = engine(h);
        dtb::Instance *I = e.get(inst);
        need(t_steps == I->t_steps(), "dtb_run_noalloc keeps the instance's substeps: change t_steps with "
                                      "dtb_configure (or one dtb_run) first");
        const bool use_ou = (flags & DTB_RUN_USE_OU) != 0;
        const bool want_samples =
            (flags & (DTB_RUN_SAMPLE_SPIKE | DTB_RUN_SAMPLE_VMEMB | DTB_RUN_SAMPLE_ISYN | DTB_RUN_SAMPLE_IOU)) != 0;
        if ((flags & DTB_RUN_SAMPLE_IOU) && !use_ou)
            throw dtb::Error(dtb::kInvalidParameter, "i_ou samples need the OU background");
        if (want_samples && I->samples().empty())
            throw dtb::Error(dtb::kUnsetSampleFailure, "samples are not set");
        const bool freq = (flags & DTB_RUN_FREQ) != 0 && out && out->freq;
        const bool samples = want_samples && out &&
                             (out->sample_spike || out->sample_vmemb || out->sample_isynaptic || out->sample_iou);
        float *means[dtb::kStatFields] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        uint32_t stat_mask = 0;
        if (out) {
            if ((flags & DTB_RUN_VMEAN) && out->vmean) {
                stat_mask |= dtb::kStatV;
                means[0] = out->vmean;
            }
            if ((flags & DTB_RUN_IMEAN) && out->imean) {
                stat_mask |= dtb::kStatI;
                means[1] = out->imean;
            }
            if ((flags & DTB_RUN_RECEPTOR_IMEAN) && out->ampa_imean && out->nmda_imean && out->gabaa_imean &&
                out->gabab_imean) {
                stat_mask |= dtb::kStatReceptor;
                means[2] = out->ampa_imean;
                means[3] = out->nmda_imean;
                means[4] = out->gabaa_imean;
                means[5] = out->gabab_imean;
            }
        }
        // The call's records in the static array, each at a 256-byte
        // boundary: counts, statistic sums, per-population divisors and
        // means, sample rows.
        const size_t rows = (size_t)iters, np = I->npop(), ns = samples ? I->samples().size() : 0;
        size_t at = 0;
        auto carve = [&](size_t bytes) {
            const size_t o = at;
            at += (bytes + 255) / 256 * 256;
            return o;
        };
        const size_t o_cnt = freq ? carve(rows * np * sizeof(uint32_t)) : 0;
        const size_t o_sum = stat_mask ? carve(rows * np * dtb::kStatFields * sizeof(long long)) : 0;
        const size_t o_div = stat_mask ? carve(np * sizeof(unsigned long long)) : 0;
        size_t o_mean[dtb::kStatFields] = {0, 0, 0, 0, 0, 0};
        for (int k = 0; k < dtb::kStatFields; ++k)
            if (means[k])
                o_mean[k] = carve(rows * np * sizeof(float));
        const size_t o_spk = samples ? carve(rows * ns) : 0;
        const size_t o_v = samples ? carve(rows * ns * sizeof(float)) : 0;
        const size_t o_is = samples ? carve(rows * ns * sizeof(float)) : 0;
        const size_t o_io = samples ? carve(rows * ns * sizeof(float)) : 0;
        if (at > dtb::kScratchBytes)
            throw dtb::Error(dtb::kInvalidParameter,
                             "dtb_run_noalloc records " + std::to_string(at) + " bytes, more than its " +
                                 std::to_string(dtb::kScratchBytes) + "-byte static buffer: run fewer steps per call");

        dtb::devmem::NoAlloc no_alloc;   // a device allocation anywhere below throws
        DTB_HIP(hipSetDevice(e.device()));
        char *base = static_cast<char *>(dtb::scratch_address(e.device()));
        e.settle();
        I->configure_run(t_steps, use_ou, e.stream());   // same substeps: no parameter re-derivation
        dtb::Instance::RecordBuffers b{};
        b.counts = freq ? reinterpret_cast<uint32_t *>(base + o_cnt) : nullptr;
        b.stat = stat_mask ? reinterpret_cast<long long *>(base + o_sum) : nullptr;
        if (samples) {
            b.spike = reinterpret_cast<uint8_t *>(base + o_spk);
            b.v = reinterpret_cast<float *>(base + o_v);
            b.isyn = reinterpret_cast<float *>(base + o_is);
            b.iou = reinterpret_cast<float *>(base + o_io);
        }
        // The tables point at the static array only for this call.
        struct Scope {
            dtb::Engine &e;
            dtb::Instance *I;
            ~Scope()
            {
                I->end_record();
                e.mark_dirty();
            }
        } scope{e, I};
        I->begin_record_at(iters, freq, samples, stat_mask, b, e.stream());
        e.mark_dirty();
        e.run(iters, 0);   // eager launches: a graph would be instantiated (allocated) here
        if (stat_mask & (dtb::kStatI | dtb::kStatReceptor))
            e.stat_tail();
        hipStream_t s = e.stream();
        if (stat_mask) {
            float *dm[dtb::kStatFields];
            for (int k = 0; k < dtb::kStatFields; ++k)
                dm[k] = means[k] ? reinterpret_cast<float *>(base + o_mean[k]) : nullptr;
            DTB_HIP(dtb::launch_stat_means(I->blocked_dev(), I->pops_dev(), (uint32_t)np, b.stat, (uint32_t)rows,
                                           reinterpret_cast<unsigned long long *>(base + o_div), dm, s));
            for (int k = 0; k < dtb::kStatFields; ++k)
                if (means[k])
                    DTB_HIP(hipMemcpyAsync(means[k], dm[k], rows * np * sizeof(float), hipMemcpyDeviceToHost, s));
        }
        if (freq)
            DTB_HIP(hipMemcpyAsync(out->freq, b.counts, rows * np * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        if (samples) {
            if ((flags & DTB_RUN_SAMPLE_SPIKE) && out->sample_spike)
                DTB_HIP(hipMemcpyAsync(out->sample_spike, b.spike, rows * ns, hipMemcpyDeviceToHost, s));
            if ((flags & DTB_RUN_SAMPLE_VMEMB) && out->sample_vmemb)
                DTB_HIP(hipMemcpyAsync(out->sample_vmemb, b.v, rows * ns * sizeof(float), hipMemcpyDeviceToHost, s));
            if ((flags & DTB_RUN_SAMPLE_ISYN) && out->sample_isynaptic)
                DTB_HIP(hipMemcpyAsync(out->sample_isynaptic, b.isyn, rows * ns * sizeof(float),
                                       hipMemcpyDeviceToHost, s));
            if ((flags & DTB_RUN_SAMPLE_IOU) && out->sample_iou)
                DTB_HIP(hipMemcpyAsync(out->sample_iou, b.iou, rows * ns * sizeof(float), hipMemcpyDeviceToHost, s));
        }
        e.sync();
    });
}

int dtb_run_stream(dtb_engine_t h, int inst, int32_t iters, int32_t t_steps, uint32_t flags, int32_t chunk,
                   dtb_run_chunk_fn fn, void *user)
{
    return guarded([&] {
        need(iters > 0, "iter must be positive");
        need(chunk > 0, "chunk must be positive");
        need(fn != nullptr, "callback is NULL");
        dtb::Engine &e = engine(h);
        dtb::Instance *I = e.get(inst);
        const size_t np = I->npop(), ns = I->samples().size();
        const size_t rows = (size_t)std::min(chunk, iters);
        // host rows of one chunk, handed to the callback
        std::vector<uint32_t> freq((flags & DTB_RUN_FREQ) ? rows * np : 0);
        std::vector<uint8_t> spike((flags & DTB_RUN_SAMPLE_SPIKE) ? rows * ns : 0);
        std::vector<float> vm((flags & DTB_RUN_SAMPLE_VMEMB) ? rows * ns : 0);
        std::vector<float> is((flags & DTB_RUN_SAMPLE_ISYN) ? rows * ns : 0);
        std::vector<float> io((flags & DTB_RUN_SAMPLE_IOU) ? rows * ns : 0);
        std::vector<float> means[dtb::kStatFields];
        if (flags & DTB_RUN_VMEAN)
            means[0].resize(rows * np);
        if (flags & DTB_RUN_IMEAN)
            means[1].resize(rows * np);
        if (flags & DTB_RUN_RECEPTOR_IMEAN)
            for (int k = 2; k < dtb::kStatFields; ++k)
                means[k].resize(rows * np);
        auto ptr = [](auto &v) { return v.empty() ? nullptr : v.data(); };
        dtb_run_outputs o{};
        o.freq = ptr(freq);
        o.sample_spike = ptr(spike);
        o.sample_vmemb = ptr(vm);
        o.sample_isynaptic = ptr(is);
        o.sample_iou = ptr(io);
        o.vmean = ptr(means[0]);
        o.imean = ptr(means[1]);
        o.ampa_imean = ptr(means[2]);
        o.nmda_imean = ptr(means[3]);
        o.gabaa_imean = ptr(means[4]);
        o.gabab_imean = ptr(means[5]);
        for (int32_t done = 0; done < iters;) {
            const int32_t m = std::min(chunk, iters - done);
            run_recorded(e, I, m, t_steps, flags, &o);
            if (fn(user, done, m, &o) != 0)
                break;   // the caller stopped the run
            done += m;
        }
    });
}

int dtb_engine_run(dtb_engine_t h, int64_t steps, int32_t graph_chunk)
{
    return guarded([&] {
        need(steps >= 0, "negative step count");
        engine(h).run(steps, graph_chunk);
    });
}

int dtb_configure(dtb_engine_t h, int inst, int32_t t_steps, int32_t use_ou)
{
    return guarded([&] {
        dtb::Engine &e = engine(h);
        e.get(inst);
        e.settle_for(inst);
        if (e.get(inst)->configure_run(t_steps, use_ou != 0, e.stream_for(inst)))
            e.mark_dirty();
    });
}

int dtb_engine_sync(dtb_engine_t h)
{
    return guarded([&] { engine(h).sync(); });
}

int dtb_engine_run_span(dtb_engine_t h, int64_t steps, int32_t graph_chunk, float *ms)
{
    return guarded([&] {
        need(steps >= 0, "negative step count");
        engine(h).run_span(steps, graph_chunk, ms);
    });
}

int dtb_engine_run_span_steps(dtb_engine_t h, int64_t steps, int32_t graph_chunk, float *ms, float *ms_steps)
{
    return guarded([&] {
        need(steps >= 0, "negative step count");
        engine(h).run_span_units(steps, graph_chunk, ms, ms_steps);
    });
}

int dtb_engine_time_kernels(dtb_engine_t h, int64_t steps, float *ms_step)
{
    return guarded([&] {
        need(steps >= 0, "negative step count");
        engine(h).run_timed(steps, ms_step);
    });
}

int dtb_engine_check(dtb_engine_t h)
{
    return guarded([&] {
        dtb::Engine &e = engine(h);
        e.sync();
    });
}

int dtb_engine_queue_capacity(dtb_engine_t h, uint32_t floor, uint32_t shift)
{
    return guarded([&] { engine(h).set_queue_capacity(floor, shift); });
}

int dtb_engine_split(dtb_engine_t h, int32_t on, uint32_t capacity)
{
    return guarded([&] { engine(h).set_split(on != 0, capacity); });
}

int dtb_engine_get_split(dtb_engine_t h, int32_t *on, uint32_t *capacity)
{
    return guarded([&] {
        need(on != nullptr, "on is NULL");
        *on = engine(h).split() ? 1 : 0;
        if (capacity)
            *capacity = engine(h).split_cap();
    });
}

int dtb_engine_pull(dtb_engine_t h, int32_t on, uint32_t defer_min)
{
    return guarded([&] { engine(h).set_pull(on != 0, defer_min); });
}

int dtb_engine_get_pull(dtb_engine_t h, int32_t *on, uint32_t *defer_min)
{
    return guarded([&] {
        const dtb::Engine &e = engine(h);
        if (on)
            *on = e.pull_on() ? 1 : 0;
        if (defer_min)
            *defer_min = e.defer_min();
    });
}

int dtb_instance_pull_info(dtb_engine_t h, int inst, int32_t *allowed, int32_t *built, int32_t *in_use,
                           uint64_t *bytes)
{
    return guarded([&] {
        dtb::Engine &e = engine(h);
        const dtb::Instance *p = e.get(inst);
        if (allowed)
            *allowed = p->pull_allowed() ? 1 : 0;
        if (built)
            *built = p->pull_built() ? 1 : 0;
        if (in_use)
            *in_use = e.uses_pull(*p) ? 1 : 0;
        if (bytes)
            *bytes = p->pull_bytes();
    });
}

int dtb_engine_schedule(dtb_engine_t h, int32_t pass_k, int32_t layout)
{
    return guarded([&] {
        need(pass_k > 0, "pass size must be 1, 2 or 4");
        engine(h).set_schedule((uint32_t)pass_k, layout);
    });
}

int dtb_engine_get_schedule(dtb_engine_t h, int32_t *pass_k, int32_t *layout, int32_t *layout_in_use)
{
    return guarded([&] {
        dtb::Engine &e = engine(h);
        if (pass_k)
            *pass_k = (int32_t)e.pass_k();
        if (layout)
            *layout = e.count_layout();
        if (layout_in_use) {
            *layout_in_use = e.any_gather() ? DTB_LAYOUT_GATHER
                                 : (e.count_shift() == dtb::kCountShiftSpread ? DTB_LAYOUT_SPREAD : DTB_LAYOUT_DENSE);
        }
    });
}

int dtb_instance_propagation(dtb_engine_t h, int inst, int32_t *gather_allowed, int32_t *gather_in_use,
                             float *group_coverage)
{
    return guarded([&] {
        dtb::Engine &e = engine(h);
        dtb::Instance *p = e.get(inst);
        if (gather_allowed)
            *gather_allowed = p->gather_allowed() ? 1 : 0;
        if (gather_in_use)
            *gather_in_use = e.use_gather(*p) ? 1 : 0;
        if (group_coverage)
            *group_coverage = (float)p->group_coverage();
    });
}

int dtb_engine_autotune(dtb_engine_t h, int64_t trial_steps, int32_t *pass_k, int32_t *layout, float *us_per_step)
{
    return guarded([&] {
        uint32_t k = 0;
        int l = 0;
        engine(h).autotune(trial_steps, &k, &l, us_per_step);
        if (pass_k)
            *pass_k = (int32_t)k;
        if (layout)
            *layout = l;
    });
}

int dtb_instance_tile_synapses(dtb_engine_t h, int inst, uint32_t *out)
{
    return guarded([&] {
        need(out != nullptr, "out is NULL");
        const dtb::Instance *p = engine(h).get(inst);
        for (uint32_t b = 0, t = 0; b < p->n(); b += dtb::kBlock, ++t)
            out[t] = p->tile_synapses(t);
    });
}

int dtb_engine_spilled(dtb_engine_t h, uint64_t *synapses)
{
    return guarded([&] {
        need(synapses != nullptr, "synapses is NULL");
        *synapses = engine(h).spilled();
    });
}

int dtb_engine_propagation_stats(dtb_engine_t h, uint64_t *spilled, uint64_t *deferred_tiles, uint64_t *pulled_tiles)
{
    return guarded([&] {
        const dtb::CtlTotals c = engine(h).propagation_stats();
        if (spilled)
            *spilled = c.spilled;
        if (deferred_tiles)
            *deferred_tiles = c.deferred;
        if (pulled_tiles)
            *pulled_tiles = c.pulled;
    });
}

int dtb_read_state(dtb_engine_t h, int inst, float *v, int32_t *t_ref_count, uint8_t *f, uint8_t *emitted,
                   float *i_ou, float *j, float *i_syn)
{
    return guarded([&] {
        dtb::Engine &e = engine(h);
        e.get(inst);
        e.settle();   // the pending delta of the last step is folded into J
        e.get(inst)->read_state(v, t_ref_count, f, emitted, i_ou, j, i_syn, e.stream());
    });
}

int dtb_write_state(dtb_engine_t h, int inst, const float *v, const int32_t *t_ref_count, const uint8_t *f,
                    const float *i_ou, const float *j)
{
    return guarded([&] {
        dtb::Engine &e = engine(h);
        e.get(inst);
        e.settle();
        e.get(inst)->write_state(v, t_ref_count, f, i_ou, j, e.stream());
        if (j)
            e.clear_queues(inst);   // J already includes the delta of the restored step
    });
}

int dtb_set_step(dtb_engine_t h, int inst, int64_t step)
{
    return guarded([&] {
        need(step >= 0, "negative step");
        dtb::Engine &e = engine(h);
        e.get(inst);
        e.settle();   // pending input (incl. deferred rows, keyed by step) is folded first
        e.get(inst)->set_step(step, e.stream());
    });
}

int dtb_counters(dtb_engine_t h, int inst, int64_t *step, uint64_t *events, uint64_t *spikes)
{
    return guarded([&] {
        dtb::Engine &e = engine(h);
        e.sync();
        e.get(inst)->counters(step, events, spikes, e.stream());
    });
}

int dtb_ou_normals(int device, uint64_t seed, uint32_t neuron0, uint64_t substep, uint32_t count, float *out)
{
    return guarded([&] {
        need(out != nullptr || count == 0, "out is NULL");
        int nd = 0;
        DTB_HIP(hipGetDeviceCount(&nd));
        if (device < 0 || device >= nd)
            throw dtb::Error(dtb::kNoDevice, "device index out of range");
        DTB_HIP(hipSetDevice(device));
        if (count == 0)
            return;
        float *d = nullptr;
        DTB_HIP(hipMalloc(reinterpret_cast<void **>(&d), sizeof(float) * count));
        hipError_t err = dtb::launch_normals(seed, neuron0, substep, count, d, nullptr);
        if (err == hipSuccess)
            err = hipMemcpy(out, d, sizeof(float) * count, hipMemcpyDeviceToHost);
        (void)hipFree(d);
        DTB_HIP(err);
    });
}

int dtb_gen_uniform(int64_t n, uint32_t k, uint64_t seed, const float *g_ui, uint32_t sub_base, float *props,
                    int64_t *rowptr, uint32_t *colinds, float *weights, uint32_t *connkinds)
{
    return guarded([&] {
        need(g_ui && props && rowptr && colinds && weights && connkinds, "output arrays are required");
        dtb::gen_uniform(n, k, seed, g_ui, sub_base, props, rowptr, colinds, weights, connkinds);
    });
}

int dtb_gen_gamma(uint64_t seed, uint32_t call, uint32_t first, uint32_t count, float alpha, float beta, float *out)
{
    return guarded([&] {
        need(out != nullptr || count == 0, "out is NULL");
        dtb::gen_gamma(seed, call, first, count, alpha, beta, out);
    });
}

int dtb_gamma_prop_cols(dtb_engine_t h, int inst, const uint32_t *prop_ids, const uint32_t *brain_ids,
                        const float *concentration, const float *rate, int32_t n)
{
    return guarded([&] {
        need(n >= 0 && (n == 0 || (prop_ids && brain_ids && concentration && rate)), "gamma arrays are required");
        dtb::Engine &e = engine(h);
        e.get(inst);
        e.settle_for(inst);
        e.get(inst)->gamma_prop_cols(prop_ids, brain_ids, concentration, rate, (size_t)n, e.stream_for(inst));
        e.mark_dirty();   // the parameter tables may have been reallocated
    });
}

int dtb_read_props(dtb_engine_t h, int inst, float *props)
{
    return guarded([&] {
        need(props != nullptr, "props is NULL");
        engine(h).get(inst)->read_props(props);
    });
}

int dtb_gen_voxel(int64_t n, uint32_t nvox, uint32_t k, float intra_frac, float density, uint64_t seed,
                  const float *g_ui, float *props, int64_t *rowptr, uint32_t *colinds, float *weights,
                  uint32_t *connkinds, uint32_t *vox_sizes)
{
    return guarded([&] {
        need(g_ui && props && rowptr && colinds && weights && connkinds, "output arrays are required");
        dtb::gen_voxel(n, nvox, k, intra_frac, density, seed, g_ui, props, rowptr, colinds, weights, connkinds,
                       vox_sizes);
    });
}

int dtb_gen_fanout(int64_t n, uint32_t k, uint64_t seed, const float *g_ui, uint32_t sub_base, float *props,
                   int64_t *rowptr, uint32_t *colinds, float *weights, uint32_t *connkinds)
{
    return guarded([&] {
        need(g_ui && props && rowptr && colinds && weights && connkinds, "output arrays are required");
        dtb::gen_fanout(n, k, seed, g_ui, sub_base, props, rowptr, colinds, weights, connkinds);
    });
}

int dtb_npz_size(const char *path, int64_t *n, int64_t *nsyn, int32_t *nblocks)
{
    return guarded([&] {
        need(path != nullptr, "path is required");
        dtb::npz_block_size(path, n, nsyn, nblocks);
    });
}

int dtb_npz_read(const char *path, float *props, int64_t *rowptr, uint32_t *colinds, float *weights,
                 uint32_t *connkinds, int64_t *block_begin)
{
    return guarded([&] {
        need(path && props && rowptr && colinds && weights && connkinds, "output arrays are required");
        dtb::npz_block_read(path, props, rowptr, colinds, weights, connkinds, block_begin);
    });
}

int dtb_npz_weight_type(const char *path, int32_t *weight_type)
{
    return guarded([&] {
        need(path && weight_type, "path and weight_type are required");
        const uint32_t f = dtb::npz_weight_format(path);
        *weight_type = f == dtb::kWeightF16 ? DTB_WEIGHTS_F16 : f == dtb::kWeightU8 ? DTB_WEIGHTS_U8 : DTB_WEIGHTS_F32;
    });
}

int dtb_npz_read_typed(const char *path, float *props, int64_t *rowptr, uint32_t *colinds, void *weights,
                       int32_t weight_type, uint32_t *connkinds, int64_t *block_begin)
{
    return guarded([&] {
        need(path && props && rowptr && colinds && weights && connkinds, "output arrays are required");
        need(weight_type >= DTB_WEIGHTS_F32 && weight_type <= DTB_WEIGHTS_U8, "unknown weight type");
        const uint32_t f = weight_type == DTB_WEIGHTS_F16 ? dtb::kWeightF16
                           : weight_type == DTB_WEIGHTS_U8 ? dtb::kWeightU8
                                                           : dtb::kWeightF32;
        dtb::npz_block_read_typed(path, props, rowptr, colinds, weights, f, connkinds, block_begin);
    });
}

}  // extern "C"
