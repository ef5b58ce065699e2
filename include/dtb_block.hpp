// dtb_block.hpp -- C++ host class over the C ABI (header only), with the
// method names of the reference's BrainBlock (include/block.hpp:32-780) for
// the calls that exist on this path, so a C++ node that drove a BrainBlock
// drives this one the same way.  The per-stage methods of the reference
// (update_I_ou_background_gpu, update_V_membrane_gpu,
// update_J_presynaptic_inner_gpu, update_I_synaptic_gpu, stat_Freqs_gpu, ...)
// are one fused launch per step here: run() replaces the loop of
// tools/dist_simulator.cpp:2072-2315 that called them.
//
// Errors throw dtb::BlockError carrying the ABI status (0-5 are the
// reference's SnnStatus values) and message.  One thread per Block.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "dtb_abi.h"

namespace dtb {

class BlockError : public std::runtime_error {
public:
    BlockError(int status, const std::string &what) : std::runtime_error(what), status(status) {}
    int status;
};

inline void check(int status)
{
    if (status != DTB_OK)
        throw BlockError(status, dtb_last_error());
}

class Block {
public:
    // BrainBlock(block_id, gpu_id, delta_t, seed), block.cpp:19-60: the
    // engine is created on `gpu_id`; the network comes from
    // init_config_params_gpu / init_connection_table_gpu or from_network.
    Block(int gpu_id, float delta_t, uint64_t seed) : delta_t_(delta_t), seed_(seed)
    {
        check(dtb_engine_create(gpu_id, &eng_));
    }
    ~Block()
    {
        if (eng_)
            dtb_engine_destroy(eng_);
    }
    Block(const Block &) = delete;
    Block &operator=(const Block &) = delete;

    // configuration.cpp:1079-1667 (properties) and 1778-2156 (connections):
    // both read the same block file here; the first call loads properties and
    // connections, the second checks it names the same file.
    void init_config_params_gpu(const std::string &filename) { load(filename); }
    void init_connection_table_gpu(const std::string &filename) { load(filename); }

    // A network already in memory (reference block layout).
    void from_network(const dtb_network &net)
    {
        if (inst_ >= 0)
            throw BlockError(DTB_INVALID_PARAMETER, "the block already holds a network");
        check(dtb_instance_add(eng_, &net, delta_t_, seed_, &inst_));
        fetch_populations();
    }

    // block.cpp:222-281.
    void set_I_ou_current_param(const std::vector<unsigned int> &brain_indice, const std::vector<float> &means,
                                const std::vector<float> &deviations, const std::vector<float> &correlation_times)
    {
        const size_t n = brain_indice.size();
        if (means.size() != n || deviations.size() != n || correlation_times.size() != n)
            throw BlockError(DTB_INVALID_PARAMETER, "OU parameter arrays differ in length");
        check(dtb_set_ou(eng_, inst(), brain_indice.data(), means.data(), deviations.data(),
                         correlation_times.data(), (int32_t)n));
    }

    // block.cpp:1281-1331 (host arrays here; the reference takes device ones).
    void update_Props_gpu(const unsigned int *neuron_indice, const unsigned int *prop_indice, const float *prop_vals,
                          unsigned int n)
    {
        check(dtb_update_props(eng_, inst(), neuron_indice, prop_indice, prop_vals, (int32_t)n));
    }
    // block.cpp:1333-1400: multiply per population.
    void update_Prop_Cols_gpu(const unsigned int *prop_indice, const unsigned int *brain_indice,
                              const float *hp_vals, unsigned int n)
    {
        check(dtb_update_prop_cols(eng_, inst(), prop_indice, brain_indice, hp_vals, (int32_t)n, 0));
    }
    // block.cpp:1402-1470: assign per population.
    void assign_Prop_Cols_gpu(const unsigned int *prop_indice, const unsigned int *brain_indice,
                              const float *hp_vals, unsigned int n)
    {
        check(dtb_update_prop_cols(eng_, inst(), prop_indice, brain_indice, hp_vals, (int32_t)n, 1));
    }
    // block.cpp:1472-1536: Gamma(alpha, rate beta) per neuron of the population.
    void update_Gamma_Prop_Cols_gpu(const unsigned int *prop_indice, const unsigned int *brain_indice,
                                    const float *alphas, const float *betas, unsigned int n)
    {
        check(dtb_gamma_prop_cols(eng_, inst(), prop_indice, brain_indice, alphas, betas, (int32_t)n));
    }

    // The step loop of snn_run (tools/dist_simulator.cpp:2072-2315): iters
    // steps of t_steps substeps with the per-step outputs of RunResponse.
    void run(int32_t iters, int32_t t_steps, uint32_t flags, const dtb_run_outputs &out)
    {
        check(dtb_run(eng_, inst(), iters, t_steps, flags, &out));
    }
    // Updatesample (tools/dist_simulator.cpp:2703-2739).
    void set_samples(const std::vector<uint32_t> &ids)
    {
        check(dtb_set_samples(eng_, inst(), ids.data(), (int32_t)ids.size()));
    }

    // block.hpp accessors.
    void set_t_step(unsigned int step) { check(dtb_set_step(eng_, inst(), (int64_t)step)); }
    unsigned long long get_seed() const { return seed_; }
    unsigned int get_total_neurons() const { return (unsigned int)n_; }
    unsigned int get_total_subblocks() const { return (unsigned int)pop_ids_.size(); }
    const unsigned int *get_sub_bids_cpu() const { return pop_ids_.data(); }
    const unsigned int *get_sub_bcounts_cpu() const { return pop_begin_.data(); }
    // V, refractory counters, fired flags after the last step.
    void get_state(std::vector<float> &v, std::vector<int32_t> &t, std::vector<uint8_t> &f)
    {
        v.resize((size_t)n_);
        t.resize((size_t)n_);
        f.resize((size_t)n_);
        check(dtb_read_state(eng_, inst(), v.data(), t.data(), f.data(), nullptr, nullptr, nullptr, nullptr));
    }
    // (step index, synaptic events delivered, spikes) since the start.
    void counters(int64_t &step, uint64_t &events, uint64_t &spikes)
    {
        check(dtb_counters(eng_, inst(), &step, &events, &spikes));
    }

    // No counterpart in the reference (it has one schedule): the step
    // schedule (dtb_engine_schedule: pass size 1/2/4, DTB_LAYOUT_*) and the
    // autotuner (dtb_engine_autotune; its trial steps advance the block).
    void set_schedule(int32_t pass_k, int32_t layout) { check(dtb_engine_schedule(eng_, pass_k, layout)); }
    void autotune(int64_t trial_steps) { check(dtb_engine_autotune(eng_, trial_steps, nullptr, nullptr, nullptr)); }

    dtb_engine_t engine() const { return eng_; }
    int instance() const { return inst_; }

private:
    int inst() const
    {
        if (inst_ < 0)
            throw BlockError(DTB_INVALID_PARAMETER, "the block holds no network yet");
        return inst_;
    }
    void load(const std::string &filename)
    {
        if (inst_ >= 0) {
            if (filename != file_)
                throw BlockError(DTB_INVALID_PARAMETER, "properties and connections come from one block file");
            return;
        }
        int64_t n = 0, s = 0;
        int32_t nb = 0;
        check(dtb_npz_size(filename.c_str(), &n, &s, &nb));
        std::vector<float> props((size_t)n * 22), w((size_t)s * 2);
        std::vector<int64_t> rowptr((size_t)n + 1);
        std::vector<uint32_t> col((size_t)s), kinds((size_t)(s + 31) / 32 + 1);
        check(dtb_npz_read(filename.c_str(), props.data(), rowptr.data(), col.data(), w.data(), kinds.data(), nullptr));
        const dtb_network net{n, props.data(), s, rowptr.data(), col.data(), w.data(), kinds.data()};
        from_network(net);
        file_ = filename;
    }
    void fetch_populations()
    {
        int64_t ns = 0;
        int32_t np = 0;
        check(dtb_instance_info(eng_, inst_, &n_, &ns, &np, nullptr, nullptr));
        pop_ids_.resize((size_t)np);
        pop_begin_.resize((size_t)np + 1);
        check(dtb_instance_info(eng_, inst_, nullptr, nullptr, nullptr, pop_ids_.data(), pop_begin_.data()));
    }

    dtb_engine_t eng_ = nullptr;
    int inst_ = -1;
    float delta_t_;
    uint64_t seed_;
    int64_t n_ = 0;
    std::string file_;
    std::vector<uint32_t> pop_ids_, pop_begin_;
};

}  // namespace dtb
