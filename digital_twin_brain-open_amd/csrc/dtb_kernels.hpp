// dtb_kernels.hpp -- host-callable launchers of the gfx950 step kernels.
#pragma once

#include "dtb_types.hpp"

namespace dtb {

hipError_t launch_neuron_step(const WgDesc *wgs, uint32_t nwg, const InstDev *insts, FiredDev *fired,
                              uint32_t fired_cap, StepCtl *ctl, hipStream_t stream);

hipError_t launch_propagate(const InstDev *insts, uint32_t ninst, const FiredDev *fired, uint32_t fired_cap,
                            StepCtl *ctl, uint32_t nblocks, hipStream_t stream);

hipError_t launch_sample(const InstDev *insts, const SampleDev *samples, uint32_t nsamples,
                         const SampleRecDev *recs, hipStream_t stream);

hipError_t launch_normals(uint64_t seed, uint32_t neuron0, uint64_t substep, uint32_t count, float *out,
                          hipStream_t stream);

}  // namespace dtb
