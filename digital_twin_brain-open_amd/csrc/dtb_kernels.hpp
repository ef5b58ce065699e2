// dtb_kernels.hpp -- host-callable launchers of the gfx950 step kernels.
#pragma once

#include "dtb_types.hpp"

namespace dtb {

// Neuron step of every tile (queue accumulation + fold + OU + LIF + fired list).
hipError_t launch_step_kernel(const TileDesc *tiles, uint32_t ntiles, const InstDev *insts, const QItem *queue,
                              uint32_t *qcnt, FiredDev *fired, uint32_t fired_cap, StepCtl *ctl,
                              hipStream_t stream);

// Binning of the fired rows into per-tile queues; also advances the step.
hipError_t launch_bin(const InstDev *insts, uint32_t ninst, const FiredDev *fired, uint32_t fired_cap,
                      const TileDesc *tiles, uint32_t *qcnt, QItem *queue, StepCtl *ctl, uint32_t nblocks,
                      hipStream_t stream);

// Pending delta of every neuron into InstDev::dlt (queues left intact).
hipError_t launch_materialize(const TileDesc *tiles, uint32_t ntiles, const InstDev *insts, const QItem *queue,
                              const uint32_t *qcnt, hipStream_t stream);

// Pending delta folded into J now, queues emptied (before parameter changes).
hipError_t launch_settle(const TileDesc *tiles, uint32_t ntiles, const InstDev *insts, const QItem *queue,
                         uint32_t *qcnt, hipStream_t stream);

hipError_t launch_sample(const InstDev *insts, const SampleDev *samples, uint32_t nsamples,
                         const SampleRecDev *recs, hipStream_t stream);

hipError_t launch_normals(uint64_t seed, uint32_t neuron0, uint64_t substep, uint32_t count, float *out,
                          hipStream_t stream);

}  // namespace dtb
