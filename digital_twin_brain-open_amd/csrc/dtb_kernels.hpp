// dtb_kernels.hpp -- host-callable launchers of the gfx950 step kernels.
#pragma once

#include "dtb_types.hpp"

namespace dtb {

// One step of every tile, launch parity `par` (queue accumulation + fold +
// OU + LIF + binning of the fired rows into queue set par ^ 1).
// wfmt: the weight format of every instance in the tables, or kWeightAny
// (each launch's kernels are specialised for it).
hipError_t launch_step_kernel(const TileDesc *tiles, uint32_t ntiles, const InstDev *insts, const QueueSet &qs,
                              uint32_t par, StepCtl *ctl, uint32_t pass_k, uint32_t wfmt, hipStream_t stream);

// launch_step_kernel with `start` / `stop` recorded at the kernel's own
// dispatch boundaries (kernel time without launch gaps).
hipError_t launch_step_kernel_timed(const TileDesc *tiles, uint32_t ntiles, const InstDev *insts, const QueueSet &qs,
                                    uint32_t par, StepCtl *ctl, uint32_t pass_k, uint32_t wfmt, hipEvent_t start,
                                    hipEvent_t stop, hipStream_t stream);

// Split binning: bins the passes listed by the k_step launch of parity
// `par` (grid workgroups, each wave one pass at a time); nothing when the
// queue set has no pass list.  Events may be null.
hipError_t launch_bin_kernel(const TileDesc *tiles, uint32_t grid, const InstDev *insts, const QueueSet &qs,
                             uint32_t par, StepCtl *ctl, uint32_t pass_k, uint32_t wfmt, hipEvent_t start,
                             hipEvent_t stop, hipStream_t stream);

// Gather propagation: delta of fired list `par` into InstDev::dlt for every
// group (runs before the k_step launch of parity `par`, and before
// materialize / settle of set `par`); empties fired list par ^ 1.
hipError_t launch_gather(const GroupDesc *groups, uint32_t ngroups, const InstDev *insts, uint32_t par, uint32_t wfmt,
                         hipStream_t stream);
hipError_t launch_gather_timed(const GroupDesc *groups, uint32_t ngroups, const InstDev *insts, uint32_t par,
                               uint32_t wfmt, hipEvent_t start, hipEvent_t stop, hipStream_t stream);

// Group offsets of the gather propagation ([n][ngroups + 1], rows ordered
// by target group).
// Gather records ([nsyn][grec_words(wfmt)]: column, weight pair bits).
hipError_t launch_build_grec(const uint32_t *col, const void *w, uint32_t wfmt, uint64_t nsyn, uint32_t *rec,
                             hipStream_t stream);

hipError_t launch_build_segidx(const int64_t *rowptr, const uint32_t *col, uint32_t n, uint32_t ngroups,
                               uint16_t *segidx, hipStream_t stream);

// In-edges of the pull propagation from the presynaptic-major CSR: cptr
// [ntiles + 1] holds each target tile's first in-edge; cursor [ntiles] is
// zeroed and counts the records placed per tile.
hipError_t launch_build_pull(const int64_t *rowptr, const uint32_t *col, const void *w, uint32_t wfmt, uint32_t n,
                             const int64_t *cptr, uint32_t *cursor, uint32_t *rec, hipStream_t stream);

// Pending delta of queue set `set` into InstDev::dlt (queues left intact).
hipError_t launch_materialize(const TileDesc *tiles, uint32_t ntiles, const InstDev *insts, const QueueSet &qs,
                              uint32_t set, hipStream_t stream);

// Pending delta of queue set `set` folded into J now, that set emptied
// (before parameter changes).
hipError_t launch_settle(const TileDesc *tiles, uint32_t ntiles, const InstDev *insts, const QueueSet &qs,
                         uint32_t set, hipStream_t stream);

// I_syn-based statistics of the last step run into the recording (after
// launch_materialize of the pending set); `par` = parity of the next launch.
hipError_t launch_stat_tail(const TileDesc *tiles, uint32_t ntiles, const InstDev *insts, uint32_t par,
                            hipStream_t stream);

// Samples of the step run by the k_step launch of parity `par`.
hipError_t launch_sample(const InstDev *insts, const SampleDev *samples, uint32_t nsamples,
                         const SampleRecDev *recs, uint32_t par, hipStream_t stream);

hipError_t launch_normals(uint64_t seed, uint32_t neuron0, uint64_t substep, uint32_t count, float *out,
                          hipStream_t stream);

}  // namespace dtb
