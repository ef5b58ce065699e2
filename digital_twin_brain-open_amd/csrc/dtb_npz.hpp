// dtb_npz.hpp -- reader of the reference's block files (see dtb_npz.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "dtb_types.hpp"

namespace dtb {

// block_0.npz, block_1.npz, ... of a directory (in block order), or the one
// file named by `path`.
std::vector<std::string> npz_block_files(const std::string &path);

// Totals over the block files: neurons, synapses, blocks.
void npz_block_size(const std::string &path, int64_t *n, int64_t *nsyn, int32_t *nblocks);

// Fills a network in the engine's input layout: the blocks' neurons
// concatenated in block order (block_begin [nblocks+1] may be null),
// presynaptic-major CSR with every row sorted by postsynaptic id, fp32 weight
// pairs and the channel bit packed 32 per word.
void npz_block_read(const std::string &path, float *props, int64_t *rowptr, uint32_t *colinds, float *weights,
                    uint32_t *connkinds, int64_t *block_begin);

// Weight format the blocks keep on the device (kWeightU8 for uint8 / int8
// files, kWeightF16 for float16, else kWeightF32).
uint32_t npz_weight_format(const std::string &path);

// npz_block_read with the weight pairs written in format wfmt: kWeightF32 as
// above; kWeightF16 / kWeightU8 copy the stored codes (the blocks must store
// that type).
void npz_block_read_typed(const std::string &path, float *props, int64_t *rowptr, uint32_t *colinds, void *weights,
                          uint32_t wfmt, uint32_t *connkinds, int64_t *block_begin);

}  // namespace dtb
