// Launch interface of the CPU emulation (test infrastructure).
#pragma once

#include <cstdint>
#include <string>

#include "hip_emu.hpp"

namespace emu {

struct Stats {
    uint64_t barriers = 0, collectives = 0, switches = 0;
};

// Runs `body(arg)` as workgroup `block_id` of a grid of `grid` workgroups
// of `threads` lanes each.  One workgroup per call: `__shared__` variables
// are function statics here, so workgroups cannot interleave (the batched
// kernels share nothing across workgroups but their item counter and the
// per-workgroup workspace slot that `block_id` selects).  `order` picks the
// order in which runnable lanes run between synchronization points: 0
// ascending, 1 descending, 2 shuffled every round (seeded by `seed`).
// Returns 0, or 1 with the error in *err.
int launch(int block_id, int grid, int threads, int order, unsigned seed, void (*body)(void *), void *arg,
           std::string *err, Stats *stats);

}  // namespace emu
