// The batch handle: many small periodic structures of uneven size,
// concatenated (atoms of structure s are off[s] .. off[s+1]-1), shared by the
// batched Tersoff (batch.hip) and batched Ewald (ewald_batch.hip) paths.
#pragma once
#include <cstdint>
#include <vector>

#include "agpu_internal.hpp"

struct agpu_batch {
    int nstruct = 0;
    int ntot = 0;
    std::vector<int> h_off;        // atom offsets (nstruct + 1)
    std::vector<double> h_cell;    // 9 per structure: cell[9*s + 3*k + x]
    agpu::DevBuf off, sof, cell, rec, r, counts, offsets, nbr, scan_tmp, eatom,
        partial, gacc, flags;
    // per list entry: bond i -> j as {unit vector, length} in FP64 and FP32,
    // the index of the reverse entry j -> i (-1 if absent) and the owner i
    agpu::DevBuf bnd64, bnd32, rev, own;
    // per list entry: {fc, dfc}, {dE/dzeta, dE/dr}, the bond energy and the
    // force the owner's energy puts on the neighbour; per atom: its own force
    agpu::DevBuf fcd, pp, ebond, fslot, fown;
    // first atom of each chunk of whole rows with at most 256 list entries
    // (nchunk = 0 when some row is longer: the unfused path)
    agpu::DevBuf chunk;
    int nchunk = 0;
    // per list entry, backward: fc, dfc/dr and their R and S derivatives
    agpu::DevBuf cut;
    agpu::HostStats stats;
    double nl_cut = -1.0;          // cutoff of the current list (the requested one + skin)
    uint64_t pos_rev = 0, nl_rev = ~0ull;
    uint64_t struct_rev = 0;       // bumped by every set_structures
    int64_t total = 0;
    // batched Ewald state (k-vector runs per structure), owned
    void *ewb = nullptr;
    void (*ewb_free)(void *) = nullptr;
    ~agpu_batch()
    {
        if (ewb && ewb_free) ewb_free(ewb);
    }
};
