// mcx_pt.h -- device state of a replica-exchange object.
#pragma once

#include "mcx_internal.h"

namespace mcx {

struct PTArgs {
    int32_t n_sets;
    int32_t R;           // replicas per set
    ChainRec* chains;    // chain of (set, replica r) = set * R + r
    int32_t* idx;        // [n_sets * R] ladder index held by replica r (1-based)
    int32_t* pos;        // [n_sets * R] replica holding ladder index k + 1
    const double* ladder;// [n_sets * R] beta of ladder index k + 1
    int32_t* steps;      // [n_sets * (R-1)] attempts per edge
    int32_t* acc;        // [n_sets * (R-1)] accepted exchanges per edge
    int32_t* stage;      // [n_sets]
    uint64_t* round;     // [n_sets] exchange rounds done (the PT stream counter)
    const int32_t* E;    // energies: E_pair + D * Q per chain ...
    const int32_t* Q;
    const double* x;     // ... or caller-supplied energies (non-null overrides)
};

hipError_t launch_pt_exchange(const PTArgs& a, hipStream_t st);

} // namespace mcx
