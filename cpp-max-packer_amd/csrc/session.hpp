// Internal view of a placement session for code in the same library that
// proposes orientation subsets and regions on the device (the learned
// proposal).  Not part of the C ABI.
#pragma once

#include <functional>

#include <hip/hip_runtime.h>

#include "fft_kernels.hpp"

struct sp_hip_session;

namespace sphip {

constexpr int kMaxRegionsPerAxis = 64;

struct SessionView {
    const unsigned char* env;  // occupancy, X-fastest, ext[0] * ext[1] * ext[2]
    int ext[3];
    int orientations;
    int fitting;               // orientations whose kernel fits the window
    int* act;                  // subset list: kActHeader ints + one per slot
    unsigned char* region;     // region mask, up to kMaxRegionsPerAxis^3 bytes
    int max_regions_per_axis;
    const OrientDesc* descs;   // one per orientation (lx <= 0: does not fit)
    const unsigned char* kcells;      // kernel cells, X-fastest, at descs[o].cells_offset
    unsigned long long* orient_best;  // per orientation, the last step's lowest feasible key (~0: none)
    StepDeviceResult* result;         // the step's result (key read by the recheck and the stamp)
    hipStream_t stream;
    std::uint64_t serial;     // unique per session in the process
    std::uint64_t reset_rev;  // changes when feasibility changes other than by stamping
};

// Prepares the session if needed and returns its device view.
SessionView session_view(sp_hip_session* s);

// Runs `propose` (which fills act / region on the session stream) followed by
// the subset search over `slots` grid slots; recorded once per owner as a
// hipGraph when the session captures graphs.  merge (nullable) runs after the
// search and before the recheck of the chosen corner and the stamp: it may
// replace the result's key.
void session_launch_proposed(sp_hip_session* s, const void* owner, const std::function<void(hipStream_t)>& propose,
                             int slots, bool stamp, const std::function<void(hipStream_t)>* merge = nullptr);

// The plain search over every orientation, whatever restriction is set.
void session_launch_full(sp_hip_session* s, bool stamp);

// Result of the last step (waits for the session stream).
void session_read_result(sp_hip_session* s, sp_step_result& out);

}  // namespace sphip
