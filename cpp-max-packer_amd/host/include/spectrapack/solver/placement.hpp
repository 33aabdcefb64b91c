// The greedy spectral placement loops on one MI355X.
//
// place_parts: the plain loop, the role engine/pack_solver/src/gap_fill.cpp:241-588
// (fill_gaps) plays.  Each proposal is the lowest kernel-window corner in
// (z, y, x) order over all orientations, the earliest orientation winning ties
// (gap_fill.cpp:461-498); the caller's verdict decides whether it is stamped
// (accepted), blocked for that orientation (rejected) or ends the run (stop),
// as at gap_fill.cpp:512-523.  With `incremental` the overlap fields are
// patched by the correlation of each placement's new cells, as fill_gaps does
// with its blocked offsets; results are identical to recomputing them.
//
// rank_parts: the ranked loop of the spectral pipeline (spectral.cpp:1485-1590
// over spectral_pipeline.cpp:1740-1857), for containers given as meshes: the
// best page head by (.65 * top - .35 * mean proximity, orientation, z, y, x)
// whose direct recheck finds no overlap and that `validate` accepts is placed.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <span>
#include <string_view>
#include <vector>

#include "spectrapack/compute/correlation.hpp"
#include "spectrapack/geometry/exact_validation.hpp"
#include "spectrapack/runtime/operation_control.hpp"

struct sp_hip_session;

namespace spectrapack::solver::gpu {

enum class GapVerdict { accepted, rejected, stop };

// One oriented kernel: {0,1} cells (X-fastest) of its trimmed window.
struct OrientedKernel {
    std::array<double, 4> rotation_xyzw {};
    compute::Index3 first {};
    compute::Shape3 shape {};
    std::vector<std::uint8_t> cells;
};

struct Environment {
    compute::Index3 first {};
    compute::Shape3 shape {};
    double pitch_mm {};
    std::vector<std::uint8_t> occupied;  // X-fastest, 1 = blocked
};

struct Proposal {
    std::uint32_t orientation {};
    std::array<double, 4> rotation_xyzw {};
    compute::Index3 translation {};           // kernel cell i sits at world cell translation + first + i
    std::array<double, 3> translation_mm {};  // pitch * translation (gap_fill.cpp:507-509)
};

struct PlacementRequest {
    std::uint64_t capacity {};
    int precision {};         // 0 auto (a-priori bound), 32 or 64
    bool incremental { true };  // update overlap fields by each placement's new cells (same results)
    double pair_clearance_mm {};  // copies block cells within this distance (constraints().pair_clearance_mm)
};

struct PlacementStats {
    std::uint64_t steps {}, proposals {}, accepted {};
    int precision {};
    double error_bound {}, max_residual {};
    std::string_view outcome { "GAP_FILL_NOT_RUN" };
};

// Runs steps until capacity, exhaustion, a stop verdict or the control's
// stop/deadline (polled between graph launches).  Outcome strings follow
// gap_fill.cpp: GAP_FILL_CAPACITY, GAP_FILL_EXHAUSTED, GAP_FILL_STOPPED,
// GAP_FILL_INTERRUPTED, GAP_FILL_CORRELATION (device or numeric failure).
PlacementStats place_parts(const Environment& environment, std::span<const OrientedKernel> kernels,
                           const PlacementRequest& request, const runtime::OperationControl& control,
                           const std::function<GapVerdict(const Proposal&)>& accept,
                           compute::FftExecution* execution = nullptr, std::vector<std::uint8_t>* final_occupancy = nullptr);

// The exact host validation as place_parts' accept verdict (the native
// validation callback of gap_fill.cpp:512-523, validation.cpp:517-634): a
// proposal is accepted when `validator` certifies its pose (rotation of its
// orientation, translation pitch * t) against the container and every copy
// accepted before, and rejected otherwise (invalid or unresolved).  The last
// report is kept in *last when given.
std::function<GapVerdict(const Proposal&)> exact_accept(geometry::exact::Validator& validator,
                                                        geometry::exact::ValidationReport* last = nullptr);

struct RankingRequest {
    std::uint64_t capacity {};
    int precision {};             // 0 auto (a-priori bound), 32 or 64: the binary overlaps
    bool incremental { true };
    double pair_clearance_mm {};
    double lattice_origin_z_mm {};  // lattice.origin_mm[2]
    double container_min_z_mm {};   // container bounds in z
    double container_max_z_mm {};
};

struct RankedPlacement {
    std::uint32_t orientation {};
    std::array<double, 4> rotation_xyzw {};
    compute::Index3 translation {};
    std::array<double, 3> translation_mm {};
    double score {};
};

struct RankingStats {
    std::uint64_t placed {}, validations {};
    std::string_view outcome { "SPECTRAL_NOT_RUN" };
};

// Places copies until capacity (SPECTRAL_COPY_LIMIT), no page has a candidate
// (SPECTRAL_PIPELINE_COMPLETE), the control's stop/deadline (polled between
// placements: SPECTRAL_INTERRUPTED) or a device failure (SPECTRAL_DEVICE).
// object_max_z_mm: each orientation's oriented-bounds max z in mm.
// validate(p) is the caller's exact check (spectral.cpp:1576 try_pose).
RankingStats rank_parts(const Environment& environment, std::span<const OrientedKernel> kernels,
                        std::span<const double> object_max_z_mm, const RankingRequest& request,
                        const runtime::OperationControl& control, const std::function<bool(const Proposal&)>& validate,
                        compute::FftExecution* execution = nullptr, std::vector<RankedPlacement>* placements = nullptr,
                        std::vector<std::uint8_t>* final_occupancy = nullptr);

// Rasterises a closed mesh (local frame) for every rotation on the device
// (conservative_fields.cpp:919-1003).  Empty on failure.
std::vector<OrientedKernel> voxelize_kernels(std::span<const std::array<double, 3>> vertices,
                                             std::span<const std::array<std::uint32_t, 3>> triangles,
                                             std::span<const std::array<double, 4>> rotations, double pitch_mm,
                                             compute::FftExecution& execution);

}  // namespace spectrapack::solver::gpu
