// Placement loops over the batched device step (see placement.hpp).
#include "spectrapack/solver/placement.hpp"

#include <algorithm>
#include <cstring>

#include "spectrapack_hip.h"

namespace spectrapack::solver::gpu {

namespace {

// A device session over the environment and kernels; owns the execution when
// the caller passes none.  `failure` names the step that failed.
struct OpenSession {
    std::unique_ptr<compute::FftExecution> owned;
    sp_hip_session* s {};
    const char* failure {};
    OpenSession(const OpenSession&) = delete;
    OpenSession& operator=(const OpenSession&) = delete;
    OpenSession(const Environment& env, std::span<const OrientedKernel> kernels, int precision, bool incremental,
                double pair_clearance_mm, compute::FftExecution* execution)
    {
        if (!execution) {
            auto made = compute::FftExecution::open(0);
            if (std::holds_alternative<compute::CorrelationFailure>(made)) {
                failure = "device";
                return;
            }
            owned = std::move(std::get<std::unique_ptr<compute::FftExecution>>(made));
            execution = owned.get();
        }
        sp_session_desc desc {};
        for (int a = 0; a != 3; ++a) {
            desc.environment_shape[a] = env.shape[a];
            desc.environment_first[a] = env.first[a];
        }
        desc.orientations = std::uint32_t(kernels.size());
        desc.precision = precision;
        desc.capture_graph = 1;
        if (sp_hip_session_create(execution->device_context(), &desc, &s) != SP_OK) {
            s = nullptr;
            failure = "device";
            return;
        }
        if (sp_hip_session_set_environment(s, env.occupied.data(), env.occupied.size()) != SP_OK ||
            (pair_clearance_mm != 0.0 && sp_hip_session_set_pair_clearance(s, pair_clearance_mm, env.pitch_mm) != SP_OK)) {
            failure = "field";
            return;
        }
        if (incremental && sp_hip_session_set_incremental(s, 1) != SP_OK) {
            failure = "device";
            return;
        }
        for (std::uint32_t o = 0; o != kernels.size(); ++o) {
            const auto& k = kernels[o];
            const std::int64_t first[3] = { k.first[0], k.first[1], k.first[2] };
            if (sp_hip_session_set_kernel(s, o, k.cells.data(), k.cells.size(), k.shape.data(), first) != SP_OK) {
                failure = "field";
                return;
            }
        }
    }
    ~OpenSession()
    {
        if (s) sp_hip_session_destroy(s);
    }
};

}  // namespace

PlacementStats place_parts(const Environment& env, std::span<const OrientedKernel> kernels,
                           const PlacementRequest& request, const runtime::OperationControl& control,
                           const std::function<GapVerdict(const Proposal&)>& accept, compute::FftExecution* execution,
                           std::vector<std::uint8_t>* final_occupancy)
{
    PlacementStats stats;
    if (kernels.empty() || request.capacity == 0) {
        stats.outcome = "GAP_FILL_UNAVAILABLE";
        return stats;
    }
    OpenSession open(env, kernels, request.precision, request.incremental, request.pair_clearance_mm, execution);
    if (open.failure) {
        stats.outcome = std::string_view(open.failure) == "field" ? "GAP_FILL_FIELD" : "GAP_FILL_CORRELATION";
        return stats;
    }
    sp_hip_session* s = open.s;
    control.phase(runtime::Phase::placing);
    std::uint64_t copies = 0;
    while (copies < request.capacity) {
        if (control.poll() != runtime::StopCause::none) {
            stats.outcome = "GAP_FILL_INTERRUPTED";
            break;
        }
        sp_step_result r {};
        const int rc = sp_hip_session_step(s, 0, &r);
        ++stats.steps;
        stats.precision = r.precision;
        stats.error_bound = r.error_bound;
        stats.max_residual = std::max(stats.max_residual, r.max_residual);
        if (rc != SP_OK) {
            stats.outcome = control.poll() != runtime::StopCause::none ? "GAP_FILL_INTERRUPTED" : "GAP_FILL_CORRELATION";
            break;
        }
        if (r.orientation < 0) {
            stats.outcome = "GAP_FILL_EXHAUSTED";
            break;
        }
        Proposal p;
        p.orientation = std::uint32_t(r.orientation);
        p.rotation_xyzw = kernels[p.orientation].rotation_xyzw;
        for (int a = 0; a != 3; ++a) {
            p.translation[a] = r.translation[a];
            p.translation_mm[a] = env.pitch_mm * static_cast<double>(r.translation[a]);
        }
        ++stats.proposals;
        const GapVerdict v = accept(p);
        if (v == GapVerdict::stop) {
            stats.outcome = "GAP_FILL_STOPPED";
            break;
        }
        const int rc2 = v == GapVerdict::accepted ? sp_hip_session_stamp(s, p.orientation, r.translation)
                                                  : sp_hip_session_exclude(s, p.orientation, r.translation);
        if (rc2 != SP_OK) {
            stats.outcome = "GAP_FILL_CORRELATION";
            break;
        }
        if (v == GapVerdict::accepted) {
            ++stats.accepted;
            ++copies;
            control.progress("placing", copies, request.capacity);
        }
    }
    if (copies >= request.capacity) stats.outcome = "GAP_FILL_CAPACITY";
    if (final_occupancy) {
        final_occupancy->resize(env.occupied.size());
        (void)sp_hip_session_get_environment(s, final_occupancy->data(), final_occupancy->size());
    }
    return stats;
}

RankingStats rank_parts(const Environment& env, std::span<const OrientedKernel> kernels,
                        std::span<const double> object_max_z_mm, const RankingRequest& request,
                        const runtime::OperationControl& control, const std::function<bool(const Proposal&)>& validate,
                        compute::FftExecution* execution, std::vector<RankedPlacement>* placements,
                        std::vector<std::uint8_t>* final_occupancy)
{
    RankingStats stats;
    if (kernels.empty() || object_max_z_mm.size() != kernels.size()) {
        stats.outcome = "SPECTRAL_UNAVAILABLE";
        return stats;
    }
    OpenSession open(env, kernels, request.precision, request.incremental, request.pair_clearance_mm, execution);
    if (open.failure) {
        stats.outcome = "SPECTRAL_DEVICE";
        return stats;
    }
    sp_hip_session* s = open.s;
    const sp_ranking_desc rank { request.lattice_origin_z_mm, env.pitch_mm, request.container_min_z_mm,
                                 request.container_max_z_mm };
    if (sp_hip_session_set_ranking(s, &rank, object_max_z_mm.data(), std::uint32_t(object_max_z_mm.size())) != SP_OK) {
        stats.outcome = "SPECTRAL_PIPELINE_CONTAINER_HEIGHT";
        return stats;
    }
    struct Hook {
        const std::function<bool(const Proposal&)>* validate;
        std::span<const OrientedKernel> kernels;
        double pitch;
        std::uint64_t calls;
        static int call(void* user, std::uint32_t o, const std::int64_t t[3])
        {
            auto* h = static_cast<Hook*>(user);
            ++h->calls;
            if (!*h->validate) return 1;
            Proposal p;
            p.orientation = o;
            p.rotation_xyzw = h->kernels[o].rotation_xyzw;
            for (int a = 0; a != 3; ++a) {
                p.translation[a] = t[a];
                p.translation_mm[a] = h->pitch * static_cast<double>(t[a]);
            }
            return (*h->validate)(p) ? 1 : 0;
        }
    } hook { &validate, kernels, env.pitch_mm, 0 };
    control.phase(runtime::Phase::placing);
    stats.outcome = "SPECTRAL_COPY_LIMIT";
    while (stats.placed < request.capacity) {
        if (control.poll() != runtime::StopCause::none) {
            stats.outcome = "SPECTRAL_INTERRUPTED";
            break;
        }
        // one placement per call: the pages start over after every placement
        sp_placement pl {};
        std::uint32_t n = 0;
        if (sp_hip_session_ranked_fill(s, 1, &Hook::call, &hook, &pl, 1, &n) != SP_OK) {
            stats.outcome = "SPECTRAL_DEVICE";
            break;
        }
        if (n == 0) {
            stats.outcome = "SPECTRAL_PIPELINE_COMPLETE";
            break;
        }
        ++stats.placed;
        if (placements) {
            RankedPlacement r;
            r.orientation = pl.orientation;
            r.rotation_xyzw = kernels[pl.orientation].rotation_xyzw;
            for (int a = 0; a != 3; ++a) {
                r.translation[a] = pl.translation[a];
                r.translation_mm[a] = env.pitch_mm * static_cast<double>(pl.translation[a]);
            }
            r.score = pl.score;
            placements->push_back(r);
        }
        control.progress("placing", stats.placed, request.capacity);
    }
    stats.validations = hook.calls;
    if (final_occupancy) {
        final_occupancy->resize(env.occupied.size());
        (void)sp_hip_session_get_environment(s, final_occupancy->data(), final_occupancy->size());
    }
    return stats;
}

std::function<GapVerdict(const Proposal&)> exact_accept(geometry::exact::Validator& validator,
                                                        geometry::exact::ValidationReport* last)
{
    return [&validator, last](const Proposal& p) {
        geometry::exact::Pose pose;
        pose.translation_mm = p.translation_mm;
        pose.rotation_xyzw = p.rotation_xyzw;
        const auto r = validator.try_add(pose);
        if (last) *last = r;
        return r.validity == geometry::exact::Validity::valid ? GapVerdict::accepted : GapVerdict::rejected;
    };
}

std::vector<OrientedKernel> voxelize_kernels(std::span<const std::array<double, 3>> vertices,
                                             std::span<const std::array<std::uint32_t, 3>> triangles,
                                             std::span<const std::array<double, 4>> rotations, double pitch_mm,
                                             compute::FftExecution& execution)
{
    const std::size_t n = rotations.size();
    std::vector<std::int64_t> firsts(3 * n);
    std::vector<std::uint32_t> shapes(3 * n);
    std::vector<std::uint64_t> offsets(n);
    std::uint64_t required = 0;
    const double origin[3] = { 0, 0, 0 };
    auto* v = reinterpret_cast<const double*>(vertices.data());
    auto* t = reinterpret_cast<const std::uint32_t*>(triangles.data());
    auto* q = reinterpret_cast<const double*>(rotations.data());
    int rc = sp_hip_voxelize_objects(execution.device_context(), v, std::uint32_t(vertices.size()), t,
                                     std::uint32_t(triangles.size()), origin, pitch_mm, q, std::uint32_t(n),
                                     firsts.data(), shapes.data(), offsets.data(), nullptr, 0, &required);
    if (rc != SP_CAPACITY && rc != SP_OK) return {};
    std::vector<std::uint8_t> cells(required);
    rc = sp_hip_voxelize_objects(execution.device_context(), v, std::uint32_t(vertices.size()), t,
                                 std::uint32_t(triangles.size()), origin, pitch_mm, q, std::uint32_t(n), firsts.data(),
                                 shapes.data(), offsets.data(), cells.data(), cells.size(), &required);
    if (rc != SP_OK) return {};
    std::vector<OrientedKernel> out(n);
    for (std::size_t k = 0; k != n; ++k) {
        out[k].rotation_xyzw = rotations[k];
        std::uint64_t m = 1;
        for (int a = 0; a != 3; ++a) {
            out[k].first[a] = firsts[3 * k + a];
            out[k].shape[a] = shapes[3 * k + a];
            m *= shapes[3 * k + a];
        }
        out[k].cells.assign(cells.begin() + offsets[k], cells.begin() + offsets[k] + m);
    }
    return out;
}

}  // namespace spectrapack::solver::gpu
