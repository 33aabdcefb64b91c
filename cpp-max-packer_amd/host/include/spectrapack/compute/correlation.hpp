// Drop-in replacement of engine/pack_compute/include/spectrapack/compute/correlation.hpp
// (lines 17-117): the same types, names and failure variants, executed by the
// MI355X library behind include/spectrapack_hip.h.  pack_solver compiles
// unchanged against this header and links libspectrapack_compute.so instead of
// pack_compute.  The `_cpu` names are kept because they are the interface the
// solver calls (spectral_pipeline.cpp:1515,1563; gap_fill.cpp:392,432;
// slab_lattice.cpp:675); the `_hip` names are the same functions.
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <optional>
#include <span>
#include <string_view>
#include <variant>
#include <vector>

#include "spectrapack/runtime/operation_control.hpp"

struct sp_hip_ctx;

namespace spectrapack::compute {

using Shape3 = std::array<std::uint32_t, 3>;
using Index3 = std::array<std::int64_t, 3>;

// Storage is X-fastest. The cell at local kernel index i occupies world index
// translation + kernel_first + i; environment cell (0,0,0) is environment_first.
struct CorrelationSpec {
    Shape3 environment_shape;
    Shape3 kernel_shape;
    Index3 environment_first;
    Index3 kernel_first;
};

struct CorrelationLimits {
    std::uint64_t max_working_bytes { 512ULL << 20 };
    std::uint64_t reserved_bytes {};
    std::uint64_t max_padded_cells { 16'777'216 };
    std::uint64_t max_direct_terms { 200'000'000 };
};

struct CorrelationStats {
    std::uint64_t padded_cells {};
    std::uint64_t working_bytes_peak {};
    std::uint64_t direct_terms {};
};

struct NumericReport {
    double max_integer_residual {};
    double mass_residual {};
    double max_probe_error {};
};

struct CorrelationFailure {
    std::string_view code;
    std::string_view message;
    CorrelationStats stats;
};

struct CorrelationResult {
    Index3 translation_first {};
    Shape3 shape {};
    std::vector<double> values;
    NumericReport numeric;
    CorrelationStats stats;
};

using CorrelationOutcome = std::variant<CorrelationResult, CorrelationFailure>;

// The owner of one device execution: a HIP context (stream, buffers, FFT
// tables) on one GPU.  One coordinator owns it and serializes calls; after a
// device failure it is closed and cannot be reused (CORRELATION_EXECUTION_CLOSED).
class FftExecution {
public:
    ~FftExecution();
    FftExecution(const FftExecution&) = delete;
    FftExecution& operator=(const FftExecution&) = delete;
    [[nodiscard]] std::uint32_t thread_count() const noexcept { return 1; }
    [[nodiscard]] std::uint64_t reserved_bytes() const noexcept { return 0; }
    [[nodiscard]] sp_hip_ctx* device_context() const noexcept { return ctx_; }
    [[nodiscard]] bool closed() const noexcept { return closed_; }
    void close() noexcept { closed_ = true; }
    // Device 0 unless SPECTRAPACK_HIP_DEVICE names another.
    static std::variant<std::unique_ptr<FftExecution>, CorrelationFailure> open(int device);

private:
    explicit FftExecution(sp_hip_ctx* ctx) : ctx_(ctx) {}
    sp_hip_ctx* ctx_ {};
    bool closed_ {};
};

[[nodiscard]] std::optional<std::uint64_t> estimate_fft_execution_bytes(std::uint32_t) noexcept;
[[nodiscard]] std::variant<std::unique_ptr<FftExecution>, CorrelationFailure> make_fft_execution(
    std::uint32_t count, std::uint64_t available_bytes);

struct CorrelationEstimate {
    Shape3 padded_shape;
    std::uint64_t padded_cells {}, working_bytes {};
};
using CorrelationEstimateOutcome = std::variant<CorrelationEstimate, CorrelationFailure>;
[[nodiscard]] CorrelationEstimateOutcome estimate_correlation_cpu(const CorrelationSpec&, std::uint32_t fft_threads = 1);

[[nodiscard]] CorrelationOutcome correlate_binary_cpu(const CorrelationSpec&, std::span<const std::uint8_t> environment,
                                                      std::span<const std::uint8_t> kernel,
                                                      const CorrelationLimits& = {},
                                                      const runtime::OperationControl& = {}, FftExecution* = nullptr);

[[nodiscard]] CorrelationOutcome correlate_proximity_cpu(const CorrelationSpec&, std::span<const double> environment,
                                                         std::span<const std::uint8_t> kernel,
                                                         const CorrelationLimits& = {},
                                                         const runtime::OperationControl& = {},
                                                         FftExecution* = nullptr);

// The same calls under their device names.
inline CorrelationOutcome correlate_binary_hip(const CorrelationSpec& s, std::span<const std::uint8_t> e,
                                               std::span<const std::uint8_t> k, const CorrelationLimits& l = {},
                                               const runtime::OperationControl& c = {}, FftExecution* x = nullptr)
{
    return correlate_binary_cpu(s, e, k, l, c, x);
}
inline CorrelationOutcome correlate_proximity_hip(const CorrelationSpec& s, std::span<const double> e,
                                                  std::span<const std::uint8_t> k, const CorrelationLimits& l = {},
                                                  const runtime::OperationControl& c = {}, FftExecution* x = nullptr)
{
    return correlate_proximity_cpu(s, e, k, l, c, x);
}

}  // namespace spectrapack::compute
