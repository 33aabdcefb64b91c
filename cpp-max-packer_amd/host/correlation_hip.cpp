// spectrapack::compute over the C ABI: the drop-in of
// engine/pack_compute/src/correlation.cpp:596-655 (estimate, binary and
// proximity correlation) and fft_execution.cpp:224-249 (execution owner).
// Failures are values with the reference's static code strings; no exception
// escapes.
#include <cstdlib>
#include <type_traits>
#include <new>

#include "spectrapack/compute/correlation.hpp"
#include "spectrapack_hip.h"

namespace spectrapack::compute {
namespace {

std::string_view message_for(int status)
{
    switch (status) {
    case SP_CORRELATION_MEMORY_LIMIT: return "Correlation buffers and FFT workspace exceed their memory limit.";
    case SP_CORRELATION_SHAPE_OVERFLOW: return "Correlation extents must be positive and representable.";
    case SP_CORRELATION_INDEX_OVERFLOW: return "The first output translation is not representable.";
    case SP_CORRELATION_INVALID_FIELD: return "Shape does not match supplied cells.";
    case SP_CORRELATION_INVALID_ENVIRONMENT: return "Environment cells must be finite values in [0,1].";
    case SP_CORRELATION_INVALID_KERNEL: return "Kernel cells must be binary.";
    case SP_CORRELATION_NUMERIC_RANGE: return "Binary correlation mass is not exactly representable.";
    case SP_CORRELATION_PADDED_CELL_LIMIT: return "Padded field exceeds its cell limit.";
    case SP_CORRELATION_NUMERIC: return "FFT output failed its integer, range, mass or direct check.";
    case SP_CORRELATION_DIRECT_TERM_LIMIT: return "Deterministic direct checks exhausted their work limit.";
    case SP_CORRELATION_ALLOCATION: return "Correlation allocation failed.";
    case SP_CORRELATION_EXECUTION_CLOSED: return "The device execution failed earlier and is closed.";
    case SP_FFT_LENGTH_UNSUPPORTED: return "A padded FFT length exceeds the device kernels' maximum.";
    case SP_DEVICE_ERROR: return "A HIP runtime call failed.";
    default: return "Correlation operation failed.";
    }
}

CorrelationFailure failure(int status, CorrelationStats stats = {})
{
    // Device-side codes without a reference counterpart surface as the
    // reference's generic operation failure.
    const int mapped = (status == SP_DEVICE_ERROR || status == SP_INVALID_ARGUMENT || status == SP_CAPACITY)
                           ? SP_CORRELATION_OPERATION
                           : status == SP_FFT_LENGTH_UNSUPPORTED ? SP_CORRELATION_PADDED_CELL_LIMIT : status;
    return CorrelationFailure { sp_hip_status_code(mapped), message_for(status), stats };
}

std::optional<CorrelationFailure> interrupted(const runtime::OperationControl& control)
{
    const auto cause = control.poll();
    if (cause == runtime::StopCause::none) return std::nullopt;
    return CorrelationFailure { cause == runtime::StopCause::user_stopped ? "OPERATION_CANCELLED" : "DEADLINE_EXCEEDED",
                                "Correlation interrupted.", {} };
}

int default_device()
{
    const char* v = std::getenv("SPECTRAPACK_HIP_DEVICE");
    return v ? std::atoi(v) : 0;
}

// A per-thread execution for callers that pass none (the reference runs a
// serial execution in that case, fft_execution.cpp:212-223).
FftExecution* thread_execution(CorrelationFailure& error)
{
    thread_local std::unique_ptr<FftExecution> owned;
    if (!owned) {
        auto made = FftExecution::open(default_device());
        if (auto* f = std::get_if<CorrelationFailure>(&made)) {
            error = *f;
            return nullptr;
        }
        owned = std::move(std::get<std::unique_ptr<FftExecution>>(made));
    }
    return owned.get();
}


// The reference counts these objects of its caller in every admission
// (correlation.cpp:67-72); the mirror's types have the reference's layout.
constexpr std::uint64_t kMetadataBytes = 2 * sizeof(CorrelationOutcome) + sizeof(std::vector<double>) +
                                         sizeof(runtime::OperationControl);

sp_correlation_spec to_c(const CorrelationSpec& s)
{
    sp_correlation_spec c {};
    for (int a = 0; a != 3; ++a) {
        c.environment_shape[a] = s.environment_shape[a];
        c.kernel_shape[a] = s.kernel_shape[a];
        c.environment_first[a] = s.environment_first[a];
        c.kernel_first[a] = s.kernel_first[a];
    }
    return c;
}

template <typename E>
CorrelationOutcome correlate(const CorrelationSpec& spec, std::span<const E> environment,
                             std::span<const std::uint8_t> kernel, const CorrelationLimits& limits,
                             const runtime::OperationControl& control, FftExecution* execution)
{
    if (auto i = interrupted(control)) return *i;
    CorrelationFailure error {};
    FftExecution* exec = execution ? execution : thread_execution(error);
    if (!exec) return error;
    if (exec->closed()) return failure(SP_CORRELATION_EXECUTION_CLOSED);
    const sp_correlation_spec c = to_c(spec);
    const sp_correlation_limits l { limits.max_working_bytes, limits.reserved_bytes, limits.max_padded_cells,
                                    limits.max_direct_terms, kMetadataBytes };
    sp_correlation_estimate est {};
    std::uint64_t cells = 0;
    if (sp_hip_estimate(&c, 1, kMetadataBytes, &est) == SP_OK) cells = est.padded_cells;
    CorrelationResult result;
    sp_numeric_report report {};
    sp_correlation_stats stats {};
    std::int64_t first[3];
    std::uint32_t shape[3];
    try {
        if (cells && cells <= limits.max_padded_cells) result.values.resize(cells);
    }
    catch (const std::bad_alloc&) {
        return failure(SP_CORRELATION_ALLOCATION);
    }
    control.phase(runtime::Phase::planning_fft);
    int rc;
    if constexpr (std::is_same_v<E, std::uint8_t>)
        rc = sp_hip_correlate_binary(exec->device_context(), &c, environment.data(), environment.size(), kernel.data(),
                                     kernel.size(), &l, SP_PRECISION_AUTO, result.values.data(), result.values.size(),
                                     first, shape, &report, &stats);
    else
        rc = sp_hip_correlate_proximity(exec->device_context(), &c, environment.data(), environment.size(),
                                        kernel.data(), kernel.size(), &l, result.values.data(), result.values.size(),
                                        first, shape, &report, &stats);
    const CorrelationStats st { stats.padded_cells, stats.working_bytes_peak, stats.direct_terms };
    if (rc == SP_DEVICE_ERROR) exec->close();  // a failed execution cannot be reused
    if (rc != SP_OK) {
        if (auto i = interrupted(control)) {
            i->stats = st;
            return *i;
        }
        return failure(rc, st);
    }
    if (auto i = interrupted(control)) {
        i->stats = st;
        return *i;
    }
    for (int a = 0; a != 3; ++a) {
        result.translation_first[a] = first[a];
        result.shape[a] = shape[a];
    }
    result.numeric = { report.max_integer_residual, report.mass_residual, report.max_probe_error };
    result.stats = st;
    return result;
}

}  // namespace

std::variant<std::unique_ptr<FftExecution>, CorrelationFailure> FftExecution::open(int device)
{
    sp_hip_ctx* ctx = nullptr;
    const int rc = sp_hip_ctx_create(device, 0, &ctx);
    if (rc != SP_OK) return failure(rc);
    return std::unique_ptr<FftExecution>(new FftExecution(ctx));
}

FftExecution::~FftExecution() { sp_hip_ctx_destroy(ctx_); }

std::optional<std::uint64_t> estimate_fft_execution_bytes(std::uint32_t count) noexcept
{
    if (count < 1 || count > 8) return std::nullopt;
    return std::uint64_t { 0 };  // device buffers are owned by the HIP context
}

std::variant<std::unique_ptr<FftExecution>, CorrelationFailure> make_fft_execution(std::uint32_t count,
                                                                                   std::uint64_t)
{
    if (count < 1 || count > 8)
        return CorrelationFailure { "CPU_THREAD_COUNT_UNSUPPORTED", "Unsupported FFT thread count.", {} };
    return FftExecution::open(default_device());
}

CorrelationEstimateOutcome estimate_correlation_cpu(const CorrelationSpec& spec, std::uint32_t fft_threads)
{
    if (!estimate_fft_execution_bytes(fft_threads))
        return CorrelationFailure { "CPU_THREAD_COUNT_UNSUPPORTED", "Unsupported FFT thread count.", {} };
    const sp_correlation_spec c = to_c(spec);
    sp_correlation_estimate e {};
    const int rc = sp_hip_estimate(&c, fft_threads, kMetadataBytes, &e);
    if (rc != SP_OK) return failure(rc);
    CorrelationEstimate out;
    for (int a = 0; a != 3; ++a) out.padded_shape[a] = e.padded_shape[a];
    out.padded_cells = e.padded_cells;
    out.working_bytes = e.working_bytes;
    return out;
}

CorrelationOutcome correlate_binary_cpu(const CorrelationSpec& spec, std::span<const std::uint8_t> environment,
                                        std::span<const std::uint8_t> kernel, const CorrelationLimits& limits,
                                        const runtime::OperationControl& control, FftExecution* execution)
{
    return correlate(spec, environment, kernel, limits, control, execution);
}

CorrelationOutcome correlate_proximity_cpu(const CorrelationSpec& spec, std::span<const double> environment,
                                           std::span<const std::uint8_t> kernel, const CorrelationLimits& limits,
                                           const runtime::OperationControl& control, FftExecution* execution)
{
    return correlate(spec, environment, kernel, limits, control, execution);
}

}  // namespace spectrapack::compute
