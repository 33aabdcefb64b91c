// Context, FFT length plans, twiddle tables, the FP32 error bound and the
// status-code table of the C ABI.
#include <cmath>
#include <cstring>

#include "common.hpp"

namespace sphip {

bool make_plan(int n, FftPlan& plan)
{
    plan = FftPlan {};
    plan.n = n;
    if (n < 1 || n > kMaxFftLength) return false;
    int m = n;
    while (m % 8 == 0) {
        plan.radix[plan.count++] = 8;
        m /= 8;
    }
    while (m % 4 == 0) {
        plan.radix[plan.count++] = 4;
        m /= 4;
    }
    while (m % 2 == 0) {
        plan.radix[plan.count++] = 2;
        m /= 2;
    }
    while (m % 3 == 0) {
        plan.radix[plan.count++] = 3;
        m /= 3;
    }
    while (m % 5 == 0) {
        plan.radix[plan.count++] = 5;
        m /= 5;
    }
    if (m != 1 || plan.count > kMaxFactors) return false;
    return true;
}

int good_length(int n)
{
    for (int m = std::max(n, 1); m <= kMaxFftLength; ++m) {
        int r = m;
        for (int p : { 2, 3, 5 })
            while (r % p == 0) r /= p;
        if (r == 1) return m;
    }
    return 0;
}

namespace {
std::mutex g_tw_mu;
std::map<std::tuple<int, int, int>, DeviceBuffer>& tw_cache()
{
    static auto* m = new std::map<std::tuple<int, int, int>, DeviceBuffer>();
    return *m;
}
}  // namespace

template <typename T> const cx<T>* twiddles(int n)
{
    int dev = 0;
    SP_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_tw_mu);
    auto key = std::make_tuple(dev, n, int(sizeof(T)));
    auto& cache = tw_cache();
    auto it = cache.find(key);
    if (it != cache.end()) return it->second.as<cx<T>>();
    std::vector<cx<T>> host(n);
    fill_twiddles<T>(n, host.data());
    DeviceBuffer buf(sizeof(cx<T>) * n);
    SP_HIP_CHECK(hipMemcpy(buf.get(), host.data(), sizeof(cx<T>) * n, hipMemcpyHostToDevice));
    const cx<T>* p = buf.as<cx<T>>();
    cache.emplace(key, std::move(buf));
    return p;
}

template const cx<float>* twiddles<float>(int);
template const cx<double>* twiddles<double>(int);

template <typename T> void fill_twiddles(int n, cx<T>* out)
{
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < n; ++k) {
        const long double a = 2.0L * pi * static_cast<long double>(k) / static_cast<long double>(n);
        out[k].x = T(std::cos(a));
        out[k].y = T(-std::sin(a));
    }
}

template void fill_twiddles<float>(int, cx<float>*);
template void fill_twiddles<double>(int, cx<double>*);

void use_device(const Context& c) { SP_HIP_CHECK(hipSetDevice(c.device)); }

// A-priori bound on |computed - exact| of one overlap count when the
// correlation runs as three FP32 FFT stages (forward B, forward K, inverse),
// under the probabilistic rounding model of Higham & Mary (2019, "A new
// approach to probabilistic rounding error analysis"): each of the
// L = log2(N) butterfly levels of the three transforms rounds every value
// with an independent, mean-zero relative error bounded by c*u (c = 2 covers
// the complex multiply-add and the rounded twiddle), and the transforms
// preserve root-mean-square magnitude (Parseval).  The output error of one
// count then has standard deviation at most c*u*sqrt(3L)*rms(C), where
// rms(C)^2 = sum C^2 / N <= min(mB, mK) * mB * mK / N (every count is at most
// min(mB, mK) and all counts sum to mB*mK).  A Gaussian tail with
// lambda = sqrt(2 ln(2N / 2^-40)) bounds all N outputs at once except with
// probability 2^-40.
double fp32_error_bound(std::uint64_t fft_cells, std::uint64_t env_mass, std::uint64_t kernel_mass)
{
    if (fft_cells == 0 || kernel_mass == 0 || env_mass == 0) return 0.0;
    const double u = std::ldexp(1.0, -24);
    const double c = 2.0;
    const double N = double(fft_cells);
    const double L = std::max(1.0, std::log2(N));
    const double mb = double(env_mass), mk = double(kernel_mass);
    const double rms = std::sqrt(std::min(mb, mk) * mb * mk / N);
    const double lambda = std::sqrt(2.0 * (std::log(2.0 * N) + 40.0 * std::log(2.0)));
    return lambda * c * u * std::sqrt(3.0 * L) * rms;
}

}  // namespace sphip

extern "C" {

const char* sp_hip_status_code(int status)
{
    switch (status) {
    case SP_OK: return "OK";
    case SP_CORRELATION_MEMORY_LIMIT: return "CORRELATION_MEMORY_LIMIT";
    case SP_CORRELATION_SHAPE_OVERFLOW: return "CORRELATION_SHAPE_OVERFLOW";
    case SP_CORRELATION_INDEX_OVERFLOW: return "CORRELATION_INDEX_OVERFLOW";
    case SP_CORRELATION_INVALID_FIELD: return "CORRELATION_INVALID_FIELD";
    case SP_CORRELATION_INVALID_ENVIRONMENT: return "CORRELATION_INVALID_ENVIRONMENT";
    case SP_CORRELATION_INVALID_KERNEL: return "CORRELATION_INVALID_KERNEL";
    case SP_CORRELATION_NUMERIC_RANGE: return "CORRELATION_NUMERIC_RANGE";
    case SP_CORRELATION_PADDED_CELL_LIMIT: return "CORRELATION_PADDED_CELL_LIMIT";
    case SP_CORRELATION_NUMERIC: return "CORRELATION_NUMERIC";
    case SP_CORRELATION_DIRECT_TERM_LIMIT: return "CORRELATION_DIRECT_TERM_LIMIT";
    case SP_CORRELATION_ALLOCATION: return "CORRELATION_ALLOCATION";
    case SP_CORRELATION_OPERATION: return "CORRELATION_OPERATION";
    case SP_CORRELATION_FLOATING_ENVIRONMENT: return "CORRELATION_FLOATING_ENVIRONMENT";
    case SP_CORRELATION_EXECUTION_CLOSED: return "CORRELATION_EXECUTION_CLOSED";
    case SP_CPU_THREAD_COUNT_UNSUPPORTED: return "CPU_THREAD_COUNT_UNSUPPORTED";
    case SP_FFT_THREAD_START_FAILURE: return "FFT_THREAD_START_FAILURE";
    case SP_OPERATION_CANCELLED: return "OPERATION_CANCELLED";
    case SP_DEADLINE_EXCEEDED: return "DEADLINE_EXCEEDED";
    case SP_DEVICE_ERROR: return "DEVICE_ERROR";
    case SP_INVALID_ARGUMENT: return "INVALID_ARGUMENT";
    case SP_FIELD_INPUT: return "FIELD_INPUT";
    case SP_FIELD_INDEX_OVERFLOW: return "FIELD_INDEX_OVERFLOW";
    case SP_FFT_LENGTH_UNSUPPORTED: return "FFT_LENGTH_UNSUPPORTED";
    case SP_CAPACITY: return "CAPACITY";
    case SP_SCRATCH_LIMIT: return "SCRATCH_LIMIT";
    case SP_FIELD_COPY_ID: return "FIELD_COPY_ID";
    case SP_FIELD_COUNT_UNDERFLOW: return "FIELD_COUNT_UNDERFLOW";
    default: return "UNKNOWN";
    }
}

int sp_hip_ctx_create(int device, uint64_t bytes_cap, sp_hip_ctx** out)
{
    if (!out) return SP_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return SP_DEVICE_ERROR;
    auto* ctx = new (std::nothrow) sp_hip_ctx;
    if (!ctx) return SP_CORRELATION_ALLOCATION;
    ctx->c.device = device;
    ctx->c.bytes_cap = bytes_cap;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return SP_DEVICE_ERROR;
    }
    if (const int rc = sphip::init_noalloc_scratch(device); rc != SP_OK) {
        ctx->c.last_error = "hipGetSymbolAddress of the allocation-free scratch failed";
        (void)hipStreamDestroy(ctx->c.stream);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return SP_OK;
}

void sp_hip_ctx_destroy(sp_hip_ctx* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    for (auto& b : ctx->c.scratch) b.release();
    (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
}

const char* sp_hip_last_error(const sp_hip_ctx* ctx) { return ctx ? ctx->c.last_error.c_str() : ""; }

double sp_hip_fp32_error_bound(uint64_t fft_cells, uint64_t env_mass, uint64_t kernel_mass)
{
    return sphip::fp32_error_bound(fft_cells, env_mass, kernel_mass);
}

}  // extern "C"
