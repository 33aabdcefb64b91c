// Internal host-side plumbing shared by the HIP translation units: the context
// object behind sp_hip_ctx, device buffers, FFT length plans and twiddles.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/spectrapack_hip.h"
#include "sp_fft.hpp"

namespace sphip {

struct DeviceError {
    std::string what;
};

#define SP_HIP_CHECK(expr)                                                                                    \
    do {                                                                                                      \
        hipError_t sp_err_ = (expr);                                                                          \
        if (sp_err_ != hipSuccess)                                                                            \
            throw ::sphip::DeviceError { std::string(#expr) + ": " + hipGetErrorString(sp_err_) };            \
    } while (0)

// Owned device allocation.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(std::size_t bytes) { allocate(bytes); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) { o.ptr_ = nullptr, o.bytes_ = 0; }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
    {
        if (this != &o) {
            release();
            ptr_ = o.ptr_;
            bytes_ = o.bytes_;
            o.ptr_ = nullptr;
            o.bytes_ = 0;
        }
        return *this;
    }
    void allocate(std::size_t bytes)
    {
        release();
        if (bytes == 0) return;
        SP_HIP_CHECK(hipMalloc(&ptr_, bytes));
        bytes_ = bytes;
    }
    // Grow-only reuse.
    void ensure(std::size_t bytes)
    {
        if (bytes > bytes_) allocate(bytes);
    }
    void release()
    {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        bytes_ = 0;
    }
    template <typename T> T* as() const { return static_cast<T*>(ptr_); }
    void* get() const { return ptr_; }
    std::size_t bytes() const { return bytes_; }

private:
    void* ptr_ {};
    std::size_t bytes_ {};
};

// Largest FFT length the LDS tiles support (per precision).
constexpr int kMaxFftLength = 2048;

// Factor n into radices 8/4/2/3/5 (8 first); false when n has another prime.
bool make_plan(int n, FftPlan& plan);
// Smallest m >= n with only factors 2, 3, 5 (m <= kMaxFftLength), or 0.
int good_length(int n);

// Per-device, per-precision twiddle tables exp(-2 pi i k/n), k in [0, n).
template <typename T> const cx<T>* twiddles(int n);
// The same values on the host (n entries).
template <typename T> void fill_twiddles(int n, cx<T>* out);

// Address of the allocation-free correlation's static scratch on `device`
// (the current device), looked up once per device; SP_OK or a status.
int init_noalloc_scratch(int device);

struct Context {
    int device {};
    hipStream_t stream {};
    std::uint64_t bytes_cap {};
    std::string last_error;
    // scratch reused across correlate calls
    DeviceBuffer scratch[6];
    DeviceBuffer pinned_flag;
    // autotuned (radix, tile width) per pass and shape, shared by sessions
    std::map<std::string, std::pair<int, int>> tiles;
};

// Apply the context's device for the current host thread.
void use_device(const Context& c);

// validation_kernel.cpp:3492-3522: an offset d (cells) can reach within the
// clearance unless pitch^2 * sum(max(|d_a| - 1, 0)^2), rounded down, exceeds
// clearance^2, rounded up.
inline bool offset_may_reach(const int d[3], double pitch, double clearance)
{
    std::uint64_t sum = 0;
    for (int a = 0; a != 3; ++a) {
        const std::uint64_t m = std::uint64_t(d[a] < 0 ? -d[a] : d[a]);
        const std::uint64_t g = m > 1 ? m - 1 : 0;
        sum += g * g;
    }
    const double p2 = std::nextafter(pitch * pitch, -HUGE_VAL);
    const double dist = std::nextafter(p2 * static_cast<double>(sum), -HUGE_VAL);
    const double c2 = std::nextafter(clearance * clearance, HUGE_VAL);
    return !(dist > c2);
}

// Clearance halo in cells (conservative_fields.cpp:189-204): 0 for no clearance.
inline int clearance_halo_cells(double clearance, double pitch)
{
    if (clearance == 0.0) return 0;
    return int(std::ceil(static_cast<long double>(clearance) / pitch) + 1.0L);
}

// A-priori FP32 bound (see the header).
double fp32_error_bound(std::uint64_t fft_cells, std::uint64_t env_mass, std::uint64_t kernel_mass);

}  // namespace sphip

struct sp_hip_ctx {
    sphip::Context c;
};
