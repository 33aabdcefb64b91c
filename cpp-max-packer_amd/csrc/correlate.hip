// Full-linear checked correlation on gfx950: the drop-in for
// correlate_binary_cpu / correlate_proximity_cpu
// (engine/pack_compute/src/correlation.cpp:238-592, :643-655).
//
// The reference transforms at P = E + K - 1 per axis.  Any FFT length
// P' >= E + K - 1 yields the same linear correlation (no circular alias), so
// the device transforms at the next 2^a 3^b 5^c length and reads the linear
// output at circular slot (x - (K - 1)) mod P'.  Every check of the reference
// runs on the device: finite/range/residual per value, total mass, and eight
// deterministic direct probes.
//
// sp_hip_correlate_binary_noalloc runs the same kernels on the static device
// array `scratch` below instead of allocated buffers, for callers that cannot
// allocate device memory.
#include <algorithm>
#include <array>
#include <cstring>
#include <limits>
#include <mutex>

#include "common.hpp"
#include "fft_kernels.hpp"

// Working memory of the allocation-free correlation: one copy per GPU, held
// from the library's first use on that GPU until the process ends.  Host code
// never uses this name as an address; it uses the per-device pointer that
// hipGetSymbolAddress returns (init_noalloc_scratch).
static __device__ __align__(256) unsigned char scratch[SP_NOALLOC_SCRATCH_BYTES];

namespace sphip {
namespace {

constexpr int kThreads = 256;
constexpr std::size_t kLdsBudget = 64 * 1024;
constexpr std::uint64_t kExactInteger = 1ULL << 53;

int choose_w(int n, std::size_t elem, int buffers, bool pairs, int cap, int min_w)
{
    int best = min_w;
    for (int w = min_w; w <= cap; w *= 2) {
        const int pencils = pairs ? w / 2 : w;
        if (elem * (std::size_t(n) + std::size_t(buffers) * n * (pencils + 1)) <= kLdsBudget) best = w;
    }
    return best;
}
template <typename T> std::size_t lds_bytes(int n, int buffers, int pencils)
{
    return sizeof(cx<T>) * (std::size_t(n) + std::size_t(buffers) * n * (pencils + 1));
}

struct Stats {
    unsigned long long ones_env;    // binary: ones; proximity: unused
    unsigned long long ones_kernel;
    unsigned int bad_env, bad_kernel;
    double env_sum;                 // proximity environment mass
};

template <typename E>
__global__ void k_validate(const E* __restrict__ env, long long ne, const unsigned char* __restrict__ ker, long long nk,
                           Stats* st)
{
    unsigned long long ones = 0, kones = 0;
    unsigned int bad = 0, kbad = 0;
    double sum = 0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < ne; i += stride) {
        const double v = double(env[i]);
        if (!(v >= 0.0 && v <= 1.0)) bad = 1;  // also catches NaN
        sum += v;
        ones += (v == 1.0);
    }
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nk; i += stride) {
        if (ker[i] > 1) kbad = 1;
        kones += ker[i];
    }
    for (int off = 32; off > 0; off >>= 1) {
        ones += __shfl_down(ones, off, 64);
        kones += __shfl_down(kones, off, 64);
        sum += __shfl_down(sum, off, 64);
        bad |= __shfl_down(bad, off, 64);
        kbad |= __shfl_down(kbad, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (ones) atomicAdd(&st->ones_env, ones);
        if (kones) atomicAdd(&st->ones_kernel, kones);
        if (bad) atomicOr(&st->bad_env, bad);
        if (kbad) atomicOr(&st->bad_kernel, kbad);
        if (sum != 0.0) atomicAdd(&st->env_sum, sum);
    }
}

// uint8 / double environment scattered into a zero-padded real grid is not
// needed: the R2C pass reads the environment in place with zero padding.

struct Readout {
    unsigned long long sum_rounded;  // binary: exact integer mass of the output
    double sum_raw;                  // sum of raw values (report)
    unsigned long long max_res_bits; // double bits of the max residual
    unsigned long long argmax;       // (magnitude key << 32) | ~index
    unsigned int flags;              // 1 nonfinite, 2 residual, 4 range
    unsigned int pad;
};

// Clears `words` 32-bit words (the check accumulators of a call), so that
// memory reused from an earlier call starts at zero.
__global__ void k_zero_words(unsigned int* p, int words)
{
    for (int i = threadIdx.x; i < words; i += blockDim.x) p[i] = 0u;
}

// Values of the full-linear output (shape ox, oy, oz) from the circular field
// R (shape px, py, pz) at slot (x - (K - 1)) mod P'; rounding and the
// per-value checks of correlation.cpp:494-540.
template <typename T>
__global__ void k_readout(const T* __restrict__ R, int px, int py, int pz, int ox, int oy, int oz, int kx, int ky, int kz,
                          int binary, double max_out, double tol, double* __restrict__ values, Readout* rd)
{
    const long long n = (long long)ox * oy * oz;
    double raw = 0, maxres = 0;
    unsigned long long rsum = 0, best = 0;
    unsigned int flags = 0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) {
        const int x = int(i % ox), y = int((i / ox) % oy), z = int(i / ((long long)ox * oy));
        const int tx = (x - (kx - 1) + px) % px, ty = (y - (ky - 1) + py) % py, tz = (z - (kz - 1) + pz) % pz;
        const double v = double(R[tx + (long long)px * (ty + (long long)py * tz)]);
        if (!isfinite(v)) {
            flags |= 1u;
            values[i] = 0;
            continue;
        }
        raw += v;
        double out;
        unsigned long long mag;
        if (binary) {
            const double r = round(v);
            const double res = fabs(v - r);
            maxres = fmax(maxres, res);
            if (!(res < 0.25)) flags |= 2u;
            if (r < 0.0 || r > max_out) flags |= 4u;
            out = r;
            if (r >= 0.0) rsum += (unsigned long long)r;
            mag = (unsigned long long)fmax(r, 0.0);
        }
        else {
            if (v < -tol || v > max_out + tol) flags |= 4u;
            out = v;
            mag = (unsigned long long)__float_as_uint(float(fabs(v)));
        }
        values[i] = out;
        const unsigned long long key = (mag << 32) | (0xffffffffULL - (unsigned long long)i);
        best = key > best ? key : best;
    }
    for (int off = 32; off > 0; off >>= 1) {
        raw += __shfl_down(raw, off, 64);
        rsum += __shfl_down(rsum, off, 64);
        maxres = fmax(maxres, __shfl_down(maxres, off, 64));
        const unsigned long long ob = __shfl_down(best, off, 64);
        best = ob > best ? ob : best;
        flags |= __shfl_down(flags, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&rd->sum_raw, raw);
        atomicAdd(&rd->sum_rounded, rsum);
        atomicMax(&rd->max_res_bits, (unsigned long long)__double_as_longlong(maxres));
        atomicMax(&rd->argmax, best);
        if (flags) atomicOr(&rd->flags, flags);
    }
}

// Direct probes (correlation.cpp:153-218): the exact overlap sum of one output
// index, one block per probe.
template <typename E>
__global__ void k_probe(const E* __restrict__ env, int ex, int ey, int ez, const unsigned char* __restrict__ ker, int kx,
                        int ky, int kz, int ox, int oy, const long long* __restrict__ probes, double* __restrict__ out)
{
    const long long idx = probes[blockIdx.x];
    const int x = int(idx % ox), y = int((idx / ox) % oy), z = int(idx / ((long long)ox * oy));
    const int sx = x - (kx - 1), sy = y - (ky - 1), sz = z - (kz - 1);
    const long long nk = (long long)kx * ky * kz;
    double sum = 0;
    for (long long i = threadIdx.x; i < nk; i += blockDim.x) {
        const int a = int(i % kx), b = int((i / kx) % ky), c = int(i / ((long long)kx * ky));
        const int e0 = a + sx, e1 = b + sy, e2 = c + sz;
        if (e0 < 0 || e1 < 0 || e2 < 0 || e0 >= ex || e1 >= ey || e2 >= ez) continue;
        sum += double(env[e0 + (long long)ex * (e1 + (long long)ey * e2)]) * double(ker[i]);
    }
    __shared__ double part[kThreads];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = part[0];
}

bool checked_mul(std::uint64_t a, std::uint64_t b, std::uint64_t& r)
{
    if (a != 0 && b > std::numeric_limits<std::uint64_t>::max() / a) return false;
    r = a * b;
    return true;
}
bool checked_add(std::uint64_t a, std::uint64_t b, std::uint64_t& r)
{
    if (b > std::numeric_limits<std::uint64_t>::max() - a) return false;
    r = a + b;
    return true;
}
bool checked_sub(std::int64_t a, std::int64_t b, std::int64_t& r)
{
    if ((b > 0 && a < std::numeric_limits<std::int64_t>::min() + b) ||
        (b < 0 && a > std::numeric_limits<std::int64_t>::max() + b))
        return false;
    r = a - b;
    return true;
}

struct Geometry {
    std::uint32_t padded[3];  // reference P = E + K - 1
    int fft[3];               // device P' >= P
    std::int64_t first[3];
    std::uint64_t padded_count, fft_cells;
};

// Metadata checks of correlation.cpp:281-314 and :596-641.
int geometry(const sp_correlation_spec& s, Geometry& g, std::uint64_t& env_count, std::uint64_t& ker_count)
{
    env_count = ker_count = g.padded_count = 1;
    for (int a = 0; a != 3; ++a) {
        if (s.environment_shape[a] == 0 || s.kernel_shape[a] == 0 ||
            !checked_mul(env_count, s.environment_shape[a], env_count) ||
            !checked_mul(ker_count, s.kernel_shape[a], ker_count))
            return SP_CORRELATION_SHAPE_OVERFLOW;
        const std::uint64_t len = std::uint64_t(s.environment_shape[a]) + s.kernel_shape[a] - 1;
        if (len > 0xffffffffULL || !checked_mul(g.padded_count, len, g.padded_count)) return SP_CORRELATION_SHAPE_OVERFLOW;
        g.padded[a] = std::uint32_t(len);
        std::int64_t diff;
        if (!checked_sub(s.environment_first[a], s.kernel_first[a], diff) ||
            !checked_sub(diff, std::int64_t(s.kernel_shape[a]) - 1, g.first[a]))
            return SP_CORRELATION_INDEX_OVERFLOW;
        if (g.first[a] > std::numeric_limits<std::int64_t>::max() - std::int64_t(len - 1)) return SP_CORRELATION_INDEX_OVERFLOW;
    }
    g.fft_cells = 1;
    for (int a = 0; a != 3; ++a) {
        g.fft[a] = g.padded[a] <= std::uint32_t(kMaxFftLength) ? good_length(int(g.padded[a])) : 0;
        g.fft_cells *= std::uint64_t(g.fft[a] ? g.fft[a] : 1);
    }
    return SP_OK;
}

constexpr std::uint64_t kAlign = 256;
constexpr int kMaxProbes = 8;
inline std::uint64_t align_up(std::uint64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// Byte offsets of one correlation's device buffers from one base.  The
// inputs and their validation counters come first, because their size does
// not depend on the precision, which is chosen after validation.  Then the two
// spectra, the real field, the output values, the check accumulators, the
// probes and, for allocation-free calls only, the three twiddle tables.
struct Layout {
    std::uint64_t env, ker, stats, inputs_end;
    std::uint64_t sb, sk, r, vals, readout, probe, probe_idx, total;
    std::uint64_t tw[3], total_with_twiddles;
};

// Sizes of the input part; false when they are not representable.
bool input_layout(std::uint64_t env_count, std::uint64_t ker_count, std::size_t env_elem, Layout& L)
{
    std::uint64_t env_bytes;
    if (!checked_mul(env_count, env_elem, env_bytes) || env_bytes > (1ULL << 62) || ker_count > (1ULL << 62))
        return false;
    L = Layout {};
    std::uint64_t o = 0;
    L.env = o;
    o += align_up(env_bytes);
    L.ker = o;
    o += align_up(ker_count);
    L.stats = o;
    o += align_up(sizeof(Stats));
    L.inputs_end = o;
    return true;
}

// The rest, once the FFT grid (each axis at most kMaxFftLength) and the
// precision (`real` bytes per real value) are known.
void compute_layout(const Geometry& g, std::size_t real, Layout& L)
{
    const std::uint64_t nh = std::uint64_t(g.fft[2] / 2 + 1);
    const std::uint64_t spec = nh * std::uint64_t(g.fft[1]) * std::uint64_t(g.fft[0]) * 2 * real;
    std::uint64_t o = L.inputs_end;
    L.sb = o;
    o += align_up(spec);
    L.sk = o;
    o += align_up(spec);
    L.r = o;
    o += align_up(g.fft_cells * real);
    L.vals = o;
    o += align_up(g.padded_count * sizeof(double));
    L.readout = o;
    o += align_up(sizeof(Readout));
    L.probe = o;
    o += align_up(kMaxProbes * sizeof(double));
    L.probe_idx = o;
    o += align_up(kMaxProbes * sizeof(long long));
    L.total = o;
    for (int a = 0; a != 3; ++a) {
        L.tw[a] = o;
        o += align_up(std::uint64_t(g.fft[a]) * 2 * real);
    }
    L.total_with_twiddles = o;
}

// Device bytes one correlation takes at `real` bytes per real value.
std::uint64_t device_bytes(const Geometry& g, std::uint64_t env_count, std::uint64_t ker_count, std::size_t real,
                           std::size_t env_elem)
{
    Layout L;
    if (!input_layout(env_count, ker_count, env_elem, L)) return std::numeric_limits<std::uint64_t>::max();
    compute_layout(g, real, L);
    return L.total;
}

// correlation.cpp:220-235: pocketfft's workspace bound for a one-axis call
// over the longest padded axis L, 64 complex<double> values per point plus
// a fixed 64 KiB.
bool pocketfft_workspace(std::uint64_t longest_axis, std::uint64_t& bytes)
{
    std::uint64_t values;
    return checked_mul(longest_axis, 64, values) && checked_mul(values, 16, bytes) &&
           checked_add(bytes, 64ULL << 10, bytes);
}

// estimate_correlation_cpu's working bytes (correlation.cpp:596-641): two
// complex<double> buffers and the double output per padded cell, one
// pocketfft workspace per FFT thread, the execution owner's reserve (0: the
// device context owns the device buffers), the caller's metadata objects and
// the probe index array of eight size_t.  False: not representable
// (CORRELATION_SHAPE_OVERFLOW).
bool reference_working_bytes(const Geometry& g, std::uint32_t fft_threads, std::uint64_t metadata, std::uint64_t& out)
{
    std::uint64_t ws;
    const std::uint64_t longest = std::max({ g.padded[0], g.padded[1], g.padded[2] });
    return checked_mul(g.padded_count, 2 * 16 + 8, out) && pocketfft_workspace(longest, ws) &&
           checked_mul(ws, fft_threads, ws) && checked_add(out, ws, out) && checked_add(out, metadata, out) &&
           checked_add(out, kMaxProbes * sizeof(std::uint64_t), out) &&
           g.padded_count <= std::uint64_t(std::numeric_limits<std::ptrdiff_t>::max()) / 16;
}

template <typename T> struct Buffers {
    cx<T>* SB;
    cx<T>* SK;
    T* R;
    double* vals;
    Readout* rd;
    double* probe;
    long long* probe_idx;
    const cx<T>* tw[3];
};

template <typename T> Buffers<T> buffers_at(unsigned char* base, const Layout& L, std::uint64_t origin)
{
    Buffers<T> b {};
    b.SB = reinterpret_cast<cx<T>*>(base + (L.sb - origin));
    b.SK = reinterpret_cast<cx<T>*>(base + (L.sk - origin));
    b.R = reinterpret_cast<T*>(base + (L.r - origin));
    b.vals = reinterpret_cast<double*>(base + (L.vals - origin));
    b.rd = reinterpret_cast<Readout*>(base + (L.readout - origin));
    b.probe = reinterpret_cast<double*>(base + (L.probe - origin));
    b.probe_idx = reinterpret_cast<long long*>(base + (L.probe_idx - origin));
    return b;
}

template <typename T, typename E>
int run(sphip::Context& C, const sp_correlation_spec& s, const Geometry& g, const E* denv, const unsigned char* dker,
        const Buffers<T>& B, bool binary, double max_out, double expected_mass, std::uint64_t expected_int_mass,
        const sp_correlation_limits& lim, double* host_values, sp_numeric_report& rep, sp_correlation_stats& st)
{
    using T2 = cx<T>;
    const int nx = g.fft[0], ny = g.fft[1], nz = g.fft[2];
    const int nh = nz / 2 + 1;
    FftPlan px, py, pz;
    if (!make_plan(nx, px) || !make_plan(ny, py) || !make_plan(nz, pz)) return SP_FFT_LENGTH_UNSUPPORTED;
    const T2* twx = B.tw[0];
    const T2* twy = B.tw[1];
    const T2* twz = B.tw[2];
    T2* SB = B.SB;
    T2* SK = B.SK;
    T* R = B.R;
    double* dvals = B.vals;
    Readout* rd = B.rd;
    double* dprobe = B.probe;
    long long* dprobe_idx = B.probe_idx;
    hipStream_t st_ = C.stream;
    const std::size_t e = sizeof(T2);
    // forward transforms of both fields
    for (int which = 0; which != 2; ++which) {
        T2* S = which == 0 ? SB : SK;
        const int W = choose_w(nz, e, 2, true, 32, 2);
        if (which == 0)
            hipLaunchKernelGGL((k_r2c_z<T, E>), dim3((nx + W - 1) / W, ny, 1), dim3(kThreads), lds_bytes<T>(nz, 2, W / 2),
                               st_, denv, 0LL, int(s.environment_shape[0]), int(s.environment_shape[1]),
                               int(s.environment_shape[0]), int(s.environment_shape[1]), int(s.environment_shape[2]), S,
                               0LL, nx, ny, nx, ny, nullptr, nullptr, twz, pz, W);
        else
            hipLaunchKernelGGL((k_r2c_z<T, unsigned char>), dim3((nx + W - 1) / W, ny, 1), dim3(kThreads),
                               lds_bytes<T>(nz, 2, W / 2), st_, dker, 0LL, int(s.kernel_shape[0]), int(s.kernel_shape[1]),
                               int(s.kernel_shape[0]), int(s.kernel_shape[1]), int(s.kernel_shape[2]), S, 0LL, nx, ny, nx,
                               ny, nullptr, nullptr, twz, pz, W);
        const int Wy = choose_w(ny, e, 2, false, 16, 1);
        hipLaunchKernelGGL((k_c2c<T>), dim3((nx + Wy - 1) / Wy, nh, 1), dim3(kThreads), lds_bytes<T>(ny, 2, Wy), st_, S, 0LL,
                           (long long)nx, 1LL, (long long)nx * ny, S, 0LL, (long long)nx, 1LL, (long long)nx * ny, nx, ny,
                           ny, nullptr, 0, nullptr, twy, py, Wy, 1);
        if (which == 0) {
            const int Wx = choose_w(nx, e, 2, false, 16, 1);
            hipLaunchKernelGGL((k_c2c<T>), dim3((ny + Wx - 1) / Wx, nh, 1), dim3(kThreads), lds_bytes<T>(nx, 2, Wx), st_, S,
                               0LL, 1LL, (long long)nx, (long long)nx * ny, S, 0LL, 1LL, (long long)nx,
                               (long long)nx * ny, ny, nx, nx, nullptr, 0, nullptr, twx, px, Wx, 1);
        }
    }
    // kernel x transform + product with conj + inverse x, into SK
    {
        const int W = choose_w(nx, e, 4, false, 16, 1);
        hipLaunchKernelGGL((k_mul_x<T>), dim3((ny + W - 1) / W, nh, 1), dim3(kThreads), lds_bytes<T>(nx, 4, W), st_, SK, 0LL,
                           nx, ny, SB, nx, ny, SK, 0LL, nullptr, 0, nx, nullptr, twx, px, W);
    }
    {
        const int W = choose_w(ny, e, 2, false, 16, 1);
        hipLaunchKernelGGL((k_c2c<T>), dim3((nx + W - 1) / W, nh, 1), dim3(kThreads), lds_bytes<T>(ny, 2, W), st_, SK, 0LL,
                           (long long)nx, 1LL, (long long)nx * ny, SK, 0LL, (long long)nx, 1LL, (long long)nx * ny, nx,
                           ny, ny, nullptr, 0, nullptr, twy, py, W, 0);
        const int W3 = choose_w(nz, e, 2, true, 32, 2);
        const T scale = T(1.0 / (double(nx) * double(ny) * double(nz)));
        hipLaunchKernelGGL((k_c2r_z_store<T>), dim3((nx + W3 - 1) / W3, ny, 1), dim3(kThreads), lds_bytes<T>(nz, 2, W3 / 2),
                           st_, SK, nx, ny, scale, R, twz, pz, W3);
    }
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st_, reinterpret_cast<unsigned int*>(rd),
                       int(sizeof(Readout) / sizeof(unsigned int)));
    SP_HIP_CHECK(hipGetLastError());
    const double prox_tol = 1e-10 * std::max(1.0, max_out);
    hipLaunchKernelGGL((k_readout<T>), dim3(1024), dim3(kThreads), 0, st_, R, nx, ny, nz, int(g.padded[0]),
                       int(g.padded[1]), int(g.padded[2]), int(s.kernel_shape[0]), int(s.kernel_shape[1]),
                       int(s.kernel_shape[2]), binary ? 1 : 0, max_out, prox_tol, dvals, rd);
    Readout r {};
    SP_HIP_CHECK(hipMemcpyAsync(&r, rd, sizeof r, hipMemcpyDeviceToHost, st_));
    SP_HIP_CHECK(hipStreamSynchronize(st_));
    double max_res;
    std::memcpy(&max_res, &r.max_res_bits, sizeof max_res);
    rep.max_integer_residual = binary ? max_res : 0.0;
    if (r.flags) return SP_CORRELATION_NUMERIC;
    // total mass (correlation.cpp:542-547): exact on the rounded integers for
    // the binary field; within 1e-10 relative for proximity
    rep.mass_residual = std::abs(r.sum_raw - expected_mass);
    if (binary) {
        if (r.sum_rounded != expected_int_mass) return SP_CORRELATION_NUMERIC;
    }
    else if (!(rep.mass_residual <= 1e-10 * std::max(1.0, std::abs(expected_mass))))
        return SP_CORRELATION_NUMERIC;
    // deterministic probes (correlation.cpp:549-577)
    const std::uint64_t n = g.padded_count;
    const std::uint64_t max_index = 0xffffffffULL - (r.argmax & 0xffffffffULL);
    std::array<long long, kMaxProbes> probes {};
    int nprobe = 0;
    for (std::uint64_t c : { std::uint64_t(0), n - 1, n / 2, max_index, n / 4, n / 3, (n / 3) * 2, (n / 4) * 3 })
        if (std::find(probes.begin(), probes.begin() + nprobe, (long long)c) == probes.begin() + nprobe &&
            nprobe < kMaxProbes)
            probes[nprobe++] = (long long)c;
    // direct-term budget: each probe visits its in-range kernel cells
    std::uint64_t terms = 0;
    for (int i = 0; i != nprobe; ++i) {
        const long long p = probes[i];
        std::uint64_t cnt = 1;
        const long long o[3] = { p % g.padded[0], (p / g.padded[0]) % g.padded[1],
                                 p / (static_cast<long long>(g.padded[0]) * g.padded[1]) };
        for (int a = 0; a != 3; ++a) {
            const long long sft = o[a] - (static_cast<long long>(s.kernel_shape[a]) - 1);
            const long long lo = std::max<long long>(0, -sft);
            const long long hi = std::min<long long>(s.kernel_shape[a], static_cast<long long>(s.environment_shape[a]) - sft);
            cnt *= std::uint64_t(std::max<long long>(0, hi - lo));
        }
        if (cnt > lim.max_direct_terms - terms) {
            st.direct_terms = lim.max_direct_terms;
            return SP_CORRELATION_DIRECT_TERM_LIMIT;
        }
        terms += cnt;
    }
    st.direct_terms = terms;
    SP_HIP_CHECK(hipMemcpyAsync(dprobe_idx, probes.data(), nprobe * sizeof(long long), hipMemcpyHostToDevice, st_));
    hipLaunchKernelGGL((k_probe<E>), dim3(unsigned(nprobe)), dim3(kThreads), 0, st_, denv, int(s.environment_shape[0]),
                       int(s.environment_shape[1]), int(s.environment_shape[2]), dker, int(s.kernel_shape[0]),
                       int(s.kernel_shape[1]), int(s.kernel_shape[2]), int(g.padded[0]), int(g.padded[1]), dprobe_idx,
                       dprobe);
    std::array<double, kMaxProbes> expected {}, got {};
    SP_HIP_CHECK(hipMemcpyAsync(expected.data(), dprobe, nprobe * sizeof(double), hipMemcpyDeviceToHost, st_));
    for (int i = 0; i != nprobe; ++i)
        SP_HIP_CHECK(hipMemcpyAsync(&got[i], dvals + probes[i], sizeof(double), hipMemcpyDeviceToHost, st_));
    SP_HIP_CHECK(hipStreamSynchronize(st_));
    double max_probe = 0;
    for (int i = 0; i != nprobe; ++i) {
        const double err = std::abs(got[i] - expected[i]);
        max_probe = std::max(max_probe, err);
        const double tol = binary ? 0.25 : 1e-10 * std::max(1.0, std::abs(expected[i]));
        if (!std::isfinite(err) || (binary ? err >= tol : err > tol)) return SP_CORRELATION_NUMERIC;
    }
    rep.max_probe_error = max_probe;
    SP_HIP_CHECK(hipMemcpyAsync(host_values, dvals, n * sizeof(double), hipMemcpyDeviceToHost, st_));
    SP_HIP_CHECK(hipStreamSynchronize(st_));
    return SP_OK;
}

// Per-device address of `scratch` (hipGetSymbolAddress at set-up) and a lock
// that keeps calls of one process on one GPU from overlapping in it.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_init_mu;
unsigned char* g_scratch_ptr[kMaxDevices] {};
std::mutex g_scratch_mu[kMaxDevices];

// Twiddles exp(-2 pi i k / n) written into the scratch's table at `dst`, by
// the same host formula as the cached tables, one axis at a time through a
// stack buffer.
template <typename T> void upload_twiddles(hipStream_t st, cx<T>* dst, int n)
{
    std::array<cx<T>, kMaxFftLength> host;
    fill_twiddles<T>(n, host.data());
    SP_HIP_CHECK(hipMemcpyAsync(dst, host.data(), sizeof(cx<T>) * n, hipMemcpyHostToDevice, st));
    SP_HIP_CHECK(hipStreamSynchronize(st));
}

template <typename E>
int correlate(sp_hip_ctx* ctx, const sp_correlation_spec* spec, const E* env, std::uint64_t env_n, const std::uint8_t* ker,
              std::uint64_t ker_n, const sp_correlation_limits* limits, int precision, double* out_values,
              std::uint64_t capacity, std::int64_t out_first[3], std::uint32_t out_shape[3], sp_numeric_report* report,
              sp_correlation_stats* stats, bool noalloc)
{
    constexpr bool binary = std::is_same_v<E, std::uint8_t>;
    if (!ctx || !spec || !out_first || !out_shape || !report || !stats) return SP_INVALID_ARGUMENT;
    sp_correlation_limits lim { 512ULL << 20, 0, 16777216, 200000000, 0 };
    if (limits) lim = *limits;
    *stats = sp_correlation_stats {};
    *report = sp_numeric_report {};
    auto& C = ctx->c;
    std::uint64_t meta;
    if (!checked_add(lim.reserved_bytes, lim.metadata_bytes, meta) || meta > lim.max_working_bytes)
        return SP_CORRELATION_MEMORY_LIMIT;
    Geometry g {};
    std::uint64_t env_count, ker_count;
    if (int rc = geometry(*spec, g, env_count, ker_count)) return rc;
    if (env_n != env_count || ker_n != ker_count || (env_n && !env) || (ker_n && !ker)) return SP_CORRELATION_INVALID_FIELD;
    std::unique_lock<std::mutex> hold;
    try {
        use_device(C);
        Layout L;
        if (!input_layout(env_count, ker_count, sizeof(E), L)) return SP_CORRELATION_SHAPE_OVERFLOW;
        unsigned char* in_base;
        if (noalloc) {
            if (C.device < 0 || C.device >= kMaxDevices || !g_scratch_ptr[C.device]) {
                C.last_error = "the allocation-free scratch of this device is not set up";
                return SP_DEVICE_ERROR;
            }
            if (L.inputs_end > SP_NOALLOC_SCRATCH_BYTES) {
                C.last_error = "inputs exceed the allocation-free scratch";
                return SP_SCRATCH_LIMIT;
            }
            hold = std::unique_lock<std::mutex>(g_scratch_mu[C.device]);
            in_base = g_scratch_ptr[C.device];
        }
        else {
            C.scratch[0].ensure(L.inputs_end);
            in_base = C.scratch[0].as<unsigned char>();
        }
        E* denv = reinterpret_cast<E*>(in_base + L.env);
        unsigned char* dker = in_base + L.ker;
        Stats* dst = reinterpret_cast<Stats*>(in_base + L.stats);
        SP_HIP_CHECK(hipMemcpyAsync(denv, env, env_count * sizeof(E), hipMemcpyHostToDevice, C.stream));
        SP_HIP_CHECK(hipMemcpyAsync(dker, ker, ker_count, hipMemcpyHostToDevice, C.stream));
        hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, C.stream, reinterpret_cast<unsigned int*>(dst),
                           int(sizeof(Stats) / sizeof(unsigned int)));
        hipLaunchKernelGGL((k_validate<E>), dim3(512), dim3(kThreads), 0, C.stream, denv, (long long)env_count, dker,
                           (long long)ker_count, dst);
        Stats h {};
        SP_HIP_CHECK(hipMemcpyAsync(&h, dst, sizeof h, hipMemcpyDeviceToHost, C.stream));
        SP_HIP_CHECK(hipStreamSynchronize(C.stream));
        if (h.bad_env) return SP_CORRELATION_INVALID_ENVIRONMENT;
        if (h.bad_kernel) return SP_CORRELATION_INVALID_KERNEL;
        double expected_mass, env_mass, max_out;
        std::uint64_t int_mass = 0;
        if constexpr (binary) {
            if (!checked_mul(h.ones_env, h.ones_kernel, int_mass) || int_mass > kExactInteger)
                return SP_CORRELATION_NUMERIC_RANGE;
            env_mass = double(h.ones_env);
            expected_mass = double(int_mass);
            max_out = double(std::min(h.ones_env, h.ones_kernel));
        }
        else {
            env_mass = h.env_sum;
            expected_mass = env_mass * double(h.ones_kernel);
            if (!std::isfinite(env_mass) || !std::isfinite(expected_mass)) return SP_CORRELATION_NUMERIC_RANGE;
            max_out = std::min(env_mass, double(h.ones_kernel));
        }
        stats->padded_cells = g.padded_count;
        if (g.padded_count > lim.max_padded_cells) return SP_CORRELATION_PADDED_CELL_LIMIT;
        // correlation.cpp:367-385: the caller's reserve plus the estimate's
        // working bytes must fit max_working_bytes
        std::uint64_t working = 0;
        if (!reference_working_bytes(g, 1, lim.metadata_bytes, working)) return SP_CORRELATION_SHAPE_OVERFLOW;
        if (!checked_add(lim.reserved_bytes, working, working) || working > lim.max_working_bytes)
            return SP_CORRELATION_MEMORY_LIMIT;
        stats->working_bytes_peak = working;
        if (!g.fft[0] || !g.fft[1] || !g.fft[2]) return SP_FFT_LENGTH_UNSUPPORTED;
        int prec = binary ? precision : 64;
        const double bound = fp32_error_bound(g.fft_cells, binary ? h.ones_env : 1, h.ones_kernel);
        if (prec == SP_PRECISION_AUTO) prec = bound > 0.4 ? 64 : 32;
        const std::size_t real = prec == 64 ? 8 : 4;
        compute_layout(g, real, L);
        if (noalloc) {
            if (L.total_with_twiddles > SP_NOALLOC_SCRATCH_BYTES) {
                C.last_error = "the correlation needs " + std::to_string(L.total_with_twiddles) +
                               " bytes, more than the allocation-free scratch holds";
                return SP_SCRATCH_LIMIT;
            }
        }
        else if (C.bytes_cap && L.total > C.bytes_cap)
            return SP_CORRELATION_MEMORY_LIMIT;
        for (int a = 0; a != 3; ++a) {
            out_first[a] = g.first[a];
            out_shape[a] = g.padded[a];
            report->fft_shape[a] = std::uint32_t(g.fft[a]);
        }
        report->precision = prec;
        report->error_bound = prec == 32 ? bound : 0.0;
        if (capacity < g.padded_count || !out_values) return SP_CAPACITY;
        unsigned char* base;
        std::uint64_t origin;
        if (noalloc) {
            base = in_base;
            origin = 0;
        }
        else {
            C.scratch[1].ensure(L.total - L.inputs_end);
            base = C.scratch[1].as<unsigned char>();
            origin = L.inputs_end;
        }
        auto go = [&](auto tag) {
            using T = decltype(tag);
            Buffers<T> B = buffers_at<T>(base, L, origin);
            for (int a = 0; a != 3; ++a) {
                if (noalloc) {
                    cx<T>* t = reinterpret_cast<cx<T>*>(base + L.tw[a]);
                    upload_twiddles<T>(C.stream, t, g.fft[a]);
                    B.tw[a] = t;
                }
                else
                    B.tw[a] = twiddles<T>(g.fft[a]);
            }
            return run<T, E>(C, *spec, g, denv, dker, B, binary, max_out, expected_mass, int_mass, lim, out_values,
                             *report, *stats);
        };
        return prec == 64 ? go(double {}) : go(float {});
    }
    catch (const DeviceError& e) {
        C.last_error = e.what;
        return SP_DEVICE_ERROR;
    }
    catch (const std::bad_alloc&) {
        return SP_CORRELATION_ALLOCATION;
    }
}

}  // namespace

// Set-up of the allocation-free scratch for one device (the library's
// context creation calls it with that device current): its address on that
// device, once.
int init_noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return SP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(g_scratch_init_mu);
    if (g_scratch_ptr[device]) return SP_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SP_DEVICE_ERROR;
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    return SP_OK;
}

}  // namespace sphip

extern "C" {

int sp_hip_estimate(const sp_correlation_spec* spec, uint32_t fft_threads, uint64_t metadata_bytes,
                    sp_correlation_estimate* out)
{
    if (!spec || !out) return SP_INVALID_ARGUMENT;
    *out = sp_correlation_estimate {};
    if (fft_threads < 1 || fft_threads > 8) return SP_CPU_THREAD_COUNT_UNSUPPORTED;
    sphip::Geometry g {};
    std::uint64_t ec, kc;
    if (int rc = sphip::geometry(*spec, g, ec, kc)) return rc;
    for (int a = 0; a != 3; ++a) out->padded_shape[a] = g.padded[a];
    out->padded_cells = g.padded_count;
    if (!sphip::reference_working_bytes(g, fft_threads, metadata_bytes, out->working_bytes)) {
        out->working_bytes = 0;
        return SP_CORRELATION_SHAPE_OVERFLOW;
    }
    if (g.fft[0] && g.fft[1] && g.fft[2]) out->device_bytes = sphip::device_bytes(g, ec, kc, 8, 1);
    return SP_OK;
}

int sp_hip_correlate_binary(sp_hip_ctx* ctx, const sp_correlation_spec* spec, const uint8_t* environment,
                            uint64_t environment_count, const uint8_t* kernel, uint64_t kernel_count,
                            const sp_correlation_limits* limits, int precision, double* out_values, uint64_t out_capacity,
                            int64_t out_first[3], uint32_t out_shape[3], sp_numeric_report* report,
                            sp_correlation_stats* stats)
{
    if (precision != SP_PRECISION_AUTO && precision != SP_PRECISION_FP32 && precision != SP_PRECISION_FP64)
        return SP_INVALID_ARGUMENT;
    return sphip::correlate<std::uint8_t>(ctx, spec, environment, environment_count, kernel, kernel_count, limits,
                                          precision, out_values, out_capacity, out_first, out_shape, report, stats,
                                          false);
}

int sp_hip_correlate_binary_noalloc(sp_hip_ctx* ctx, const sp_correlation_spec* spec, const uint8_t* environment,
                                    uint64_t environment_count, const uint8_t* kernel, uint64_t kernel_count,
                                    const sp_correlation_limits* limits, int precision, double* out_values,
                                    uint64_t out_capacity, int64_t out_first[3], uint32_t out_shape[3],
                                    sp_numeric_report* report, sp_correlation_stats* stats)
{
    if (precision != SP_PRECISION_AUTO && precision != SP_PRECISION_FP32 && precision != SP_PRECISION_FP64)
        return SP_INVALID_ARGUMENT;
    return sphip::correlate<std::uint8_t>(ctx, spec, environment, environment_count, kernel, kernel_count, limits,
                                          precision, out_values, out_capacity, out_first, out_shape, report, stats,
                                          true);
}

int sp_hip_correlate_proximity(sp_hip_ctx* ctx, const sp_correlation_spec* spec, const double* environment,
                               uint64_t environment_count, const uint8_t* kernel, uint64_t kernel_count,
                               const sp_correlation_limits* limits, double* out_values, uint64_t out_capacity,
                               int64_t out_first[3], uint32_t out_shape[3], sp_numeric_report* report,
                               sp_correlation_stats* stats)
{
    return sphip::correlate<double>(ctx, spec, environment, environment_count, kernel, kernel_count, limits,
                                    SP_PRECISION_FP64, out_values, out_capacity, out_first, out_shape, report, stats,
                                    false);
}

}  // extern "C"
