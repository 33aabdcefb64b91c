// Oracle (test infrastructure): the checked full-linear FFT correlation of
// engine/pack_compute/src/correlation.cpp:238-655 and the independent direct
// integer oracle of tests/support/src/integer_correlation.cpp:76-130.
#include <algorithm>
#include <cmath>
#include <limits>
#include <type_traits>

#include "sp_oracle.hpp"

namespace sporacle {
namespace {

using C = std::complex<double>;
constexpr std::uint64_t kExactInteger = 1ULL << 53;

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

// correlation.cpp:117-132 (Kahan-style compensated sum)
struct Compensated {
    double value {}, correction {};
    void add(double v)
    {
        const double adjusted = v - correction;
        const double next = value + adjusted;
        correction = (next - value) - adjusted;
        value = next;
    }
};

std::size_t lin(const std::array<std::uint32_t, 3>& s, std::uint32_t x, std::uint32_t y, std::uint32_t z)
{
    return x + std::size_t(s[0]) * (y + std::size_t(s[1]) * z);
}

// The metadata term of correlation.cpp:67-72 depends on the reference's
// object sizes; this restatement uses a fixed 256-byte stand-in.
constexpr std::uint64_t kMetadataBytes = 256;
constexpr std::uint64_t kPocketWorkspaceFixed = 64ULL << 10;

// correlation.cpp:153-218
template <typename E>
bool direct_probe(const CorrelationSpec& s, const E* env, const std::uint8_t* ker,
                  const std::array<std::uint32_t, 3>& out_shape, std::size_t index, const CorrelationLimits& l,
                  std::uint64_t& terms, double& expected)
{
    const std::uint32_t x = index % out_shape[0];
    const std::size_t yz = index / out_shape[0];
    const std::uint32_t y = yz % out_shape[1], z = yz / out_shape[1];
    const std::int64_t shift[3] = { std::int64_t(x) - (std::int64_t(s.kernel_shape[0]) - 1),
                                    std::int64_t(y) - (std::int64_t(s.kernel_shape[1]) - 1),
                                    std::int64_t(z) - (std::int64_t(s.kernel_shape[2]) - 1) };
    std::uint64_t isum = 0;
    Compensated rsum;
    for (std::uint32_t kz = 0; kz != s.kernel_shape[2]; ++kz)
        for (std::uint32_t ky = 0; ky != s.kernel_shape[1]; ++ky)
            for (std::uint32_t kx = 0; kx != s.kernel_shape[0]; ++kx) {
                const std::int64_t e[3] = { kx + shift[0], ky + shift[1], kz + shift[2] };
                if (e[0] < 0 || e[1] < 0 || e[2] < 0 || e[0] >= s.environment_shape[0] ||
                    e[1] >= s.environment_shape[1] || e[2] >= s.environment_shape[2])
                    continue;
                if (terms == l.max_direct_terms) return false;
                ++terms;
                const auto eo = lin(s.environment_shape, e[0], e[1], e[2]);
                const auto ko = lin(s.kernel_shape, kx, ky, kz);
                if constexpr (std::is_same_v<E, std::uint8_t>) isum += std::uint64_t(env[eo]) * ker[ko];
                else rsum.add(env[eo] * ker[ko]);
            }
    if constexpr (std::is_same_v<E, std::uint8_t>) {
        if (isum > kExactInteger) return false;
        expected = double(isum);
    }
    else
        expected = rsum.value;
    return true;
}

void add_probe(std::size_t* probes, std::size_t& count, std::size_t c)
{
    for (std::size_t i = 0; i != count; ++i)
        if (probes[i] == c) return;
    if (count == 8) return;
    probes[count++] = c;
}

CorrelationOut fail(const char* code, std::uint64_t padded = 0, std::uint64_t peak = 0, std::uint64_t terms = 0)
{
    CorrelationOut o;
    o.code = code;
    o.padded_cells = padded;
    o.working_bytes_peak = peak;
    o.direct_terms = terms;
    return o;
}

// correlation.cpp:238-592
template <typename E>
CorrelationOut correlate(const CorrelationSpec& s, const E* env, std::uint64_t env_n, const std::uint8_t* ker,
                         std::uint64_t ker_n, const CorrelationLimits& l, NumericFault fault)
{
    constexpr bool binary = std::is_same_v<E, std::uint8_t>;
    std::uint64_t meta;
    if (!checked_add(l.reserved_bytes, kMetadataBytes, meta) || meta > l.max_working_bytes)
        return fail("CORRELATION_MEMORY_LIMIT");
    std::uint64_t env_count = 1, ker_count = 1, padded_count = 1;
    std::array<std::uint32_t, 3> padded {};
    std::array<std::int64_t, 3> tfirst {};
    for (int a = 0; a != 3; ++a) {
        if (s.environment_shape[a] == 0 || s.kernel_shape[a] == 0 || !checked_mul(env_count, s.environment_shape[a], env_count) ||
            !checked_mul(ker_count, s.kernel_shape[a], ker_count))
            return fail("CORRELATION_SHAPE_OVERFLOW");
        const std::uint64_t len = std::uint64_t(s.environment_shape[a]) + s.kernel_shape[a] - 1;
        if (len > 0xffffffffULL || !checked_mul(padded_count, len, padded_count)) return fail("CORRELATION_SHAPE_OVERFLOW");
        padded[a] = std::uint32_t(len);
        std::int64_t diff;
        if (!checked_sub(s.environment_first[a], s.kernel_first[a], diff) ||
            !checked_sub(diff, std::int64_t(s.kernel_shape[a]) - 1, tfirst[a]))
            return fail("CORRELATION_INDEX_OVERFLOW");
        if (tfirst[a] > std::numeric_limits<std::int64_t>::max() - std::int64_t(len - 1))
            return fail("CORRELATION_INDEX_OVERFLOW");
    }
    if (env_n != env_count || ker_n != ker_count) return fail("CORRELATION_INVALID_FIELD");
    std::uint64_t env_ones = 0, ker_ones = 0;
    Compensated env_mass_sum;
    for (std::uint64_t i = 0; i != env_n; ++i) {
        const double v = double(env[i]);
        if (!std::isfinite(v) || v < 0.0 || v > 1.0) return fail("CORRELATION_INVALID_ENVIRONMENT");
        if constexpr (binary) env_ones += env[i];
        else env_mass_sum.add(v);
    }
    for (std::uint64_t i = 0; i != ker_n; ++i) {
        if (ker[i] > 1) return fail("CORRELATION_INVALID_KERNEL");
        ker_ones += ker[i];
    }
    double expected_mass, env_mass;
    if constexpr (binary) {
        std::uint64_t prod;
        if (!checked_mul(env_ones, ker_ones, prod) || prod > kExactInteger) return fail("CORRELATION_NUMERIC_RANGE");
        env_mass = double(env_ones);
        expected_mass = double(prod);
    }
    else {
        env_mass = env_mass_sum.value;
        expected_mass = env_mass * double(ker_ones);
        if (!std::isfinite(env_mass) || !std::isfinite(expected_mass)) return fail("CORRELATION_NUMERIC_RANGE");
    }
    if (padded_count > l.max_padded_cells) return fail("CORRELATION_PADDED_CELL_LIMIT", padded_count);
    const Estimate est = estimate_correlation(s, 1);
    if (!est.code.empty()) return fail(est.code.c_str(), padded_count);
    std::uint64_t working = l.reserved_bytes;
    if (!checked_add(working, est.working_bytes, working) || working > l.max_working_bytes)
        return fail("CORRELATION_MEMORY_LIMIT", padded_count);

    std::vector<C> blocked(padded_count), occupied(padded_count);
    for (std::uint32_t z = 0; z != s.environment_shape[2]; ++z)
        for (std::uint32_t y = 0; y != s.environment_shape[1]; ++y)
            for (std::uint32_t x = 0; x != s.environment_shape[0]; ++x)
                blocked[lin(padded, x, y, z)] = double(env[lin(s.environment_shape, x, y, z)]);
    for (std::uint32_t z = 0; z != s.kernel_shape[2]; ++z)
        for (std::uint32_t y = 0; y != s.kernel_shape[1]; ++y)
            for (std::uint32_t x = 0; x != s.kernel_shape[0]; ++x)
                occupied[lin(padded, x, y, z)] = double(ker[lin(s.kernel_shape, x, y, z)]);
    for (int a = 0; a != 3; ++a) {
        fft_axis(blocked.data(), padded, a, true, 1.0);
        fft_axis(occupied.data(), padded, a, true, 1.0);
    }
    for (std::size_t i = 0; i != blocked.size(); ++i) blocked[i] *= std::conj(occupied[i]);
    const double inv = (fault == NumericFault::bad_normalization ? 2.0 : 1.0) / double(padded_count);
    for (int a = 0; a != 3; ++a) fft_axis(blocked.data(), padded, a, false, a == 0 ? inv : 1.0);
    if (fault == NumericFault::nan) blocked[0] = C(std::numeric_limits<double>::quiet_NaN(), 0.0);
    if (fault == NumericFault::integral_corruption) blocked.back() += 1.0;

    CorrelationOut out;
    out.padded_cells = padded_count;
    out.working_bytes_peak = working;
    std::vector<double> values(padded_count);
    Compensated out_mass;
    double max_res = 0, max_mag = -1.0;
    std::size_t max_idx = 0;
    const double max_out = binary ? double(std::min(env_ones, ker_ones)) : std::min(env_mass, double(ker_ones));
    const double prox_tol = 1e-10 * std::max(1.0, max_out);
    for (std::uint32_t z = 0; z != padded[2]; ++z)
        for (std::uint32_t y = 0; y != padded[1]; ++y)
            for (std::uint32_t x = 0; x != padded[0]; ++x) {
                const std::uint32_t tx = (std::uint64_t(x) + padded[0] - (s.kernel_shape[0] - 1)) % padded[0];
                const std::uint32_t ty = (std::uint64_t(y) + padded[1] - (s.kernel_shape[1] - 1)) % padded[1];
                const std::uint32_t tz = (std::uint64_t(z) + padded[2] - (s.kernel_shape[2] - 1)) % padded[2];
                const C t = blocked[lin(padded, tx, ty, tz)];
                const double v = t.real();
                const double imag_tol = binary ? 0.25 : prox_tol;
                if (!std::isfinite(v) || !std::isfinite(t.imag()) || std::abs(t.imag()) >= imag_tol)
                    return fail("CORRELATION_NUMERIC", padded_count, working);
                const auto o = lin(padded, x, y, z);
                out_mass.add(v);
                if (std::abs(v) > max_mag) {
                    max_mag = std::abs(v);
                    max_idx = o;
                }
                if constexpr (binary) {
                    const double r = std::round(v);
                    const double res = std::abs(v - r);
                    max_res = std::max(max_res, res);
                    if (res >= 0.25 || r < 0.0 || r > max_out) return fail("CORRELATION_NUMERIC", padded_count, working);
                    values[o] = r;
                }
                else {
                    if (v < -prox_tol || v > max_out + prox_tol) return fail("CORRELATION_NUMERIC", padded_count, working);
                    values[o] = v;
                }
            }
    const double mass_res = std::abs(out_mass.value - expected_mass);
    const double mass_tol = binary ? 0.25 : 1e-10 * std::max(1.0, std::abs(expected_mass));
    if (!std::isfinite(mass_res) || (binary ? mass_res >= mass_tol : mass_res > mass_tol))
        return fail("CORRELATION_NUMERIC", padded_count, working);
    std::size_t probes[8];
    std::size_t pc = 0;
    const std::size_t n = padded_count;
    add_probe(probes, pc, 0);
    add_probe(probes, pc, n - 1);
    add_probe(probes, pc, n / 2);
    add_probe(probes, pc, max_idx);
    add_probe(probes, pc, n / 4);
    add_probe(probes, pc, n / 3);
    add_probe(probes, pc, (n / 3) * 2);
    add_probe(probes, pc, (n / 4) * 3);
    double max_probe = 0;
    std::uint64_t terms = 0;
    for (std::size_t i = 0; i != pc; ++i) {
        double expected;
        if (!direct_probe(s, env, ker, padded, probes[i], l, terms, expected))
            return fail("CORRELATION_DIRECT_TERM_LIMIT", padded_count, working, terms);
        const double err = std::abs(values[probes[i]] - expected);
        max_probe = std::max(max_probe, err);
        const double tol = binary ? 0.25 : 1e-10 * std::max(1.0, std::abs(expected));
        if (!std::isfinite(err) || (binary ? err >= tol : err > tol))
            return fail("CORRELATION_NUMERIC", padded_count, working, terms);
    }
    out.translation_first = tfirst;
    out.shape = padded;
    out.values = std::move(values);
    out.max_integer_residual = max_res;
    out.mass_residual = mass_res;
    out.max_probe_error = max_probe;
    out.direct_terms = terms;
    return out;
}

}  // namespace

// correlation.cpp:596-641 with the pocketfft workspace model of :220-236.
Estimate estimate_correlation(const CorrelationSpec& s, std::uint32_t fft_threads)
{
    Estimate e;
    if (fft_threads < 1 || fft_threads > 8) {
        e.code = "CPU_THREAD_COUNT_UNSUPPORTED";
        return e;
    }
    e.padded_cells = 1;
    for (int a = 0; a != 3; ++a) {
        const std::uint64_t len = std::uint64_t(s.environment_shape[a]) + s.kernel_shape[a] - 1;
        if (s.environment_shape[a] == 0 || s.kernel_shape[a] == 0 || len > 0xffffffffULL ||
            !checked_mul(e.padded_cells, len, e.padded_cells)) {
            e.code = "CORRELATION_SHAPE_OVERFLOW";
            return e;
        }
        e.padded_shape[a] = std::uint32_t(len);
        std::int64_t first;
        if (!checked_sub(s.environment_first[a], s.kernel_first[a], first) ||
            !checked_sub(first, std::int64_t(s.kernel_shape[a]) - 1, first) ||
            first > std::numeric_limits<std::int64_t>::max() - std::int64_t(len - 1)) {
            e.code = "CORRELATION_INDEX_OVERFLOW";
            return e;
        }
    }
    const std::uint64_t longest = *std::max_element(e.padded_shape.begin(), e.padded_shape.end());
    std::uint64_t ws;
    if (!checked_mul(e.padded_cells, 2 * sizeof(C) + sizeof(double), e.working_bytes) ||
        !checked_mul(longest, 64 * sizeof(C), ws) || !checked_add(ws, kPocketWorkspaceFixed, ws) ||
        !checked_mul(ws, fft_threads, ws) || !checked_add(e.working_bytes, ws, e.working_bytes) ||
        !checked_add(e.working_bytes, kMetadataBytes + 8 * sizeof(std::size_t), e.working_bytes)) {
        e.code = "CORRELATION_SHAPE_OVERFLOW";
        return e;
    }
    return e;
}

CorrelationOut correlate_binary(const CorrelationSpec& s, const std::uint8_t* env, std::uint64_t env_n,
                                const std::uint8_t* ker, std::uint64_t ker_n, const CorrelationLimits& l,
                                NumericFault fault)
{
    return correlate(s, env, env_n, ker, ker_n, l, fault);
}

CorrelationOut correlate_proximity(const CorrelationSpec& s, const double* env, std::uint64_t env_n,
                                   const std::uint8_t* ker, std::uint64_t ker_n, const CorrelationLimits& l,
                                   NumericFault fault)
{
    return correlate(s, env, env_n, ker, ker_n, l, fault);
}

// tests/support/src/integer_correlation.cpp:76-130: sum B(x) * O(x - t) over
// every translation with at least one overlapping cell, as exact integers.
DirectOut direct_linear_cross_correlation(const std::vector<int>& b, std::array<int, 3> be, const std::vector<int>& k,
                                          std::array<int, 3> ke, std::array<int, 3> origin)
{
    DirectOut o;
    for (int a = 0; a != 3; ++a)
        if (be[a] <= 0 || ke[a] <= 0) {
            o.code = "invalid_argument";
            return o;
        }
    if (b.size() != std::size_t(be[0]) * be[1] * be[2] || k.size() != std::size_t(ke[0]) * ke[1] * ke[2]) {
        o.code = "invalid_argument";
        return o;
    }
    auto checked_int = [&](long long v, int& out) {
        if (v < std::numeric_limits<int>::min() || v > std::numeric_limits<int>::max()) return false;
        out = int(v);
        return true;
    };
    for (int a = 0; a != 3; ++a) {
        if (!checked_int(-((long long)(ke[a]) - 1 + origin[a]), o.range[2 * a]) ||
            !checked_int((long long)(be[a]) - 1 - origin[a], o.range[2 * a + 1]) ||
            !checked_int((long long)(be[a]) + ke[a] - 1, o.extent[a])) {
            o.code = "invalid_argument";
            return o;
        }
    }
    o.values.assign(std::size_t(o.extent[0]) * o.extent[1] * o.extent[2], 0);
    for (long long tz = o.range[4]; tz <= o.range[5]; ++tz)
        for (long long ty = o.range[2]; ty <= o.range[3]; ++ty)
            for (long long tx = o.range[0]; tx <= o.range[1]; ++tx) {
                long long sum = 0;
                for (int kz = 0; kz < ke[2]; ++kz)
                    for (int ky = 0; ky < ke[1]; ++ky)
                        for (int kx = 0; kx < ke[0]; ++kx) {
                            const long long bx = tx + origin[0] + kx, by = ty + origin[1] + ky, bz = tz + origin[2] + kz;
                            if (bx < 0 || bx >= be[0] || by < 0 || by >= be[1] || bz < 0 || bz >= be[2]) continue;
                            const long long term =
                                (long long)b[bx + std::size_t(be[0]) * (by + std::size_t(be[1]) * bz)] *
                                k[kx + std::size_t(ke[0]) * (ky + std::size_t(ke[1]) * kz)];
                            if ((term > 0 && sum > std::numeric_limits<long long>::max() - term) ||
                                (term < 0 && sum < std::numeric_limits<long long>::min() - term)) {
                                o.code = "overflow_error";
                                return o;
                            }
                            sum += term;
                        }
                if (sum < std::numeric_limits<int>::min() || sum > std::numeric_limits<int>::max()) {
                    o.code = "overflow_error";
                    return o;
                }
                o.values[std::size_t(tx - o.range[0]) +
                         std::size_t(o.extent[0]) * (std::size_t(ty - o.range[2]) + std::size_t(o.extent[1]) * std::size_t(tz - o.range[4]))] =
                    int(sum);
            }
    return o;
}

}  // namespace sporacle
