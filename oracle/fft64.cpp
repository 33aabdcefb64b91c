// Oracle (test infrastructure): FP64 complex FFT along one axis of an
// X-fastest 3D array.  It plays the role pocketfft::c2c has in
// engine/pack_compute/src/fft_execution.cpp:57-121 (one axis per call, forward
// = exp(-2*pi*i*jk/n), the factor applied to the output).  pocketfft itself
// (vcpkg pocketfft 2023-09-25, mreineck/pocketfft@9efd4da5) is not vendored in
// the reference; this is an independent mixed-radix (2,3,4,5 and generic odd
// radices) Cooley-Tukey with Bluestein's algorithm for lengths with a prime
// factor above 13, the same family of methods pocketfft publishes.
#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "sp_oracle.hpp"

namespace sporacle {
namespace {

using C = std::complex<double>;

C twiddle(std::uint64_t k, std::uint64_t n, bool forward)
{
    // Reduce k/n to the first octant in long double for accurate twiddles.
    const long double pi = 3.141592653589793238462643383279502884L;
    const long double angle = 2.0L * pi * static_cast<long double>(k % n) / static_cast<long double>(n);
    const double c = static_cast<double>(std::cos(angle));
    const double s = static_cast<double>(std::sin(angle));
    return forward ? C(c, -s) : C(c, s);
}

struct Plan {
    std::size_t n {};
    bool bluestein {};
    std::vector<std::size_t> factors;       // mixed-radix factors, product n
    std::vector<C> tw_fwd, tw_inv;         // per-stage twiddles
    std::vector<std::vector<C>> roots_fwd, roots_inv;  // per-stage radix roots
    // Bluestein
    std::size_t m {};
    std::vector<C> chirp_fwd, chirp_inv;   // w_k = exp(-+ i pi k^2 / n)
    std::vector<C> filt_fwd, filt_inv;     // FFT_m of conj chirp sequence
    std::shared_ptr<Plan> sub;             // power-of-two plan of length m
};

std::vector<std::size_t> factorize(std::size_t n, std::size_t& largest)
{
    std::vector<std::size_t> f;
    largest = 1;
    while (n % 4 == 0) {
        f.push_back(4);
        n /= 4;
    }
    while (n % 2 == 0) {
        f.push_back(2);
        n /= 2;
    }
    for (std::size_t p = 3; p * p <= n; p += 2)
        while (n % p == 0) {
            f.push_back(p);
            n /= p;
            largest = std::max(largest, p);
        }
    if (n > 1) {
        f.push_back(n);
        largest = std::max(largest, n);
    }
    return f;
}

std::shared_ptr<Plan> make_plan(std::size_t n);

std::shared_ptr<Plan> get_plan(std::size_t n)
{
    static std::mutex mu;
    static std::map<std::size_t, std::shared_ptr<Plan>> cache;
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find(n);
        if (it != cache.end()) return it->second;
    }
    auto p = make_plan(n);
    std::lock_guard<std::mutex> lock(mu);
    cache[n] = p;
    return p;
}

void run(const Plan& p, C* data, bool forward, std::vector<C>& scratch);

std::shared_ptr<Plan> make_plan(std::size_t n)
{
    auto p = std::make_shared<Plan>();
    p->n = n;
    std::size_t largest;
    p->factors = factorize(n, largest);
    // Bluestein only when it is cheaper than the generic-radix passes, with
    // pocketfft's published cost model (n * sum of factors, odd factors above
    // 5 weighted 1.1; Bluestein as 2 transforms of the next 2^a length times a
    // 1.5 fudge factor).
    auto cost = [](std::size_t len) {
        std::size_t lg;
        const auto fs = factorize(len, lg);
        double c = 0;
        for (std::size_t f : fs) c += f <= 5 ? double(f) : 1.1 * double(f);
        return c * double(len);
    };
    std::size_t m2 = 1;
    while (m2 < 2 * n - 1) m2 <<= 1;
    if (largest > 5 && 1.5 * 2 * cost(m2) < cost(n)) {
        p->bluestein = true;
        std::size_t m = 1;
        while (m < 2 * n - 1) m <<= 1;
        p->m = m;
        p->sub = get_plan(m);
        p->chirp_fwd.resize(n);
        p->chirp_inv.resize(n);
        const long double pi = 3.141592653589793238462643383279502884L;
        for (std::size_t k = 0; k != n; ++k) {
            const std::uint64_t k2 = (std::uint64_t(k) * k) % (2 * n);
            const long double a = pi * static_cast<long double>(k2) / static_cast<long double>(n);
            const C w(static_cast<double>(std::cos(a)), static_cast<double>(std::sin(a)));
            p->chirp_fwd[k] = std::conj(w);
            p->chirp_inv[k] = w;
        }
        std::vector<C> scratch;
        for (int dir = 0; dir != 2; ++dir) {
            const auto& ch = dir == 0 ? p->chirp_fwd : p->chirp_inv;
            std::vector<C> b(m, C(0, 0));
            b[0] = std::conj(ch[0]);
            for (std::size_t k = 1; k != n; ++k) b[k] = b[m - k] = std::conj(ch[k]);
            run(*p->sub, b.data(), true, scratch);
            (dir == 0 ? p->filt_fwd : p->filt_inv) = std::move(b);
        }
        return p;
    }
    // Stage twiddles for a decimation-in-time Stockham formulation.
    std::size_t l = 1;
    for (std::size_t r : p->factors) {
        std::vector<C> rf(r), ri(r);
        for (std::size_t q = 0; q != r; ++q) {
            rf[q] = twiddle(q, r, true);
            ri[q] = twiddle(q, r, false);
        }
        p->roots_fwd.push_back(std::move(rf));
        p->roots_inv.push_back(std::move(ri));
        const std::size_t span = l * r;
        for (std::size_t j = 0; j != l; ++j)
            for (std::size_t q = 1; q != r; ++q) {
                p->tw_fwd.push_back(twiddle(j * q, span, true));
                p->tw_inv.push_back(twiddle(j * q, span, false));
            }
        l = span;
    }
    return p;
}

// Stockham autosort: stage with radix r, l = product of earlier radices.
void run_mixed(const Plan& p, C* data, bool forward, std::vector<C>& scratch)
{
    const std::size_t n = p.n;
    scratch.resize(n);
    C* src = data;
    C* dst = scratch.data();
    std::size_t l = 1;
    std::size_t tw_off = 0;
    const auto& tw = forward ? p.tw_fwd : p.tw_inv;
    std::vector<C> tmp;
    for (std::size_t stage = 0; stage != p.factors.size(); ++stage) {
        const std::size_t r = p.factors[stage];
        const std::size_t m = n / (l * r);  // remaining
        const auto& roots = forward ? p.roots_fwd[stage] : p.roots_inv[stage];
        tmp.resize(r);
        for (std::size_t j = 0; j != l; ++j) {
            for (std::size_t k = 0; k != m; ++k) {
                // gather the r inputs of this butterfly: src[(q*m + k)*l + j]
                for (std::size_t q = 0; q != r; ++q) {
                    C v = src[(q * m + k) * l + j];
                    if (q != 0) v *= tw[tw_off + j * (r - 1) + (q - 1)];
                    tmp[q] = v;
                }
                // radix-r DFT
                if (r == 2) {
                    dst[(k * r) * l + j] = tmp[0] + tmp[1];
                    dst[(k * r + 1) * l + j] = tmp[0] - tmp[1];
                }
                else if (r == 4) {
                    const C a = tmp[0] + tmp[2], b = tmp[0] - tmp[2];
                    const C c = tmp[1] + tmp[3];
                    C d = tmp[1] - tmp[3];
                    d = forward ? C(d.imag(), -d.real()) : C(-d.imag(), d.real());
                    dst[(k * r) * l + j] = a + c;
                    dst[(k * r + 1) * l + j] = b + d;
                    dst[(k * r + 2) * l + j] = a - c;
                    dst[(k * r + 3) * l + j] = b - d;
                }
                else {
                    for (std::size_t s = 0; s != r; ++s) {
                        C acc = tmp[0];
                        std::size_t idx = s;
                        for (std::size_t q = 1; q != r; ++q) {
                            acc += tmp[q] * roots[idx];
                            idx += s;
                            if (idx >= r) idx -= r;
                        }
                        dst[(k * r + s) * l + j] = acc;
                    }
                }
            }
        }
        tw_off += l * (r - 1);
        l *= r;
        std::swap(src, dst);
    }
    if (src != data)
        for (std::size_t i = 0; i != n; ++i) data[i] = src[i];
}

void run(const Plan& p, C* data, bool forward, std::vector<C>& scratch)
{
    if (!p.bluestein) {
        run_mixed(p, data, forward, scratch);
        return;
    }
    const std::size_t n = p.n, m = p.m;
    const auto& ch = forward ? p.chirp_fwd : p.chirp_inv;
    const auto& filt = forward ? p.filt_fwd : p.filt_inv;
    std::vector<C> a(m, C(0, 0)), sub_scratch;
    for (std::size_t k = 0; k != n; ++k) a[k] = data[k] * ch[k];
    run(*p.sub, a.data(), true, sub_scratch);
    for (std::size_t k = 0; k != m; ++k) a[k] *= filt[k];
    run(*p.sub, a.data(), false, sub_scratch);
    const double inv = 1.0 / static_cast<double>(m);
    for (std::size_t k = 0; k != n; ++k) data[k] = a[k] * inv * ch[k];
}

}  // namespace

void fft_axis(std::complex<double>* data, const std::array<std::uint32_t, 3>& shape, int axis, bool forward,
              double factor)
{
    const std::size_t n = shape[axis];
    auto plan = get_plan(n);
    const std::size_t s0 = 1, s1 = shape[0], s2 = std::size_t(shape[0]) * shape[1];
    const std::size_t stride = axis == 0 ? s0 : axis == 1 ? s1 : s2;
    std::size_t outer_a, outer_b, sa, sb;
    if (axis == 0) { outer_a = shape[1]; sa = s1; outer_b = shape[2]; sb = s2; }
    else if (axis == 1) { outer_a = shape[0]; sa = s0; outer_b = shape[2]; sb = s2; }
    else { outer_a = shape[0]; sa = s0; outer_b = shape[1]; sb = s1; }
    std::vector<C> scratch;
    if (axis == 0) {
        std::vector<C> line(n);
        for (std::size_t b = 0; b != outer_b; ++b)
            for (std::size_t a = 0; a != outer_a; ++a) {
                C* base = data + a * sa + b * sb;
                for (std::size_t i = 0; i != n; ++i) line[i] = base[i];
                run(*plan, line.data(), forward, scratch);
                for (std::size_t i = 0; i != n; ++i) base[i] = factor == 1.0 ? line[i] : line[i] * factor;
            }
        return;
    }
    // strided axes: gather a block of adjacent lines (contiguous along x) at once
    constexpr std::size_t kBlock = 16;
    std::vector<C> lines(kBlock * n);
    for (std::size_t b = 0; b != outer_b; ++b)
        for (std::size_t a0 = 0; a0 < outer_a; a0 += kBlock) {
            const std::size_t cnt = std::min(kBlock, outer_a - a0);
            C* base = data + a0 * sa + b * sb;
            for (std::size_t i = 0; i != n; ++i)
                for (std::size_t l = 0; l != cnt; ++l) lines[l * n + i] = base[l * sa + i * stride];
            for (std::size_t l = 0; l != cnt; ++l) run(*plan, lines.data() + l * n, forward, scratch);
            for (std::size_t i = 0; i != n; ++i)
                for (std::size_t l = 0; l != cnt; ++l)
                    base[l * sa + i * stride] = factor == 1.0 ? lines[l * n + i] : lines[l * n + i] * factor;
        }
}

}  // namespace sporacle
