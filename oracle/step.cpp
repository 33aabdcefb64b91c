// Oracle (test infrastructure): orientation catalogs, the plain placement step
// (free masks + lowest (z, y, x) choice + stamping) and the proximity field.
#include <algorithm>
#include <cmath>
#include <limits>

#include "sp_oracle.hpp"

namespace sporacle {

namespace {
using Matrix = std::array<int, 9>;

// engine/pack_solver/src/orientation_cube.hpp:38-78 and gap_fill.cpp:137-169:
// canonical quaternion of a signed-permutation matrix.
Quat matrix_quaternion(const Matrix& m)
{
    const int trace = m[0] + m[4] + m[8];
    std::array<int, 4> n {};
    if (trace > 0) n = { m[7] - m[5], m[2] - m[6], m[3] - m[1], trace + 1 };
    else if (m[0] > m[4] && m[0] > m[8]) n = { 1 + m[0] - m[4] - m[8], m[1] + m[3], m[2] + m[6], m[7] - m[5] };
    else if (m[4] > m[8]) n = { m[1] + m[3], 1 + m[4] - m[0] - m[8], m[5] + m[7], m[2] - m[6] };
    else n = { m[2] + m[6], m[5] + m[7], 1 + m[8] - m[0] - m[4], m[3] - m[1] };
    const auto nonzero = std::count_if(n.begin(), n.end(), [](int v) { return v != 0; });
    const double mag = nonzero == 1 ? 1.0 : nonzero == 2 ? 0.7071067811865475244 : 0.5;
    Quat q {};
    for (int i = 0; i != 4; ++i) q[i] = n[i] < 0 ? -mag : n[i] ? mag : 0.0;
    bool negate = q[3] < 0;
    if (q[3] == 0) {
        for (int i = 0; i != 3; ++i)
            if (q[i] != 0) {
                negate = q[i] < 0;
                break;
            }
    }
    for (auto& c : q) {
        if (negate) c = -c;
        if (c == 0) c = 0;
    }
    return q;
}
}  // namespace

// orientation_cube.hpp:10-80: the 24 proper signed permutations in
// lexicographic matrix order, as canonical quaternions.
std::array<Quat, 24> cube_seed_array()
{
    std::array<Matrix, 24> ms {};
    std::size_t count = 0;
    std::array<int, 3> perm { 0, 1, 2 };
    auto det = [](const Matrix& m) {
        return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
    };
    do {
        for (int sx : { -1, 1 })
            for (int sy : { -1, 1 })
                for (int sz : { -1, 1 }) {
                    Matrix m {};
                    m[perm[0]] = sx;
                    m[3 + perm[1]] = sy;
                    m[6 + perm[2]] = sz;
                    if (det(m) == 1) ms[count++] = m;
                }
    } while (std::next_permutation(perm.begin(), perm.end()));
    std::sort(ms.begin(), ms.end());
    std::array<Quat, 24> out {};
    for (std::size_t i = 0; i != 24; ++i) out[i] = matrix_quaternion(ms[i]);
    return out;
}

// orientations.cpp:38-43
bool valid_quaternion(const Quat& q)
{
    const double n = std::hypot(std::hypot(q[0], q[1]), std::hypot(q[2], q[3]));
    return std::isfinite(n) && n > 0;
}

// orientations.cpp:44-73
Quat canonical(Quat q)
{
    const double n = std::hypot(std::hypot(q[0], q[1]), std::hypot(q[2], q[3]));
    for (auto& v : q) {
        v /= n;
        if (v == 0) v = 0;
    }
    bool negate = q[3] < 0;
    if (q[3] == 0)
        for (int i = 0; i != 3; ++i)
            if (q[i] != 0) {
                negate = q[i] < 0;
                break;
            }
    if (negate)
        for (auto& v : q) v = -v;
    for (auto& v : q)
        if (v == 0) v = 0;
    return q;
}

// orientations.cpp:136-149: canonicalise, keep first occurrences in order.
std::vector<Quat> custom_catalog(const std::vector<Quat>& raw)
{
    std::vector<Quat> out;
    for (const auto& q : raw) {
        if (!valid_quaternion(q)) return {};
        const Quat c = canonical(q);
        if (std::find(out.begin(), out.end(), c) == out.end()) out.push_back(c);
    }
    return out;
}

// orientations.cpp:80-102
std::array<Quat, 96> cube45_orientations()
{
    std::array<Quat, 96> r {};
    const auto cubes = cube_seed_array();
    std::copy(cubes.begin(), cubes.end(), r.begin());
    constexpr double sine = 0.38268343236508977173;
    constexpr double cosine = 0.92387953251128675613;
    std::size_t idx = 24;
    for (int axis = 0; axis != 3; ++axis) {
        Quat turn { 0, 0, 0, cosine };
        turn[axis] = sine;
        const double x = turn[0], y = turn[1], z = turn[2], w = turn[3];
        for (const auto& q : cubes) {
            const double a = q[0], b = q[1], c = q[2], d = q[3];
            r[idx++] = canonical({ w * a + x * d + y * c - z * b, w * b - x * c + y * d + z * a,
                                   w * c + x * b - y * a + z * d, w * d - x * a - y * b - z * c });
        }
    }
    return r;
}

// gap_fill.cpp:381-420: legal translations keep the kernel window inside the
// environment window; a translation is free when its rounded overlap is zero.
bool free_mask(const std::array<std::int64_t, 3>& ef, const std::array<std::uint32_t, 3>& es, const KernelField& k,
               const CorrelationOut& c, std::vector<std::uint8_t>& mask, std::array<std::int64_t, 3>& low,
               std::array<std::uint32_t, 3>& legal)
{
    bool fits = true;
    for (int a = 0; a != 3; ++a) {
        low[a] = ef[a] - k.first[a];
        fits = fits && es[a] >= k.shape[a];
        legal[a] = fits ? es[a] - k.shape[a] + 1 : 0;
    }
    if (!fits) return false;
    mask.assign(std::size_t(legal[0]) * legal[1] * legal[2], 0);
    for (std::uint32_t z = 0; z != legal[2]; ++z)
        for (std::uint32_t y = 0; y != legal[1]; ++y)
            for (std::uint32_t x = 0; x != legal[0]; ++x) {
                const std::int64_t off[3] = { low[0] + x - c.translation_first[0], low[1] + y - c.translation_first[1],
                                              low[2] + z - c.translation_first[2] };
                bool inside = true;
                for (int a = 0; a != 3; ++a) inside = inside && off[a] >= 0 && off[a] < std::int64_t(c.shape[a]);
                mask[x + std::size_t(legal[0]) * (y + std::size_t(legal[1]) * z)] =
                    inside && std::round(c.values[off[0] + std::size_t(c.shape[0]) * (off[1] + std::size_t(c.shape[1]) * off[2])]) == 0;
            }
    return true;
}

// gap_fill.cpp:363-422 (one correlation per rotation) and :461-498 (choice):
// per rotation the first free translation in (z, y, x) order; across
// rotations the smallest key t + kernel_first compared as (z, y, x), where a
// strict comparison keeps the earliest rotation on ties.
std::string place_step(const std::vector<std::uint8_t>& env, const std::array<std::int64_t, 3>& ef,
                       const std::array<std::uint32_t, 3>& es, const std::vector<KernelField>& kernels, bool direct,
                       StepChoice& choice, std::vector<std::vector<std::int32_t>>* legal_fields,
                       const std::vector<std::pair<int, std::array<std::int64_t, 3>>>* excluded)
{
    choice = StepChoice {};
    if (legal_fields) legal_fields->assign(kernels.size(), {});
    CorrelationLimits lim;
    lim.max_working_bytes = ~0ULL >> 1;
    lim.max_padded_cells = ~0ULL >> 1;
    lim.max_direct_terms = ~0ULL;
    for (std::size_t o = 0; o != kernels.size(); ++o) {
        const auto& k = kernels[o];
        CorrelationSpec s { es, k.shape, ef, k.first };
        CorrelationOut c;
        if (direct) {
            std::vector<int> b(env.begin(), env.end()), kk(k.cells.begin(), k.cells.end());
            const std::array<int, 3> origin { int(k.first[0] - ef[0]), int(k.first[1] - ef[1]), int(k.first[2] - ef[2]) };
            bool fits = true;
            for (int a = 0; a != 3; ++a) fits = fits && es[a] >= k.shape[a];
            if (!fits) continue;
            // Only legal translations are needed: evaluate them directly.
            c.translation_first = { ef[0] - k.first[0] - (std::int64_t(k.shape[0]) - 1),
                                    ef[1] - k.first[1] - (std::int64_t(k.shape[1]) - 1),
                                    ef[2] - k.first[2] - (std::int64_t(k.shape[2]) - 1) };
            c.shape = { es[0] + k.shape[0] - 1, es[1] + k.shape[1] - 1, es[2] + k.shape[2] - 1 };
            c.values.assign(std::size_t(c.shape[0]) * c.shape[1] * c.shape[2], 1.0);
            (void)origin;
            for (std::uint32_t z = 0; z + k.shape[2] <= es[2]; ++z)
                for (std::uint32_t y = 0; y + k.shape[1] <= es[1]; ++y)
                    for (std::uint32_t x = 0; x + k.shape[0] <= es[0]; ++x) {
                        std::int64_t sum = 0;
                        for (std::uint32_t kz = 0; kz != k.shape[2]; ++kz)
                            for (std::uint32_t ky = 0; ky != k.shape[1]; ++ky) {
                                const std::uint8_t* er = env.data() + x + std::size_t(es[0]) * (y + ky + std::size_t(es[1]) * (z + kz));
                                const std::uint8_t* kr = k.cells.data() + std::size_t(k.shape[0]) * (ky + std::size_t(k.shape[1]) * kz);
                                for (std::uint32_t kx = 0; kx != k.shape[0]; ++kx) sum += er[kx] * kr[kx];
                            }
                        // translation t = ef - kf + (x,y,z) sits at output index t - translation_first
                        const std::size_t ox = x + k.shape[0] - 1, oy = y + k.shape[1] - 1, oz = z + k.shape[2] - 1;
                        c.values[ox + std::size_t(c.shape[0]) * (oy + std::size_t(c.shape[1]) * oz)] = double(sum);
                    }
        }
        else {
            c = correlate_binary(s, env.data(), env.size(), k.cells.data(), k.cells.size(), lim);
            if (!c.code.empty()) return c.code;
        }
        std::vector<std::uint8_t> mask;
        std::array<std::int64_t, 3> low;
        std::array<std::uint32_t, 3> legal;
        if (!free_mask(ef, es, k, c, mask, low, legal)) continue;
        if (legal_fields) {
            auto& f = (*legal_fields)[o];
            f.resize(mask.size());
            for (std::uint32_t z = 0; z != legal[2]; ++z)
                for (std::uint32_t y = 0; y != legal[1]; ++y)
                    for (std::uint32_t x = 0; x != legal[0]; ++x) {
                        const std::int64_t off[3] = { low[0] + x - c.translation_first[0], low[1] + y - c.translation_first[1],
                                                      low[2] + z - c.translation_first[2] };
                        f[x + std::size_t(legal[0]) * (y + std::size_t(legal[1]) * z)] = std::int32_t(
                            c.values[off[0] + std::size_t(c.shape[0]) * (off[1] + std::size_t(c.shape[1]) * off[2])]);
                    }
        }
        // gap_fill.cpp:517: a rejected proposal's translation is no longer free
        if (excluded)
            for (const auto& [eo, et] : *excluded) {
                if (eo != int(o)) continue;
                const std::int64_t u[3] = { et[0] - low[0], et[1] - low[1], et[2] - low[2] };
                if (u[0] < 0 || u[1] < 0 || u[2] < 0 || u[0] >= legal[0] || u[1] >= legal[1] || u[2] >= legal[2]) continue;
                mask[u[0] + std::size_t(legal[0]) * (u[1] + std::size_t(legal[1]) * u[2])] = 0;
            }
        std::size_t cursor = 0;
        for (; cursor != mask.size(); ++cursor)
            if (mask[cursor]) break;
        if (cursor == mask.size()) continue;
        const std::int64_t cx = cursor % legal[0], cy = (cursor / legal[0]) % legal[1], cz = cursor / (std::size_t(legal[0]) * legal[1]);
        const std::array<std::int64_t, 3> t { low[0] + cx, low[1] + cy, low[2] + cz };
        const std::array<std::int64_t, 3> key { t[2] + k.first[2], t[1] + k.first[1], t[0] + k.first[0] };
        if (choice.orientation < 0 || key < choice.key) {
            choice.orientation = int(o);
            choice.translation = t;
            choice.key = key;
        }
    }
    return {};
}

// gap_fill.cpp:98-121: mark every nonzero kernel cell inside the window.
void stamp(std::vector<std::uint8_t>& env, const std::array<std::int64_t, 3>& ef, const std::array<std::uint32_t, 3>& es,
           const KernelField& k, const std::array<std::int64_t, 3>& t)
{
    for (std::uint32_t z = 0; z != k.shape[2]; ++z)
        for (std::uint32_t y = 0; y != k.shape[1]; ++y)
            for (std::uint32_t x = 0; x != k.shape[0]; ++x) {
                if (!k.cells[x + std::size_t(k.shape[0]) * (y + std::size_t(k.shape[1]) * z)]) continue;
                const std::int64_t l[3] = { t[0] + k.first[0] + x - ef[0], t[1] + k.first[1] + y - ef[1],
                                            t[2] + k.first[2] + z - ef[2] };
                bool inside = true;
                for (int a = 0; a != 3; ++a) inside = inside && l[a] >= 0 && l[a] < std::int64_t(es[a]);
                if (inside) env[l[0] + std::size_t(es[0]) * (l[1] + std::size_t(es[1]) * l[2])] = 1;
            }
}

// engine/pack_solver/src/field_proximity.hpp:32-132: exact squared distance to
// the nearest occupied cell by per-axis integer lower envelopes, then
// exp(-sqrt(d2)/2) (0 where no cell is occupied).
void build_proximity(const std::vector<std::uint8_t>& occ, const std::array<std::uint32_t, 3>& shape,
                     std::vector<double>& values, std::vector<std::uint64_t>* squared)
{
    const std::size_t n = occ.size();
    values.assign(n, 0.0);
    for (std::size_t i = 0; i != n; ++i) values[i] = occ[i] ? 0. : std::numeric_limits<double>::infinity();
    const std::uint32_t longest = *std::max_element(shape.begin(), shape.end());
    struct Site {
        std::uint32_t coordinate;
        std::int64_t start;
    };
    std::vector<double> line(longest);
    std::vector<Site> sites;
    const std::size_t strides[3] = { 1, shape[0], std::size_t(shape[0]) * shape[1] };
    for (int axis = 0; axis != 3; ++axis) {
        const auto length = shape[axis];
        const auto stride = strides[axis];
        for (std::size_t base = 0; base != n; ++base) {
            if ((base / stride) % length != 0) continue;
            sites.clear();
            for (std::uint32_t q = 0; q != length; ++q) line[q] = values[base + q * stride];
            for (std::uint32_t q = 0; q != length; ++q) {
                if (!std::isfinite(line[q])) continue;
                std::int64_t start = std::numeric_limits<std::int64_t>::min();
                while (!sites.empty()) {
                    const std::int64_t p = sites.back().coordinate, nx = q;
                    const std::int64_t num = std::int64_t(line[q]) - std::int64_t(line[p]) + nx * nx - p * p;
                    const std::int64_t den = 2 * (nx - p);
                    start = num / den + (num % den < 0 ? -1 : 0) + 1;
                    if (start > sites.back().start) break;
                    sites.pop_back();
                }
                if (sites.empty()) start = std::numeric_limits<std::int64_t>::min();
                sites.push_back({ q, start });
            }
            std::size_t site = 0;
            for (std::uint32_t q = 0; q != length; ++q) {
                while (site + 1 < sites.size() && sites[site + 1].start <= q) ++site;
                if (sites.empty()) values[base + q * stride] = std::numeric_limits<double>::infinity();
                else {
                    const std::int64_t d = std::int64_t(q) - sites[site].coordinate;
                    values[base + q * stride] = line[sites[site].coordinate] + double(d * d);
                }
            }
        }
    }
    if (squared) {
        squared->resize(n);
        for (std::size_t i = 0; i != n; ++i)
            (*squared)[i] = std::isfinite(values[i]) ? std::uint64_t(values[i]) : ~0ULL;
    }
    for (auto& v : values) v = std::isfinite(v) ? std::exp(-std::sqrt(v) / 2.) : 0.;
}

}  // namespace sporacle
