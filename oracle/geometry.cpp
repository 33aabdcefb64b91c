// Oracle (test infrastructure): interval arithmetic, placement, boundary raster,
// component fill and container fields.  Reference paths are relative to the
// reference repository root; "vk" = engine/pack_geometry/src/validation_kernel.cpp,
// "cf" = engine/pack_geometry/src/conservative_fields.cpp.
#include <algorithm>
#include <bit>
#include <cmath>
#include <deque>
#include <limits>

#include "sp_oracle.hpp"

namespace sporacle {

namespace {
constexpr double kInf = std::numeric_limits<double>::infinity();
}

// vk:291-293
Interval point_interval(double v) { return { v, v, std::isfinite(v) }; }

// vk:295-300: round-to-nearest result widened one ulp outward on each side.
Interval widened(double low, double high)
{
    if (!std::isfinite(low) || !std::isfinite(high)) return {};
    low = std::nextafter(low, -kInf);
    high = std::nextafter(high, kInf);
    return { low, high, std::isfinite(low) && std::isfinite(high) };
}

// vk:302-305
Interval add_interval(Interval a, Interval b)
{
    if (!a.valid || !b.valid) return {};
    return widened(a.low + b.low, a.high + b.high);
}

// vk:307-310
Interval subtract_interval(Interval a, Interval b)
{
    if (!a.valid || !b.valid) return {};
    return widened(a.low - b.high, a.high - b.low);
}

// vk:312-318: min/max of the four endpoint products (first minimum wins ties).
Interval multiply_interval(Interval a, Interval b)
{
    if (!a.valid || !b.valid) return {};
    const double v[4] = { a.low * b.low, a.low * b.high, a.high * b.low, a.high * b.high };
    for (double x : v)
        if (!std::isfinite(x)) return {};
    double lo = v[0], hi = v[0];
    for (int i = 1; i != 4; ++i) {
        if (v[i] < lo) lo = v[i];
        if (hi < v[i]) hi = v[i];
    }
    return widened(lo, hi);
}

// vk:320-324
Interval divide_interval(Interval n, Interval d)
{
    if (!n.valid || !d.valid || d.low <= 0.0) return {};
    const Interval reciprocal = widened(1.0 / d.high, 1.0 / d.low);
    return multiply_interval(n, reciprocal);
}

// vk:490-508: outward enclosure of the quaternion's rotation matrix.
std::array<std::array<Interval, 3>, 3> rotation_intervals(const Quat& q)
{
    Interval v[4];
    for (int i = 0; i != 4; ++i) v[i] = point_interval(q[i]);
    auto sq = [&](int i) { return multiply_interval(v[i], v[i]); };
    auto twice = [](Interval x) { return multiply_interval(point_interval(2.0), x); };
    const Interval d = add_interval(add_interval(sq(0), sq(1)), add_interval(sq(2), sq(3)));
    std::array<std::array<Interval, 3>, 3> h;
    h[0][0] = subtract_interval(add_interval(sq(3), sq(0)), add_interval(sq(1), sq(2)));
    h[0][1] = twice(subtract_interval(multiply_interval(v[0], v[1]), multiply_interval(v[2], v[3])));
    h[0][2] = twice(add_interval(multiply_interval(v[0], v[2]), multiply_interval(v[1], v[3])));
    h[1][0] = twice(add_interval(multiply_interval(v[0], v[1]), multiply_interval(v[2], v[3])));
    h[1][1] = subtract_interval(add_interval(sq(3), sq(1)), add_interval(sq(0), sq(2)));
    h[1][2] = twice(subtract_interval(multiply_interval(v[1], v[2]), multiply_interval(v[0], v[3])));
    h[2][0] = twice(subtract_interval(multiply_interval(v[0], v[2]), multiply_interval(v[1], v[3])));
    h[2][1] = twice(add_interval(multiply_interval(v[1], v[2]), multiply_interval(v[0], v[3])));
    h[2][2] = subtract_interval(add_interval(sq(3), sq(2)), add_interval(sq(0), sq(1)));
    for (auto& row : h)
        for (auto& e : row) e = divide_interval(e, d);
    return h;
}

// vk:510-517
Interval transformed_interval(const Vec3& p, const Vec3& t, const std::array<std::array<Interval, 3>, 3>& r, int row)
{
    Interval value = point_interval(t[row]);
    for (int c = 0; c != 3; ++c) value = add_interval(value, multiply_interval(r[row][c], point_interval(p[c])));
    return value;
}

namespace {

// vk:1942-1966
IVec3 isub(const IVec3& a, const IVec3& b)
{
    return { subtract_interval(a[0], b[0]), subtract_interval(a[1], b[1]), subtract_interval(a[2], b[2]) };
}
IVec3 icross(const IVec3& a, const IVec3& b)
{
    return { subtract_interval(multiply_interval(a[1], b[2]), multiply_interval(a[2], b[1])),
             subtract_interval(multiply_interval(a[2], b[0]), multiply_interval(a[0], b[2])),
             subtract_interval(multiply_interval(a[0], b[1]), multiply_interval(a[1], b[0])) };
}
Interval idot(const IVec3& a, const IVec3& b)
{
    Interval r = point_interval(0.0);
    for (int i = 0; i != 3; ++i) r = add_interval(r, multiply_interval(a[i], b[i]));
    return r;
}

// vk:2038-2052
bool projection_separates(const std::array<IVec3, 3>& tri, const IVec3& box, const IVec3& axis)
{
    Interval tp = idot(tri[0], axis);
    for (int v = 1; v != 3; ++v) {
        const Interval p = idot(tri[v], axis);
        if (!p.valid) return false;
        tp.low = std::min(tp.low, p.low);
        tp.high = std::max(tp.high, p.high);
    }
    const Interval bp = idot(box, axis);
    return tp.valid && bp.valid && (tp.high < bp.low || bp.high < tp.low);
}

}  // namespace

// vk:2054-2079: closed-cell separating-axis test on outward enclosures.
bool triangle_box_disjoint(const std::array<IVec3, 3>& tri, const IVec3& box)
{
    for (int a = 0; a != 3; ++a) {
        double lo = tri[0][a].low, hi = tri[0][a].high;
        for (int v = 1; v != 3; ++v) {
            lo = std::min(lo, tri[v][a].low);
            hi = std::max(hi, tri[v][a].high);
        }
        if (hi < box[a].low || box[a].high < lo) return true;
    }
    const IVec3 first = isub(tri[1], tri[0]);
    const IVec3 second = isub(tri[2], tri[0]);
    if (projection_separates(tri, box, icross(first, second))) return true;
    const IVec3 axes[3] = { { point_interval(1), point_interval(0), point_interval(0) },
                            { point_interval(0), point_interval(1), point_interval(0) },
                            { point_interval(0), point_interval(0), point_interval(1) } };
    const IVec3 edges[3] = { first, second, isub(tri[2], tri[1]) };
    for (const auto& e : edges)
        for (const auto& ax : axes)
            if (projection_separates(tri, box, icross(e, ax))) return true;
    return false;
}

// vk:2081-2100: outward enclosure of closed cell idx.
bool grid_cell_interval(const Window& w, const std::array<std::int64_t, 3>& idx, IVec3& out)
{
    constexpr std::int64_t kLargest = std::int64_t { 1 } << 53;
    for (int a = 0; a != 3; ++a) {
        if (idx[a] < -kLargest || idx[a] >= kLargest) return false;
        const Interval lo = add_interval(point_interval(w.origin[a]), multiply_interval(point_interval(w.pitch),
                                                                                        point_interval(double(idx[a]))));
        const Interval hi = add_interval(point_interval(w.origin[a]),
                                         multiply_interval(point_interval(w.pitch), point_interval(double(idx[a] + 1))));
        if (!lo.valid || !hi.valid) return false;
        out[a] = { std::min(lo.low, hi.low), std::max(lo.high, hi.high), true };
    }
    return true;
}

// vk:384-438: an 8-vertex, 12-triangle axis-aligned cuboid certificate.
bool certify_cuboid(const Mesh& m, Vec3& lo, Vec3& hi)
{
    if (m.vertices.size() != 8 || m.triangles.size() != 12) return false;
    lo = hi = m.vertices.front();
    for (const auto& v : m.vertices)
        for (int a = 0; a != 3; ++a) {
            if (!std::isfinite(v[a])) return false;
            lo[a] = std::min(lo[a], v[a]);
            hi[a] = std::max(hi[a], v[a]);
        }
    for (int a = 0; a != 3; ++a)
        if (!(lo[a] < hi[a])) return false;
    bool corners[8] = {};
    for (const auto& v : m.vertices) {
        unsigned mask = 0;
        for (int a = 0; a != 3; ++a) {
            if (v[a] == hi[a]) mask |= 1U << a;
            else if (v[a] != lo[a]) return false;
        }
        if (corners[mask]) return false;
        corners[mask] = true;
    }
    for (bool c : corners)
        if (!c) return false;
    struct FaceMasks {
        unsigned values[2] {};
        int size {};
    } faces[6];
    for (const auto& t : m.triangles) {
        bool assigned = false;
        for (int a = 0; a != 3 && !assigned; ++a)
            for (int side = 0; side != 2; ++side) {
                const double plane = side == 0 ? lo[a] : hi[a];
                if (m.vertices[t[0]][a] == plane && m.vertices[t[1]][a] == plane && m.vertices[t[2]][a] == plane) {
                    unsigned mask = 0;
                    for (auto i : t) {
                        const auto& v = m.vertices[i];
                        const int b = (a + 1) % 3, c = (a + 2) % 3;
                        const unsigned corner = (v[b] == hi[b] ? 1U : 0U) | (v[c] == hi[c] ? 2U : 0U);
                        mask |= 1U << corner;
                    }
                    auto& f = faces[a * 2 + side];
                    if (f.size == 2) return false;
                    f.values[f.size++] = mask;
                    assigned = true;
                    break;
                }
            }
        if (!assigned) return false;
    }
    for (const auto& f : faces)
        if (f.size != 2 || (f.values[0] | f.values[1]) != 0xFU || std::popcount(f.values[0] & f.values[1]) != 2)
            return false;
    return true;
}

// vk:454-488: a signed-permutation rotation recognised from equal-magnitude components.
bool cardinal_rotation(const Quat& quat, std::array<int, 3>& src, std::array<int, 3>& sgn)
{
    double magnitude = 0;
    int nonzero = 0;
    int q[4] = {};
    for (int i = 0; i != 4; ++i) {
        if (quat[i] == 0.0) continue;
        const double cur = std::abs(quat[i]);
        if (nonzero == 0) magnitude = cur;
        else if (std::bit_cast<std::uint64_t>(cur) != std::bit_cast<std::uint64_t>(magnitude)) return false;
        q[i] = std::signbit(quat[i]) ? -1 : 1;
        ++nonzero;
    }
    if (nonzero != 1 && nonzero != 2 && nonzero != 4) return false;
    const int x = q[0], y = q[1], z = q[2], w = q[3], d = nonzero;
    const int h[3][3] = { { w * w + x * x - y * y - z * z, 2 * (x * y - z * w), 2 * (x * z + y * w) },
                          { 2 * (x * y + z * w), w * w - x * x + y * y - z * z, 2 * (y * z - x * w) },
                          { 2 * (x * z - y * w), 2 * (y * z + x * w), w * w - x * x - y * y + z * z } };
    bool used[3] = {};
    for (int r = 0; r != 3; ++r) {
        int found = -1;
        for (int c = 0; c != 3; ++c) {
            if (h[r][c] == d || h[r][c] == -d) {
                if (found != -1) return false;
                found = c;
                sgn[r] = h[r][c] / d;
            }
            else if (h[r][c] != 0)
                return false;
        }
        if (found < 0 || used[found]) return false;
        used[found] = true;
        src[r] = found;
    }
    return true;
}

// vk:829-943 with translation (0,0,0) (the lattice origin used for kernels).
bool place(const Mesh& m, const Quat& q, Placed& p)
{
    p = Placed {};
    p.mesh = &m;
    p.q = q;
    for (double c : q)
        if (!std::isfinite(c)) return false;
    if (q[0] == 0 && q[1] == 0 && q[2] == 0 && q[3] == 0) return false;
    if (m.vertices.empty()) return false;
    p.mesh_min = p.mesh_max = m.vertices.front();
    for (const auto& v : m.vertices)
        for (int a = 0; a != 3; ++a) {
            p.mesh_min[a] = std::min(p.mesh_min[a], v[a]);
            p.mesh_max[a] = std::max(p.mesh_max[a], v[a]);
        }
    Vec3 clo, chi;
    p.is_cuboid = certify_cuboid(m, clo, chi);
    p.is_cardinal = cardinal_rotation(q, p.source_axis, p.sign);
    const Vec3 t { 0.0, 0.0, 0.0 };
    p.intervals.resize(m.vertices.size());
    if (p.is_cuboid && p.is_cardinal) {
        for (int w = 0; w != 3; ++w) {
            const int l = p.source_axis[w], s = p.sign[w];
            const double lo = s > 0 ? clo[l] : chi[l];
            const double hi = s > 0 ? chi[l] : clo[l];
            p.cuboid_min[w] = t[w] + s * lo;
            p.cuboid_max[w] = t[w] + s * hi;
            p.bounds_min[w] = std::nextafter(t[w] + s * lo, -kInf);
            p.bounds_max[w] = std::nextafter(t[w] + s * hi, kInf);
        }
        p.world.resize(m.vertices.size());
        p.represented_world_exact = true;  // translation 0: t + s*v is exact
        for (std::size_t i = 0; i != m.vertices.size(); ++i)
            for (int a = 0; a != 3; ++a) {
                const double sum = t[a] + p.sign[a] * m.vertices[i][p.source_axis[a]];
                p.world[i][a] = sum;
                p.intervals[i][a] = widened(sum, sum);
            }
        return true;
    }
    const auto r = rotation_intervals(q);
    if (p.is_cardinal) p.world.resize(m.vertices.size());
    p.represented_world_exact = p.is_cardinal;
    bool first = true;
    for (std::size_t i = 0; i != m.vertices.size(); ++i) {
        for (int a = 0; a != 3; ++a) {
            const Interval iv = transformed_interval(m.vertices[i], t, r, a);
            if (!iv.valid) return false;
            p.intervals[i][a] = iv;
            if (p.is_cardinal) p.world[i][a] = t[a] + p.sign[a] * m.vertices[i][p.source_axis[a]];
            if (first) {
                p.bounds_min[a] = iv.low;
                p.bounds_max[a] = iv.high;
            }
            else {
                p.bounds_min[a] = std::min(p.bounds_min[a], iv.low);
                p.bounds_max[a] = std::max(p.bounds_max[a], iv.high);
            }
        }
        first = false;
    }
    return true;
}

// cf:506-519 (quotient in long double, as the reference computes it) and cf:950-960.
bool trim_window(const Placed& p, const Vec3& origin, double pitch, Window& w)
{
    constexpr long double kLargest = 9007199254740992.0L;
    w.origin = origin;
    w.pitch = pitch;
    for (int a = 0; a != 3; ++a) {
        std::int64_t idx[2];
        for (int lower = 0; lower != 2; ++lower) {
            const double bound = lower ? p.bounds_min[a] : p.bounds_max[a];
            const long double ratio = (static_cast<long double>(bound) - origin[a]) / pitch;
            if (!std::isfinite(ratio) || ratio <= -kLargest + 1 || ratio >= kLargest) return false;
            auto i = static_cast<std::int64_t>(std::floor(ratio));
            idx[lower] = lower ? i - 1 : i + 1;
        }
        const std::int64_t first = idx[1], last = idx[0];
        if (last < first || std::uint64_t(last - first) + 1 > 0xffffffffULL) return false;
        w.first[a] = first;
        w.shape[a] = static_cast<std::uint32_t>(last - first + 1);
    }
    return true;
}

// vk:3248-3340 (face candidate ranges) and vk:3668-3762 (per-cell closed SAT).
// A cell becomes 2 when some face's enclosure is not separated from it.
void rasterize_boundary(const Placed& p, const Window& w, std::vector<std::uint8_t>& tags)
{
    const std::int64_t last_w[3] = { w.first[0] + std::int64_t(w.shape[0]) - 1, w.first[1] + std::int64_t(w.shape[1]) - 1,
                                     w.first[2] + std::int64_t(w.shape[2]) - 1 };
    for (const auto& f : p.mesh->triangles) {
        const std::array<IVec3, 3> tri { p.intervals[f[0]], p.intervals[f[1]], p.intervals[f[2]] };
        std::int64_t first[3], last[3];
        bool outside = false;
        for (int a = 0; a != 3; ++a) {
            double lo = tri[0][a].low, hi = tri[0][a].high;
            for (int v = 1; v != 3; ++v) {
                lo = std::min(lo, tri[v][a].low);
                hi = std::max(hi, tri[v][a].high);
            }
            const long double l = (static_cast<long double>(lo) - w.origin[a]) / w.pitch;
            const long double h = (static_cast<long double>(hi) - w.origin[a]) / w.pitch;
            first[a] = static_cast<std::int64_t>(std::floor(l)) - 1;
            last[a] = static_cast<std::int64_t>(std::floor(h)) + 1;
            if (last[a] < w.first[a] || first[a] > last_w[a]) outside = true;
            first[a] = std::max(first[a], w.first[a]);
            last[a] = std::min(last[a], last_w[a]);
        }
        if (outside) continue;
        for (std::int64_t z = first[2]; z <= last[2]; ++z)
            for (std::int64_t y = first[1]; y <= last[1]; ++y)
                for (std::int64_t x = first[0]; x <= last[0]; ++x) {
                    const std::uint64_t flat = std::uint64_t(x - w.first[0]) +
                                               std::uint64_t(w.shape[0]) *
                                                   (std::uint64_t(y - w.first[1]) +
                                                    std::uint64_t(w.shape[1]) * std::uint64_t(z - w.first[2]));
                    if (tags[flat] == 2) continue;
                    IVec3 cell;
                    if (!grid_cell_interval(w, { x, y, z }, cell)) continue;
                    if (!triangle_box_disjoint(tri, cell)) tags[flat] = 2;
                }
    }
}

namespace {

constexpr double kSlopes[12][2] = { { 0.137, 0.271 }, { 0.223, 0.419 }, { 0.347, 0.163 }, { 0.431, 0.593 },
                                    { 0.557, 0.317 }, { 0.619, 0.733 }, { 0.709, 0.467 }, { 0.823, 0.197 },
                                    { 0.911, 0.541 }, { 0.293, 0.887 }, { 0.487, 0.773 }, { 0.677, 0.929 } };

double midpoint(double a, double b) { return (a + b) / 2; }  // std::midpoint for finite normal values

// vk:1895-1940 for a solid placed at translation 0.
Decision point_in_material(const Vec3& point, const Placed& t)
{
    if (t.is_cuboid && t.is_cardinal) {
        bool strict = true;
        for (int a = 0; a != 3; ++a) {
            if (point[a] < t.cuboid_min[a] || point[a] > t.cuboid_max[a]) return Decision::no;
            strict = strict && point[a] > t.cuboid_min[a] && point[a] < t.cuboid_max[a];
        }
        return strict ? Decision::yes : Decision::indeterminate;
    }
    if (!t.represented_world_exact) return Decision::indeterminate;
    const double span = std::max({ t.bounds_max[0] - t.bounds_min[0], t.bounds_max[1] - t.bounds_min[1],
                                   t.bounds_max[2] - t.bounds_min[2], 1.0 });
    for (const auto& s : kSlopes) {
        const Vec3 end { std::max(point[0], t.bounds_max[0]) + 2 * span + 1, point[1] + s[0] * span,
                         point[2] + s[1] * span };
        std::uint64_t crossings = 0;
        bool retry = false;
        for (const auto& f : t.mesh->triangles) {
            const std::array<Vec3, 3> tri { t.world[f[0]], t.world[f[1]], t.world[f[2]] };
            const Crossing c = segment_triangle_crossing(point, end, tri);
            if (c == Crossing::crossing) ++crossings;
            else if (c != Crossing::none) {
                retry = true;
                break;
            }
        }
        if (!retry) return crossings % 2 == 0 ? Decision::no : Decision::yes;
    }
    return Decision::indeterminate;
}

// vk:2144-2194: ray parity against the local (unplaced) mesh.
Decision local_point_in_material(const Vec3& point, const Placed& t)
{
    const Mesh& m = *t.mesh;
    const double span = std::max({ t.mesh_max[0] - t.mesh_min[0], t.mesh_max[1] - t.mesh_min[1],
                                   t.mesh_max[2] - t.mesh_min[2], 1.0 });
    for (const auto& s : kSlopes) {
        const Vec3 end { std::max(point[0], t.mesh_max[0]) + 2 * span + 1, point[1] + s[0] * span,
                         point[2] + s[1] * span };
        std::uint64_t crossings = 0;
        bool retry = false;
        for (const auto& f : m.triangles) {
            const std::array<Vec3, 3> tri { m.vertices[f[0]], m.vertices[f[1]], m.vertices[f[2]] };
            bool disjoint = false;
            for (int a = 0; a != 3; ++a) {
                const double lo = std::min({ tri[0][a], tri[1][a], tri[2][a] });
                const double hi = std::max({ tri[0][a], tri[1][a], tri[2][a] });
                disjoint = disjoint || std::max(point[a], end[a]) < lo || std::min(point[a], end[a]) > hi;
            }
            if (disjoint) continue;
            const Crossing c = segment_triangle_crossing(point, end, tri);
            if (c == Crossing::crossing) ++crossings;
            else if (c != Crossing::none) {
                retry = true;
                break;
            }
        }
        if (!retry) return crossings % 2 == 0 ? Decision::no : Decision::yes;
    }
    return Decision::indeterminate;
}

// vk:3345-3399
Decision local_interval_point_in_material(const Placed& s, const Vec3& wmin, const Vec3& wmax)
{
    for (int a = 0; a != 3; ++a)
        if (wmax[a] < s.bounds_min[a] || wmin[a] > s.bounds_max[a]) return Decision::no;
    Interval shifted[3];
    for (int a = 0; a != 3; ++a) shifted[a] = subtract_interval({ wmin[a], wmax[a], true }, point_interval(0.0));
    const auto r = rotation_intervals(s.q);
    IVec3 local;
    for (int la = 0; la != 3; ++la) {
        local[la] = point_interval(0.0);
        for (int wa = 0; wa != 3; ++wa) local[la] = add_interval(local[la], multiply_interval(r[wa][la], shifted[wa]));
        if (!local[la].valid) return Decision::indeterminate;
    }
    const Mesh& m = *s.mesh;
    for (const auto& f : m.triangles) {
        std::array<IVec3, 3> tri;
        for (int v = 0; v != 3; ++v)
            for (int a = 0; a != 3; ++a) tri[v][a] = point_interval(m.vertices[f[v]][a]);
        if (!triangle_box_disjoint(tri, local)) return Decision::indeterminate;
    }
    Vec3 witness;
    for (int a = 0; a != 3; ++a) {
        witness[a] = midpoint(local[a].low, local[a].high);
        if (!std::isfinite(witness[a])) return Decision::indeterminate;
    }
    return local_point_in_material(witness, s);
}

}  // namespace

// vk:3401-3438
Decision classify_material_witness(const Placed& p, const Vec3& wmin, const Vec3& wmax)
{
    for (int a = 0; a != 3; ++a) {
        if (!std::isfinite(wmin[a]) || !std::isfinite(wmax[a]) || wmin[a] > wmax[a]) return Decision::indeterminate;
    }
    for (int a = 0; a != 3; ++a)
        if (wmax[a] < p.bounds_min[a] || wmin[a] > p.bounds_max[a]) return Decision::no;
    if (p.represented_world_exact) {
        Vec3 rep;
        for (int a = 0; a != 3; ++a) rep[a] = midpoint(wmin[a], wmax[a]);
        return point_in_material(rep, p);
    }
    return local_interval_point_in_material(p, wmin, wmax);
}

// cf:257-360: 6-connected components of non-boundary cells in scan order; the
// first cell of each component is its witness.
void fill_components(const Placed& p, const Window& w, std::vector<std::uint8_t>& tags, std::uint64_t* components,
                     std::uint64_t* uncertain)
{
    const std::uint64_t n = w.cells();
    std::vector<std::uint8_t> seen(n, 0);
    std::vector<std::uint64_t> queue;
    queue.reserve(n);
    const std::uint32_t sx = w.shape[0], sy = w.shape[1], sz = w.shape[2];
    std::uint64_t comps = 0, unc = 0;
    for (std::uint32_t z = 0; z < sz; ++z)
        for (std::uint32_t y = 0; y < sy; ++y)
            for (std::uint32_t x = 0; x < sx; ++x) {
                const std::uint64_t start = x + std::uint64_t(sx) * (y + std::uint64_t(sy) * z);
                if (tags[start] == 2 || seen[start]) continue;
                const std::size_t begin = queue.size();
                seen[start] = 1;
                queue.push_back(start);
                for (std::size_t c = begin; c < queue.size(); ++c) {
                    const std::uint64_t cur = queue[c];
                    const std::int64_t cx = cur % sx, rest = cur / sx, cy = rest % sy, cz = rest / sy;
                    const int d[6][3] = { { -1, 0, 0 }, { 1, 0, 0 }, { 0, -1, 0 }, { 0, 1, 0 }, { 0, 0, -1 }, { 0, 0, 1 } };
                    for (const auto& dd : d) {
                        const std::int64_t nx = cx + dd[0], ny = cy + dd[1], nz = cz + dd[2];
                        if (nx < 0 || ny < 0 || nz < 0 || nx >= sx || ny >= sy || nz >= sz) continue;
                        const std::uint64_t adj = nx + std::uint64_t(sx) * (ny + std::uint64_t(sy) * nz);
                        if (!seen[adj] && tags[adj] != 2) {
                            seen[adj] = 1;
                            queue.push_back(adj);
                        }
                    }
                }
                ++comps;
                IVec3 cell;
                Decision dec = Decision::indeterminate;
                if (grid_cell_interval(w, { w.first[0] + x, w.first[1] + y, w.first[2] + z }, cell)) {
                    const Vec3 lo { cell[0].low, cell[1].low, cell[2].low };
                    const Vec3 hi { cell[0].high, cell[1].high, cell[2].high };
                    dec = classify_material_witness(p, lo, hi);
                }
                if (dec == Decision::indeterminate) unc += queue.size() - begin;
                const std::uint8_t tag = dec == Decision::indeterminate ? 3 : dec == Decision::yes ? 1 : 0;
                for (std::size_t i = begin; i < queue.size(); ++i) tags[queue[i]] = tag;
            }
    if (components) *components = comps;
    if (uncertain) *uncertain = unc;
}

// cf:919-1003
bool voxelize_object(const Mesh& m, const Vec3& origin, double pitch, const Quat& q, Window& w,
                     std::vector<std::uint8_t>& cells, std::vector<std::uint8_t>* raw_tags)
{
    if (origin[0] != 0 || origin[1] != 0 || origin[2] != 0 || !(pitch > 0)) return false;
    Placed p;
    if (!place(m, q, p)) return false;
    if (!trim_window(p, origin, pitch, w)) return false;
    std::vector<std::uint8_t> tags(w.cells(), 0);
    rasterize_boundary(p, w, tags);
    fill_components(p, w, tags);
    if (raw_tags) *raw_tags = tags;
    cells.resize(tags.size());
    for (std::size_t i = 0; i != tags.size(); ++i) cells[i] = tags[i] == 0 ? 0 : 1;
    return true;
}

// vk:3457-3490 for each cell (cf:1146-1198): free (0) only when certified
// wholly inside the box shrunk by the clearance.
bool voxelize_box_container(const Window& w, const Vec3& dims, double clearance, std::vector<std::uint8_t>& cells)
{
    if (!(dims[0] > 0 && dims[1] > 0 && dims[2] > 0) || !std::isfinite(clearance) || clearance < 0) return false;
    cells.assign(w.cells(), 1);
    for (std::uint32_t z = 0; z < w.shape[2]; ++z)
        for (std::uint32_t y = 0; y < w.shape[1]; ++y)
            for (std::uint32_t x = 0; x < w.shape[0]; ++x) {
                const std::array<std::int64_t, 3> idx { w.first[0] + x, w.first[1] + y, w.first[2] + z };
                IVec3 cell;
                if (!grid_cell_interval(w, idx, cell)) return false;
                Decision d = Decision::yes;
                for (int a = 0; a != 3 && d == Decision::yes; ++a) {
                    const Interval& c = cell[a];
                    if (c.low < clearance) {
                        if (c.high < clearance) {
                            d = Decision::no;
                            break;
                        }
                        const double v[1] = { clearance };
                        const int s[1] = { -1 };
                        if (exact_grid_sign(w.origin[a], w.pitch, idx[a], v, s, 1) < 0) {
                            d = Decision::no;
                            break;
                        }
                    }
                    const Interval permitted = subtract_interval(point_interval(dims[a]), point_interval(clearance));
                    if (permitted.valid && c.high <= permitted.low) continue;
                    if (permitted.valid && c.low > permitted.high) {
                        d = Decision::no;
                        break;
                    }
                    const double v[2] = { dims[a], clearance };
                    const int s[2] = { -1, 1 };
                    if (exact_grid_sign(w.origin[a], w.pitch, idx[a] + 1, v, s, 2) > 0) d = Decision::no;
                }
                cells[x + std::uint64_t(w.shape[0]) * (y + std::uint64_t(w.shape[1]) * z)] = d == Decision::yes ? 0 : 1;
            }
    return true;
}

// vk:3492-3522
Decision euclidean_offset_reaches(const std::array<std::int64_t, 3>& d, double pitch, double clearance)
{
    if (!(pitch > 0.0) || !std::isfinite(pitch) || clearance < 0.0 || !std::isfinite(clearance))
        return Decision::indeterminate;
    std::uint64_t sum = 0;
    for (auto v : d) {
        if (v == std::numeric_limits<std::int64_t>::min()) return Decision::indeterminate;
        const std::uint64_t mag = std::uint64_t(v < 0 ? -v : v);
        const std::uint64_t gap = mag > 1 ? mag - 1 : 0;
        sum += gap * gap;
    }
    if (sum > (std::uint64_t { 1 } << 53)) return Decision::indeterminate;
    const double p2 = std::nextafter(pitch * pitch, -kInf);
    const double dist = std::nextafter(p2 * static_cast<double>(sum), -kInf);
    const double c2 = std::nextafter(clearance * clearance, kInf);
    if (!std::isfinite(dist) || !std::isfinite(c2)) return Decision::indeterminate;
    return dist > c2 ? Decision::no : Decision::yes;
}

// cf:1200-1275: raster + fill of the container solid in the window widened by
// the clearance halo (cf:184-204), then every cell that is not certified
// material blocks the stencil around it (cf:436-504, invert_material).
bool voxelize_mesh_container(const Mesh& m, const Window& w, double clearance, std::vector<std::uint8_t>& cells)
{
    if (!std::isfinite(clearance) || clearance < 0) return false;
    std::int64_t halo = 0;
    if (clearance != 0.0) {
        const long double r = std::ceil(static_cast<long double>(clearance) / w.pitch) + 1.0L;
        halo = static_cast<std::int64_t>(r);
    }
    Window work = w;
    for (int a = 0; a != 3; ++a) {
        work.first[a] -= halo;
        work.shape[a] += static_cast<std::uint32_t>(2 * halo);
    }
    Placed p;
    if (!place(m, { 0, 0, 0, 1 }, p)) return false;
    std::vector<std::uint8_t> tags(work.cells(), 0);
    rasterize_boundary(p, work, tags);
    fill_components(p, work, tags);
    std::vector<std::array<std::int64_t, 3>> stencil;
    for (std::int64_t z = -halo; z <= halo; ++z)
        for (std::int64_t y = -halo; y <= halo; ++y)
            for (std::int64_t x = -halo; x <= halo; ++x)
                if (euclidean_offset_reaches({ x, y, z }, w.pitch, clearance) != Decision::no) stencil.push_back({ x, y, z });
    cells.assign(w.cells(), 0);
    for (std::uint32_t z = 0; z < work.shape[2]; ++z)
        for (std::uint32_t y = 0; y < work.shape[1]; ++y)
            for (std::uint32_t x = 0; x < work.shape[0]; ++x) {
                const auto v = tags[x + std::uint64_t(work.shape[0]) * (y + std::uint64_t(work.shape[1]) * z)];
                if (v == 1) continue;
                const std::int64_t src[3] = { work.first[0] + x, work.first[1] + y, work.first[2] + z };
                for (const auto& d : stencil) {
                    std::int64_t t[3];
                    bool inside = true;
                    for (int a = 0; a != 3; ++a) {
                        t[a] = src[a] + d[a];
                        if (t[a] < w.first[a] || t[a] > w.first[a] + std::int64_t(w.shape[a]) - 1) inside = false;
                    }
                    if (inside)
                        cells[std::uint64_t(t[0] - w.first[0]) +
                              std::uint64_t(w.shape[0]) *
                                  (std::uint64_t(t[1] - w.first[1]) + std::uint64_t(w.shape[1]) * std::uint64_t(t[2] - w.first[2]))] = 1;
                }
            }
    return true;
}

}  // namespace sporacle
