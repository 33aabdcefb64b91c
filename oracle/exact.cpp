// Oracle (test infrastructure): exact orientation predicates and the
// segment/triangle crossing classifier of
// engine/pack_geometry/src/exact_predicates.cpp:236-300,562-745,1056-1110.
//
// The reference decides signs with an interval filter and an exact
// multi-limb fallback.  Here the fallback is floating-point expansion
// arithmetic (Shewchuk 1997, "Adaptive Precision Floating-Point Arithmetic and
// Fast Robust Geometric Predicates": Two-Sum, Two-Product, Grow-Expansion),
// which yields the same exact sign.
#include <algorithm>
#include <bit>
#include <cmath>
#include <limits>
#include <vector>

#include "sp_oracle.hpp"

namespace sporacle {
namespace {

using Expansion = std::vector<double>;  // nonoverlapping, increasing magnitude

void two_sum(double a, double b, double& x, double& y)
{
    x = a + b;
    const double bv = x - a;
    const double av = x - bv;
    y = (a - av) + (b - bv);
}

void two_product(double a, double b, double& x, double& y)
{
    x = a * b;
    y = std::fma(a, b, -x);
}

// Grow-Expansion: e + b, zero components dropped.
Expansion grow(const Expansion& e, double b)
{
    Expansion h;
    h.reserve(e.size() + 1);
    double q = b;
    for (double ei : e) {
        double s, t;
        two_sum(q, ei, s, t);
        if (t != 0.0) h.push_back(t);
        q = s;
    }
    if (q != 0.0 || h.empty()) h.push_back(q);
    return h;
}

Expansion sum(const Expansion& e, const Expansion& f)
{
    Expansion h = e;
    for (double fi : f) h = grow(h, fi);
    return h;
}

Expansion scale(const Expansion& e, double b)
{
    Expansion h { 0.0 };
    for (double ei : e) {
        double p, q;
        two_product(ei, b, p, q);
        h = grow(h, q);
        h = grow(h, p);
    }
    return h;
}

Expansion product(const Expansion& e, const Expansion& f)
{
    Expansion h { 0.0 };
    for (double fi : f) h = sum(h, scale(e, fi));
    return h;
}

Expansion negate(Expansion e)
{
    for (double& v : e) v = -v;
    return e;
}

Expansion diff(double a, double b)
{
    double x, y;
    two_sum(a, -b, x, y);
    return y != 0.0 ? Expansion { y, x } : Expansion { x };
}

int sign_of(const Expansion& e)
{
    for (auto it = e.rbegin(); it != e.rend(); ++it)
        if (*it != 0.0) return *it > 0 ? 1 : -1;
    return 0;
}

// exact_predicates.cpp:620-647
bool same_coordinate(double a, double b)
{
    if (std::bit_cast<std::uint64_t>(a) == std::bit_cast<std::uint64_t>(b)) return true;
    return ((std::bit_cast<std::uint64_t>(a) | std::bit_cast<std::uint64_t>(b)) & 0x7fffffffffffffffULL) == 0;
}

int orient3_exact(const Vec3& a, const Vec3& b, const Vec3& c, const Vec3& d)
{
    const Expansion ax = diff(a[0], d[0]), ay = diff(a[1], d[1]), az = diff(a[2], d[2]);
    const Expansion bx = diff(b[0], d[0]), by = diff(b[1], d[1]), bz = diff(b[2], d[2]);
    const Expansion cx = diff(c[0], d[0]), cy = diff(c[1], d[1]), cz = diff(c[2], d[2]);
    const Expansion m0 = sum(product(by, cz), negate(product(bz, cy)));
    const Expansion m1 = sum(product(bx, cz), negate(product(bz, cx)));
    const Expansion m2 = sum(product(bx, cy), negate(product(by, cx)));
    const Expansion det = sum(sum(product(ax, m0), negate(product(ay, m1))), product(az, m2));
    return sign_of(det);
}

int orient2_exact(const std::array<double, 2>& a, const std::array<double, 2>& b, const std::array<double, 2>& c)
{
    const Expansion ax = diff(a[0], c[0]), ay = diff(a[1], c[1]);
    const Expansion bx = diff(b[0], c[0]), by = diff(b[1], c[1]);
    return sign_of(sum(product(ax, by), negate(product(ay, bx))));
}

Sign to_sign(int s) { return s > 0 ? Sign::positive : s < 0 ? Sign::negative : Sign::zero; }

using P2 = std::array<double, 2>;

P2 project(const Vec3& p, int dropped)
{
    if (dropped == 0) return { p[1], p[2] };
    if (dropped == 1) return { p[0], p[2] };
    return { p[0], p[1] };
}

// exact_predicates.cpp:254-262
int projection_for(const std::array<Vec3, 3>& t)
{
    for (int dropped = 0; dropped < 3; ++dropped)
        if (orient2d(project(t[0], dropped), project(t[1], dropped), project(t[2], dropped)) != Sign::zero)
            return dropped;
    return -1;
}

// exact_predicates.cpp:264-271
bool point_on_segment(const P2& p, const P2& a, const P2& b)
{
    if (orient2d(a, b, p) != Sign::zero) return false;
    return p[0] >= std::min(a[0], b[0]) && p[0] <= std::max(a[0], b[0]) && p[1] >= std::min(a[1], b[1]) &&
           p[1] <= std::max(a[1], b[1]);
}

// exact_predicates.cpp:273-288
bool segments_intersect(const P2& a, const P2& b, const P2& c, const P2& d)
{
    const Sign abc = orient2d(a, b, c), abd = orient2d(a, b, d), cda = orient2d(c, d, a), cdb = orient2d(c, d, b);
    if (abc != Sign::zero && abd != Sign::zero && abc != abd && cda != Sign::zero && cdb != Sign::zero && cda != cdb)
        return true;
    return (abc == Sign::zero && point_on_segment(c, a, b)) || (abd == Sign::zero && point_on_segment(d, a, b)) ||
           (cda == Sign::zero && point_on_segment(a, c, d)) || (cdb == Sign::zero && point_on_segment(b, c, d));
}

bool nonneg(Sign s) { return s != Sign::negative; }
bool nonpos(Sign s) { return s != Sign::positive; }

// exact_predicates.cpp:290-298
bool point_in_triangle(const P2& p, const std::array<P2, 3>& t)
{
    const Sign s0 = orient2d(t[0], t[1], p), s1 = orient2d(t[1], t[2], p), s2 = orient2d(t[2], t[0], p);
    return (nonneg(s0) && nonneg(s1) && nonneg(s2)) || (nonpos(s0) && nonpos(s1) && nonpos(s2));
}

}  // namespace

// exact_predicates.cpp:705-742 (structural zero, interval filter, exact).
Sign orient3d(const Vec3& a, const Vec3& b, const Vec3& c, const Vec3& d)
{
    for (int ax = 0; ax != 3; ++ax)
        if (same_coordinate(a[ax], b[ax]) && same_coordinate(b[ax], c[ax]) && same_coordinate(c[ax], d[ax]))
            return Sign::zero;
    Interval x[3], y[3], z[3];
    for (int ax = 0; ax != 3; ++ax) {
        x[ax] = subtract_interval(point_interval(a[ax]), point_interval(d[ax]));
        y[ax] = subtract_interval(point_interval(b[ax]), point_interval(d[ax]));
        z[ax] = subtract_interval(point_interval(c[ax]), point_interval(d[ax]));
    }
    const Interval m0 = subtract_interval(multiply_interval(y[1], z[2]), multiply_interval(y[2], z[1]));
    const Interval m1 = subtract_interval(multiply_interval(y[0], z[2]), multiply_interval(y[2], z[0]));
    const Interval m2 = subtract_interval(multiply_interval(y[0], z[1]), multiply_interval(y[1], z[0]));
    const Interval det = add_interval(subtract_interval(multiply_interval(x[0], m0), multiply_interval(x[1], m1)),
                                      multiply_interval(x[2], m2));
    if (det.valid && std::isfinite(det.low) && std::isfinite(det.high)) {
        if (det.low > 0.0) return Sign::positive;
        if (det.high < 0.0) return Sign::negative;
    }
    return to_sign(orient3_exact(a, b, c, d));
}

// exact_predicates.cpp:667-703
Sign orient2d(const P2& a, const P2& b, const P2& c)
{
    for (int ax = 0; ax != 2; ++ax)
        if (same_coordinate(a[ax], b[ax]) && same_coordinate(b[ax], c[ax])) return Sign::zero;
    const Interval ax = subtract_interval(point_interval(a[0]), point_interval(c[0]));
    const Interval ay = subtract_interval(point_interval(a[1]), point_interval(c[1]));
    const Interval bx = subtract_interval(point_interval(b[0]), point_interval(c[0]));
    const Interval by = subtract_interval(point_interval(b[1]), point_interval(c[1]));
    const Interval det = subtract_interval(multiply_interval(ax, by), multiply_interval(ay, bx));
    if (det.valid && std::isfinite(det.low) && std::isfinite(det.high)) {
        if (det.low > 0.0) return Sign::positive;
        if (det.high < 0.0) return Sign::negative;
    }
    return to_sign(orient2_exact(a, b, c));
}

// exact_predicates.cpp:1056-1110
Crossing segment_triangle_crossing(const Vec3& p, const Vec3& q, const std::array<Vec3, 3>& t)
{
    const Sign s1 = orient3d(t[0], t[1], t[2], p);
    const Sign s2 = orient3d(t[0], t[1], t[2], q);
    if (s1 == Sign::zero && s2 == Sign::zero) {
        const int dropped = projection_for(t);
        if (dropped < 0) return Crossing::degenerate;  // reference: uncertain (also a retry)
        const std::array<P2, 3> pt { project(t[0], dropped), project(t[1], dropped), project(t[2], dropped) };
        const P2 pp = project(p, dropped), qq = project(q, dropped);
        if (point_in_triangle(pp, pt) || point_in_triangle(qq, pt)) return Crossing::degenerate;
        for (int e = 0; e != 3; ++e)
            if (segments_intersect(pp, qq, pt[e], pt[(e + 1) % 3])) return Crossing::degenerate;
        return Crossing::none;
    }
    if (s1 == Sign::zero) {
        const int dropped = projection_for(t);
        if (dropped < 0) return Crossing::degenerate;
        const std::array<P2, 3> pt { project(t[0], dropped), project(t[1], dropped), project(t[2], dropped) };
        return point_in_triangle(project(p, dropped), pt) ? Crossing::boundary : Crossing::none;
    }
    if (s2 == Sign::zero) return Crossing::degenerate;
    if (s1 == s2) return Crossing::none;
    const Sign e0 = orient3d(p, q, t[0], t[1]);
    const Sign e1 = orient3d(p, q, t[1], t[2]);
    const Sign e2 = orient3d(p, q, t[2], t[0]);
    if (e0 == Sign::zero || e1 == Sign::zero || e2 == Sign::zero) return Crossing::degenerate;
    return (e0 == e1 && e1 == e2) ? Crossing::crossing : Crossing::none;
}

// vk:2102-2142: exact sign of origin + pitch*index + sum(sign_i * v_i).
int exact_grid_sign(double origin, double pitch, std::int64_t index, const double* values, const int* signs, int n)
{
    const double ri = static_cast<double>(index);
    double hi, lo;
    two_product(pitch, ri, hi, lo);
    Expansion e { origin };
    e = grow(e, lo);
    e = grow(e, hi);
    for (int i = 0; i != n; ++i) e = grow(e, signs[i] < 0 ? -values[i] : values[i]);
    return sign_of(e);
}

}  // namespace sporacle
