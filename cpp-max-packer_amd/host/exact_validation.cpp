// Exact host validation of final placements (engine/pack_geometry/src/
// validation.cpp:517-634 for the pair and containment checks; predicates in
// the manner of exact_predicates.cpp: a floating filter, exact evaluation
// when the filter cannot decide).  See exact_validation.hpp for the verdicts.
#include "spectrapack/geometry/exact_validation.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <optional>

namespace spectrapack::geometry::exact {
namespace {

// ---- arbitrary-precision integers (sign and 32-bit limbs) -----------------
class BigInt {
public:
    BigInt() = default;
    static BigInt from_u64(std::uint64_t v, int sign)
    {
        BigInt r;
        if (v == 0) return r;
        r.mag_.push_back(std::uint32_t(v));
        if (v >> 32) r.mag_.push_back(std::uint32_t(v >> 32));
        r.sign_ = sign;
        return r;
    }
    int sign() const { return sign_; }
    bool is_zero() const { return sign_ == 0; }
    BigInt neg() const
    {
        BigInt r = *this;
        r.sign_ = -r.sign_;
        return r;
    }
    // number of trailing zero bits of the magnitude (0 for zero)
    int trailing_zeros() const
    {
        if (sign_ == 0) return 0;
        int n = 0;
        std::size_t i = 0;
        while (mag_[i] == 0) {
            n += 32;
            ++i;
        }
        std::uint32_t w = mag_[i];
        while (!(w & 1u)) {
            w >>= 1;
            ++n;
        }
        return n;
    }
    BigInt shl(int bits) const
    {
        if (sign_ == 0 || bits == 0) return *this;
        BigInt r;
        r.sign_ = sign_;
        const int words = bits / 32, b = bits % 32;
        r.mag_.assign(words, 0u);
        std::uint32_t carry = 0;
        for (std::uint32_t w : mag_) {
            r.mag_.push_back(b ? (w << b) | carry : w);
            carry = b ? w >> (32 - b) : 0;
        }
        if (carry) r.mag_.push_back(carry);
        return r;
    }
    BigInt shr(int bits) const  // exact only when the low bits are zero
    {
        if (sign_ == 0 || bits == 0) return *this;
        BigInt r;
        r.sign_ = sign_;
        const int words = bits / 32, b = bits % 32;
        for (std::size_t i = std::size_t(words); i < mag_.size(); ++i) {
            std::uint32_t w = mag_[i] >> b;
            if (b && i + 1 < mag_.size()) w |= mag_[i + 1] << (32 - b);
            r.mag_.push_back(w);
        }
        r.trim();
        return r;
    }
    friend BigInt operator+(const BigInt& a, const BigInt& b)
    {
        if (a.sign_ == 0) return b;
        if (b.sign_ == 0) return a;
        if (a.sign_ == b.sign_) {
            BigInt r;
            r.sign_ = a.sign_;
            r.mag_ = add_mag(a.mag_, b.mag_);
            return r;
        }
        const int c = cmp_mag(a.mag_, b.mag_);
        if (c == 0) return BigInt {};
        BigInt r;
        if (c > 0) {
            r.sign_ = a.sign_;
            r.mag_ = sub_mag(a.mag_, b.mag_);
        }
        else {
            r.sign_ = b.sign_;
            r.mag_ = sub_mag(b.mag_, a.mag_);
        }
        r.trim();
        return r;
    }
    friend BigInt operator-(const BigInt& a, const BigInt& b) { return a + b.neg(); }
    friend BigInt operator*(const BigInt& a, const BigInt& b)
    {
        if (a.sign_ == 0 || b.sign_ == 0) return BigInt {};
        BigInt r;
        r.sign_ = a.sign_ * b.sign_;
        r.mag_.assign(a.mag_.size() + b.mag_.size(), 0u);
        for (std::size_t i = 0; i != a.mag_.size(); ++i) {
            std::uint64_t carry = 0;
            for (std::size_t j = 0; j != b.mag_.size(); ++j) {
                const std::uint64_t t = std::uint64_t(a.mag_[i]) * b.mag_[j] + r.mag_[i + j] + carry;
                r.mag_[i + j] = std::uint32_t(t);
                carry = t >> 32;
            }
            std::size_t k = i + b.mag_.size();
            while (carry) {
                const std::uint64_t t = std::uint64_t(r.mag_[k]) + carry;
                r.mag_[k++] = std::uint32_t(t);
                carry = t >> 32;
            }
        }
        r.trim();
        return r;
    }

private:
    static std::vector<std::uint32_t> add_mag(const std::vector<std::uint32_t>& a, const std::vector<std::uint32_t>& b)
    {
        std::vector<std::uint32_t> r(std::max(a.size(), b.size()) + 1, 0u);
        std::uint64_t carry = 0;
        for (std::size_t i = 0; i + 1 < r.size(); ++i) {
            const std::uint64_t t = std::uint64_t(i < a.size() ? a[i] : 0) + (i < b.size() ? b[i] : 0) + carry;
            r[i] = std::uint32_t(t);
            carry = t >> 32;
        }
        r.back() = std::uint32_t(carry);
        while (!r.empty() && r.back() == 0) r.pop_back();
        return r;
    }
    static std::vector<std::uint32_t> sub_mag(const std::vector<std::uint32_t>& a, const std::vector<std::uint32_t>& b)
    {
        std::vector<std::uint32_t> r(a.size(), 0u);
        std::int64_t borrow = 0;
        for (std::size_t i = 0; i != a.size(); ++i) {
            std::int64_t t = std::int64_t(a[i]) - (i < b.size() ? b[i] : 0) - borrow;
            borrow = t < 0;
            if (t < 0) t += (std::int64_t(1) << 32);
            r[i] = std::uint32_t(t);
        }
        return r;
    }
    static int cmp_mag(const std::vector<std::uint32_t>& a, const std::vector<std::uint32_t>& b)
    {
        if (a.size() != b.size()) return a.size() < b.size() ? -1 : 1;
        for (std::size_t i = a.size(); i-- > 0;)
            if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
        return 0;
    }
    void trim()
    {
        while (!mag_.empty() && mag_.back() == 0) mag_.pop_back();
        if (mag_.empty()) sign_ = 0;
    }
    std::vector<std::uint32_t> mag_;
    int sign_ {};
};

// ---- exact dyadic numbers m * 2^e ------------------------------------------
struct Exact {
    BigInt m;
    int e {};

    static Exact of(double v)
    {
        Exact r;
        if (v == 0.0) return r;
        int k;
        const double f = std::frexp(std::fabs(v), &k);  // |v| = f 2^k, f in [0.5, 1)
        const auto mant = static_cast<std::uint64_t>(std::ldexp(f, 53));
        r.m = BigInt::from_u64(mant, v < 0 ? -1 : 1);
        r.e = k - 53;
        r.normalise();
        return r;
    }
    int sign() const { return m.sign(); }
    void normalise()
    {
        if (m.is_zero()) {
            e = 0;
            return;
        }
        const int tz = m.trailing_zeros();
        if (tz) {
            m = m.shr(tz);
            e += tz;
        }
    }
    friend Exact operator+(const Exact& a, const Exact& b)
    {
        if (a.m.is_zero()) return b;
        if (b.m.is_zero()) return a;
        const int e = std::min(a.e, b.e);
        Exact r;
        r.m = a.m.shl(a.e - e) + b.m.shl(b.e - e);
        r.e = e;
        r.normalise();
        return r;
    }
    friend Exact operator-(const Exact& a, const Exact& b)
    {
        Exact n = b;
        n.m = n.m.neg();
        return a + n;
    }
    friend Exact operator*(const Exact& a, const Exact& b)
    {
        Exact r;
        r.m = a.m * b.m;
        r.e = a.e + b.e;
        if (r.m.is_zero()) r.e = 0;
        return r;
    }
};

// ---- outward-rounded intervals (rounded to nearest, then widened by one
// ulp per side, as validation_kernel.cpp:291-324 widens) -----------------
struct Iv {
    double lo, hi;
};
inline double down(double x) { return std::nextafter(x, -HUGE_VAL); }
inline double up(double x) { return std::nextafter(x, HUGE_VAL); }
inline Iv pt(double v) { return { v, v }; }
inline Iv operator+(Iv a, Iv b) { return { down(a.lo + b.lo), up(a.hi + b.hi) }; }
inline Iv operator-(Iv a, Iv b) { return { down(a.lo - b.hi), up(a.hi - b.lo) }; }
inline Iv operator*(Iv a, Iv b)
{
    const double p[4] = { a.lo * b.lo, a.lo * b.hi, a.hi * b.lo, a.hi * b.hi };
    return { down(*std::min_element(p, p + 4)), up(*std::max_element(p, p + 4)) };
}
inline Iv divide_positive(Iv a, Iv d)  // d.lo > 0
{
    const double p[4] = { a.lo / d.lo, a.lo / d.hi, a.hi / d.lo, a.hi / d.hi };
    return { down(*std::min_element(p, p + 4)), up(*std::max_element(p, p + 4)) };
}
inline int certain_sign(Iv v) { return v.lo > 0 ? 1 : v.hi < 0 ? -1 : 0; }

// ---- posed meshes ------------------------------------------------------------
// validation_kernel.cpp:490-517: rotation = H(q) / |q|^2.  The exact vertex is
// p = (|q|^2 t + H v) / |q|^2, kept as homogeneous (X, Y, Z, W = |q|^2).
struct Posed {
    const Mesh* mesh {};
    std::array<Exact, 4> q;  // x, y, z, w
    std::array<Exact, 3> t;
    std::array<std::array<Exact, 3>, 3> h;
    Exact w;  // |q|^2
    std::vector<std::array<Iv, 3>> iv;   // vertex enclosures
    mutable std::vector<std::array<Exact, 4>> hx;  // exact homogeneous vertices, computed on demand
    mutable std::vector<char> have;
    std::array<Iv, 3> lo {}, hi {};      // enclosure of the vertex bounds
    std::vector<std::array<Iv, 3>> tri_lo, tri_hi;  // per-triangle bounds (lo.lo .. hi.hi used)
    std::array<std::uint32_t, 3> min_v {}, max_v {};  // vertices with the exact least / greatest coordinate
};

struct Ref {
    const Posed* p;
    std::uint32_t v;
};

struct Counters {
    std::uint64_t exact = 0;
};

const std::array<Exact, 4>& exact_vertex(const Posed& P, std::uint32_t v)
{
    if (!P.have[v]) {
        const auto& x = P.mesh->vertices[v];
        const std::array<Exact, 3> ex { Exact::of(x[0]), Exact::of(x[1]), Exact::of(x[2]) };
        for (int r = 0; r != 3; ++r) {
            Exact acc = P.w * P.t[r];
            for (int c = 0; c != 3; ++c) acc = acc + P.h[r][c] * ex[c];
            P.hx[v][r] = acc;
        }
        P.hx[v][3] = P.w;
        P.have[v] = 1;
    }
    return P.hx[v];
}

Posed pose_mesh(const Mesh& mesh, const Pose& pose)
{
    Posed P;
    P.mesh = &mesh;
    for (int i = 0; i != 4; ++i) P.q[i] = Exact::of(pose.rotation_xyzw[i]);
    for (int i = 0; i != 3; ++i) P.t[i] = Exact::of(pose.translation_mm[i]);
    const auto& q = P.q;
    auto sq = [&](int i) { return q[i] * q[i]; };
    const Exact two = Exact::of(2.0);
    P.w = (sq(0) + sq(1)) + (sq(2) + sq(3));
    P.h[0][0] = (sq(3) + sq(0)) - (sq(1) + sq(2));
    P.h[0][1] = two * (q[0] * q[1] - q[2] * q[3]);
    P.h[0][2] = two * (q[0] * q[2] + q[1] * q[3]);
    P.h[1][0] = two * (q[0] * q[1] + q[2] * q[3]);
    P.h[1][1] = (sq(3) + sq(1)) - (sq(0) + sq(2));
    P.h[1][2] = two * (q[1] * q[2] - q[0] * q[3]);
    P.h[2][0] = two * (q[0] * q[2] - q[1] * q[3]);
    P.h[2][1] = two * (q[1] * q[2] + q[0] * q[3]);
    P.h[2][2] = (sq(3) + sq(2)) - (sq(0) + sq(1));
    // the same matrix as intervals (validation_kernel.cpp:490-508 order)
    std::array<Iv, 4> v;
    for (int i = 0; i != 4; ++i) v[i] = pt(pose.rotation_xyzw[i]);
    auto isq = [&](int i) { return v[i] * v[i]; };
    auto twice = [](Iv x) { return pt(2.0) * x; };
    const Iv d = (isq(0) + isq(1)) + (isq(2) + isq(3));
    std::array<std::array<Iv, 3>, 3> r;
    r[0][0] = (isq(3) + isq(0)) - (isq(1) + isq(2));
    r[0][1] = twice(v[0] * v[1] - v[2] * v[3]);
    r[0][2] = twice(v[0] * v[2] + v[1] * v[3]);
    r[1][0] = twice(v[0] * v[1] + v[2] * v[3]);
    r[1][1] = (isq(3) + isq(1)) - (isq(0) + isq(2));
    r[1][2] = twice(v[1] * v[2] - v[0] * v[3]);
    r[2][0] = twice(v[0] * v[2] - v[1] * v[3]);
    r[2][1] = twice(v[1] * v[2] + v[0] * v[3]);
    r[2][2] = (isq(3) + isq(2)) - (isq(0) + isq(1));
    for (auto& row : r)
        for (auto& e : row) e = divide_positive(e, d);
    const std::size_t n = mesh.vertices.size();
    P.iv.resize(n);
    P.hx.resize(n);
    P.have.assign(n, 0);
    for (int a = 0; a != 3; ++a) {
        P.lo[a] = { HUGE_VAL, HUGE_VAL };
        P.hi[a] = { -HUGE_VAL, -HUGE_VAL };
    }
    for (std::size_t i = 0; i != n; ++i) {
        const auto& x = mesh.vertices[i];
        for (int row = 0; row != 3; ++row) {
            Iv acc = pt(pose.translation_mm[row]);
            for (int c = 0; c != 3; ++c) acc = acc + r[row][c] * pt(x[c]);
            P.iv[i][row] = acc;
            P.lo[row].lo = std::min(P.lo[row].lo, acc.lo);
            P.lo[row].hi = std::min(P.lo[row].hi, acc.hi);
            P.hi[row].lo = std::max(P.hi[row].lo, acc.lo);
            P.hi[row].hi = std::max(P.hi[row].hi, acc.hi);
        }
    }
    P.tri_lo.resize(mesh.triangles.size());
    P.tri_hi.resize(mesh.triangles.size());
    for (std::size_t t = 0; t != mesh.triangles.size(); ++t)
        for (int a = 0; a != 3; ++a) {
            double l = HUGE_VAL, h = -HUGE_VAL;
            for (std::uint32_t k : mesh.triangles[t]) {
                l = std::min(l, P.iv[k][a].lo);
                h = std::max(h, P.iv[k][a].hi);
            }
            P.tri_lo[t][a] = { l, l };
            P.tri_hi[t][a] = { h, h };
        }
    return P;
}

// sign of X_a / W_a - X_b / W_b on axis `a`
int compare_coord(const Ref& p, const Ref& q, int axis, Counters& c)
{
    const Iv d = p.p->iv[p.v][axis] - q.p->iv[q.v][axis];
    if (const int s = certain_sign(d)) return s;
    ++c.exact;
    const auto& P = exact_vertex(*p.p, p.v);
    const auto& Q = exact_vertex(*q.p, q.v);
    return (P[axis] * Q[3] - Q[axis] * P[3]).sign();
}

// sign of X_p / W_p - value on axis `a`
int compare_value(const Ref& p, int axis, double value, Counters& c)
{
    if (const int s = certain_sign(p.p->iv[p.v][axis] - pt(value))) return s;
    ++c.exact;
    const auto& P = exact_vertex(*p.p, p.v);
    return (P[axis] - Exact::of(value) * P[3]).sign();
}

// The vertices with the exact least and greatest coordinate per axis.
void exact_bounds(Posed& P, Counters& c)
{
    const std::size_t n = P.iv.size();
    for (int a = 0; a != 3; ++a) {
        double min_hi = HUGE_VAL, max_lo = -HUGE_VAL;
        for (std::size_t i = 0; i != n; ++i) {
            min_hi = std::min(min_hi, P.iv[i][a].hi);
            max_lo = std::max(max_lo, P.iv[i][a].lo);
        }
        std::uint32_t best_min = 0, best_max = 0;
        bool have_min = false, have_max = false;
        for (std::uint32_t i = 0; i != n; ++i) {
            if (P.iv[i][a].lo <= min_hi) {
                if (!have_min || compare_coord({ &P, i }, { &P, best_min }, a, c) < 0) best_min = i;
                have_min = true;
            }
            if (P.iv[i][a].hi >= max_lo) {
                if (!have_max || compare_coord({ &P, i }, { &P, best_max }, a, c) > 0) best_max = i;
                have_max = true;
            }
        }
        P.min_v[a] = best_min;
        P.max_v[a] = best_max;
    }
}

Exact det4(const std::array<const std::array<Exact, 4>*, 4>& m)
{
    auto a = [&](int r, int c) -> const Exact& { return (*m[r])[c]; };
    const Exact s0 = a(0, 0) * a(1, 1) - a(1, 0) * a(0, 1), s1 = a(0, 0) * a(1, 2) - a(1, 0) * a(0, 2),
                s2 = a(0, 0) * a(1, 3) - a(1, 0) * a(0, 3), s3 = a(0, 1) * a(1, 2) - a(1, 1) * a(0, 2),
                s4 = a(0, 1) * a(1, 3) - a(1, 1) * a(0, 3), s5 = a(0, 2) * a(1, 3) - a(1, 2) * a(0, 3);
    const Exact c5 = a(2, 2) * a(3, 3) - a(3, 2) * a(2, 3), c4 = a(2, 1) * a(3, 3) - a(3, 1) * a(2, 3),
                c3 = a(2, 1) * a(3, 2) - a(3, 1) * a(2, 2), c2 = a(2, 0) * a(3, 3) - a(3, 0) * a(2, 3),
                c1 = a(2, 0) * a(3, 2) - a(3, 0) * a(2, 2), c0 = a(2, 0) * a(3, 1) - a(3, 0) * a(2, 1);
    return s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
}

// sign of det(b - a, c - a, d - a): positive when d lies on the side of the
// plane (a, b, c) that (b - a) x (c - a) points to.
int orient3d(const Ref& a, const Ref& b, const Ref& c, const Ref& d, Counters& k)
{
    const auto& A = a.p->iv[a.v];
    const auto& B = b.p->iv[b.v];
    const auto& C = c.p->iv[c.v];
    const auto& D = d.p->iv[d.v];
    Iv u[3], v[3], w[3];
    for (int i = 0; i != 3; ++i) {
        u[i] = B[i] - A[i];
        v[i] = C[i] - A[i];
        w[i] = D[i] - A[i];
    }
    const Iv det = u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) +
                   u[2] * (v[0] * w[1] - v[1] * w[0]);
    if (const int s = certain_sign(det)) return s;
    ++k.exact;
    // det4 of the homogeneous rows (a, b, c, d) = -det3(b - a, c - a, d - a) W_a W_b W_c W_d
    return -det4({ &exact_vertex(*a.p, a.v), &exact_vertex(*b.p, b.v), &exact_vertex(*c.p, c.v),
                   &exact_vertex(*d.p, d.v) })
                .sign();
}

// The same in the coordinate plane that leaves out `drop`: sign of det(b - a, c - a).
int orient2d(const Ref& a, const Ref& b, const Ref& c, int drop, Counters& k)
{
    const int i = drop == 0 ? 1 : 0, j = drop == 2 ? 1 : 2;
    const auto& A = a.p->iv[a.v];
    const auto& B = b.p->iv[b.v];
    const auto& C = c.p->iv[c.v];
    const Iv det = (B[i] - A[i]) * (C[j] - A[j]) - (B[j] - A[j]) * (C[i] - A[i]);
    if (const int s = certain_sign(det)) return s;
    ++k.exact;
    const auto& P = exact_vertex(*a.p, a.v);
    const auto& Q = exact_vertex(*b.p, b.v);
    const auto& R = exact_vertex(*c.p, c.v);
    // det3 of homogeneous rows (u, v, W) = det2(b - a, c - a) W_a W_b W_c
    const Exact d = P[i] * (Q[j] * R[3] - Q[3] * R[j]) - P[j] * (Q[i] * R[3] - Q[3] * R[i]) +
                    P[3] * (Q[i] * R[j] - Q[j] * R[i]);
    return d.sign();
}

bool mixed(int a, int b, int c) { return (a > 0 || b > 0 || c > 0) && (a < 0 || b < 0 || c < 0); }

// c on the closed segment ab, given the three are collinear in the plane
bool on_segment_2d(const Ref& a, const Ref& b, const Ref& c, int drop, Counters& k)
{
    for (int axis = 0; axis != 3; ++axis) {
        if (axis == drop) continue;
        const int ca = compare_coord(c, a, axis, k), cb = compare_coord(c, b, axis, k);
        if (ca * cb > 0) return false;
    }
    return true;
}

bool segments_meet_2d(const Ref& a, const Ref& b, const Ref& c, const Ref& d, int drop, Counters& k)
{
    const int d1 = orient2d(a, b, c, drop, k), d2 = orient2d(a, b, d, drop, k);
    const int d3 = orient2d(c, d, a, drop, k), d4 = orient2d(c, d, b, drop, k);
    if (d1 * d2 < 0 && d3 * d4 < 0) return true;
    if (d1 == 0 && on_segment_2d(a, b, c, drop, k)) return true;
    if (d2 == 0 && on_segment_2d(a, b, d, drop, k)) return true;
    if (d3 == 0 && on_segment_2d(c, d, a, drop, k)) return true;
    if (d4 == 0 && on_segment_2d(c, d, b, drop, k)) return true;
    return false;
}

bool point_in_triangle_2d(const Ref& x, const Ref& p, const Ref& q, const Ref& r, int drop, Counters& k)
{
    return !mixed(orient2d(p, q, x, drop, k), orient2d(q, r, x, drop, k), orient2d(r, p, x, drop, k));
}

// a coordinate plane onto which the triangle projects with nonzero area
int projection_axis(const Ref& p, const Ref& q, const Ref& r, Counters& k)
{
    for (int drop = 0; drop != 3; ++drop)
        if (orient2d(p, q, r, drop, k) != 0) return drop;
    return -1;  // degenerate triangle
}

bool segment_meets_triangle_2d(const Ref& a, const Ref& b, const Ref& p, const Ref& q, const Ref& r, int drop,
                               Counters& k)
{
    return point_in_triangle_2d(a, p, q, r, drop, k) || point_in_triangle_2d(b, p, q, r, drop, k) ||
           segments_meet_2d(a, b, p, q, drop, k) || segments_meet_2d(a, b, q, r, drop, k) ||
           segments_meet_2d(a, b, r, p, drop, k);
}

enum class Contact { none, touch, cross };

// The closed segment ab against the closed triangle pqr: none; cross (a and b
// strictly on opposite sides and the crossing strictly inside the triangle);
// touch (any other common point).
Contact segment_triangle(const Ref& a, const Ref& b, const Ref& p, const Ref& q, const Ref& r, Counters& k)
{
    const int sa = orient3d(p, q, r, a, k), sb = orient3d(p, q, r, b, k);
    if (sa * sb > 0) return Contact::none;
    if (sa == 0 && sb == 0) {
        const int drop = projection_axis(p, q, r, k);
        if (drop < 0) return Contact::touch;  // degenerate face: not certified
        return segment_meets_triangle_2d(a, b, p, q, r, drop, k) ? Contact::touch : Contact::none;
    }
    const int s1 = orient3d(a, b, p, q, k), s2 = orient3d(a, b, q, r, k), s3 = orient3d(a, b, r, p, k);
    if (mixed(s1, s2, s3)) return Contact::none;
    if (sa * sb < 0 && s1 != 0 && s2 != 0 && s3 != 0) return Contact::cross;
    return Contact::touch;
}

struct TriRef {
    const Posed* p;
    std::uint32_t t;
    Ref v(int i) const { return { p, p->mesh->triangles[t][i] }; }
};

// Closed triangle against closed triangle: none, touch (common points, no
// edge crossing another face strictly inside) or cross.
Contact triangle_triangle(const TriRef& A, const TriRef& B, Counters& k)
{
    const int a0 = orient3d(B.v(0), B.v(1), B.v(2), A.v(0), k), a1 = orient3d(B.v(0), B.v(1), B.v(2), A.v(1), k),
              a2 = orient3d(B.v(0), B.v(1), B.v(2), A.v(2), k);
    if ((a0 > 0 && a1 > 0 && a2 > 0) || (a0 < 0 && a1 < 0 && a2 < 0)) return Contact::none;
    const int b0 = orient3d(A.v(0), A.v(1), A.v(2), B.v(0), k), b1 = orient3d(A.v(0), A.v(1), A.v(2), B.v(1), k),
              b2 = orient3d(A.v(0), A.v(1), A.v(2), B.v(2), k);
    if ((b0 > 0 && b1 > 0 && b2 > 0) || (b0 < 0 && b1 < 0 && b2 < 0)) return Contact::none;
    if (a0 == 0 && a1 == 0 && a2 == 0) {
        // coplanar faces: they may touch, never cross
        const int drop = projection_axis(B.v(0), B.v(1), B.v(2), k);
        if (drop < 0) return Contact::touch;
        for (int e = 0; e != 3; ++e)
            if (segment_meets_triangle_2d(A.v(e), A.v((e + 1) % 3), B.v(0), B.v(1), B.v(2), drop, k))
                return Contact::touch;
        for (int e = 0; e != 3; ++e)
            if (point_in_triangle_2d(B.v(e), A.v(0), A.v(1), A.v(2), drop, k)) return Contact::touch;
        return Contact::none;
    }
    Contact best = Contact::none;
    for (int e = 0; e != 3; ++e) {
        const Contact c1 = segment_triangle(A.v(e), A.v((e + 1) % 3), B.v(0), B.v(1), B.v(2), k);
        if (c1 == Contact::cross) return Contact::cross;
        if (c1 == Contact::touch) best = Contact::touch;
        const Contact c2 = segment_triangle(B.v(e), B.v((e + 1) % 3), A.v(0), A.v(1), A.v(2), k);
        if (c2 == Contact::cross) return Contact::cross;
        if (c2 == Contact::touch) best = Contact::touch;
    }
    return best;
}

bool boxes_overlap(const std::array<Iv, 3>& alo, const std::array<Iv, 3>& ahi, const std::array<Iv, 3>& blo,
                   const std::array<Iv, 3>& bhi)
{
    for (int a = 0; a != 3; ++a)
        if (ahi[a].hi < blo[a].lo || bhi[a].hi < alo[a].lo) return false;
    return true;
}

// Surface contact of two posed meshes: the worst contact over triangle pairs
// whose bounds meet; the triangles of touching pairs are listed when asked.
Contact surfaces(const Posed& A, const Posed& B, ValidationReport& rep, Counters& k,
                 std::vector<std::uint32_t>* touch_a = nullptr, std::vector<std::uint32_t>* touch_b = nullptr)
{
    std::vector<std::uint32_t> la, lb;
    for (std::uint32_t t = 0; t != A.tri_lo.size(); ++t)
        if (boxes_overlap(A.tri_lo[t], A.tri_hi[t], B.lo, B.hi)) la.push_back(t);
    for (std::uint32_t t = 0; t != B.tri_lo.size(); ++t)
        if (boxes_overlap(B.tri_lo[t], B.tri_hi[t], A.lo, A.hi)) lb.push_back(t);
    Contact worst = Contact::none;
    for (std::uint32_t ta : la)
        for (std::uint32_t tb : lb) {
            if (!boxes_overlap(A.tri_lo[ta], A.tri_hi[ta], B.tri_lo[tb], B.tri_hi[tb])) continue;
            ++rep.triangle_tests;
            const Contact c = triangle_triangle({ &A, ta }, { &B, tb }, k);
            if (c == Contact::cross) return c;
            if (c == Contact::touch) {
                worst = c;
                if (touch_a) touch_a->push_back(ta);
                if (touch_b) touch_b->push_back(tb);
            }
        }
    return worst;
}

// A one-point posed set: the centroid of triangle t of P (exact homogeneous
// (X0 + X1 + X2, 3 W) within one copy, whose vertices share W).
Posed centroid_point(const Posed& P, std::uint32_t t)
{
    Posed c;
    c.mesh = P.mesh;
    c.iv.assign(1, {});
    c.hx.assign(1, {});
    c.have.assign(1, 1);
    const auto& tri = P.mesh->triangles[t];
    for (int a = 0; a != 3; ++a)
        c.iv[0][a] = divide_positive(P.iv[tri[0]][a] + P.iv[tri[1]][a] + P.iv[tri[2]][a], pt(3.0));
    const auto& X0 = exact_vertex(P, tri[0]);
    const auto& X1 = exact_vertex(P, tri[1]);
    const auto& X2 = exact_vertex(P, tri[2]);
    for (int a = 0; a != 3; ++a) c.hx[0][a] = X0[a] + X1[a] + X2[a];
    c.hx[0][3] = Exact::of(3.0) * X0[3];
    return c;
}

// Exact ray parity of a point against a closed mesh: 1 inside, 0 outside,
// -1 on the surface or no ray without a degenerate hit.
int point_in_mesh(const Ref& x, const Posed& M, Counters& k)
{
    // fixed ray slopes, none parallel to a coordinate plane
    static const double dirs[6][3] = { { 1.0, 0.3718420711, 0.1913264537 },  { -0.2837152981, 1.0, 0.4419210723 },
                                       { 0.5230984111, -0.1170993011, 1.0 }, { -1.0, -0.6291021751, 0.2510137819 },
                                       { 0.3356121773, -1.0, -0.7022813357 }, { -0.4121957237, 0.8813205177, -1.0 } };
    double span = 1.0;
    for (int a = 0; a != 3; ++a)
        span = std::max({ span, std::fabs(M.lo[a].lo), std::fabs(M.hi[a].hi), std::fabs(x.p->iv[x.v][a].lo),
                          std::fabs(x.p->iv[x.v][a].hi) });
    const double len = std::ldexp(1.0, std::ilogb(span) + 3);  // a power of two beyond the scene
    for (const auto& d : dirs) {
        // the ray's far end as a one-vertex posed mesh of its own: exact
        // coordinates X + len d W share x's homogeneous weight
        Mesh far_mesh;
        Posed far;
        far.mesh = &far_mesh;
        far.iv.assign(1, {});
        for (int a = 0; a != 3; ++a) far.iv[0][a] = x.p->iv[x.v][a] + pt(len * d[a]);
        const auto& X = exact_vertex(*x.p, x.v);
        far.hx.assign(1, {});
        for (int a = 0; a != 3; ++a) far.hx[0][a] = X[a] + Exact::of(len * d[a]) * X[3];
        far.hx[0][3] = X[3];
        far.have.assign(1, 1);
        const Ref y { &far, 0 };
        std::array<Iv, 3> slo, shi;
        for (int a = 0; a != 3; ++a) {
            slo[a] = { std::min(x.p->iv[x.v][a].lo, far.iv[0][a].lo), 0 };
            shi[a] = { 0, std::max(x.p->iv[x.v][a].hi, far.iv[0][a].hi) };
        }
        int crossings = 0;
        bool degenerate = false;
        for (std::uint32_t t = 0; t != M.tri_lo.size() && !degenerate; ++t) {
            if (!boxes_overlap(M.tri_lo[t], M.tri_hi[t], slo, shi)) continue;
            const TriRef T { &M, t };
            const Contact c = segment_triangle(x, y, T.v(0), T.v(1), T.v(2), k);
            if (c == Contact::cross) ++crossings;
            else if (c == Contact::touch) {
                if (orient3d(T.v(0), T.v(1), T.v(2), x, k) == 0) {
                    const int drop = projection_axis(T.v(0), T.v(1), T.v(2), k);
                    if (drop < 0 || point_in_triangle_2d(x, T.v(0), T.v(1), T.v(2), drop, k)) return -1;
                }
                degenerate = true;
            }
        }
        if (!degenerate) return crossings & 1;
    }
    return -1;
}

// One vertex per connected component of the mesh (components share vertices
// through triangles).
std::vector<std::uint32_t> component_vertices(const Mesh& m)
{
    std::vector<std::uint32_t> parent(m.vertices.size());
    std::iota(parent.begin(), parent.end(), 0u);
    auto find = [&](std::uint32_t v) {
        while (parent[v] != v) v = parent[v] = parent[parent[v]];
        return v;
    };
    std::vector<char> used(m.vertices.size(), 0);
    for (const auto& t : m.triangles) {
        for (std::uint32_t v : t) used[v] = 1;
        const std::uint32_t r0 = find(t[0]);
        for (int i = 1; i != 3; ++i) parent[find(t[i])] = r0;
    }
    std::vector<std::uint32_t> out;
    std::vector<char> seen(m.vertices.size(), 0);
    for (std::uint32_t v = 0; v != m.vertices.size(); ++v)
        if (used[v] && !seen[find(v)]) {
            seen[find(v)] = 1;
            out.push_back(v);
        }
    return out;
}

ValidationReport verdict(Validity v, std::string_view code, std::size_t a = kNoCopy, std::size_t b = kNoCopy)
{
    ValidationReport r;
    r.validity = v;
    r.code = code;
    r.first = a;
    r.second = b;
    return r;
}

}  // namespace

struct Validator::State {
    const Mesh* part;
    Container container;
    double pair_clearance, wall_clearance;
    std::vector<std::uint32_t> part_components;
    std::unique_ptr<Posed> container_mesh;
    std::vector<std::unique_ptr<Posed>> copies;
    Counters k;

    bool centroid_inside(const Posed& P, std::vector<std::uint32_t>& tris, const Posed& other)
    {
        std::sort(tris.begin(), tris.end());
        tris.erase(std::unique(tris.begin(), tris.end()), tris.end());
        for (std::uint32_t t : tris) {
            const Posed c = centroid_point(P, t);
            if (point_in_mesh({ &c, 0 }, other, k) == 1) return true;
        }
        return false;
    }

    // bounds_pair_certificate's role: separated exact bounds on some axis
    bool bounds_certify(const Posed& A, const Posed& B)
    {
        for (int a = 0; a != 3; ++a) {
            for (int side = 0; side != 2; ++side) {
                const Posed& L = side ? B : A;  // lower copy
                const Posed& U = side ? A : B;
                // interval test first: U.min - L.max >= clearance for sure
                if ((U.lo[a] - L.hi[a]).lo >= pair_clearance) return true;
                if ((U.lo[a] - L.hi[a]).hi < pair_clearance) continue;
                const auto& Pu = exact_vertex(U, U.min_v[a]);
                const auto& Pl = exact_vertex(L, L.max_v[a]);
                ++k.exact;
                const Exact gap = Pu[a] * Pl[3] - Pl[a] * Pu[3] - Exact::of(pair_clearance) * Pu[3] * Pl[3];
                if (gap.sign() >= 0) return true;
            }
        }
        return false;
    }

    ValidationReport pair(std::size_t ia, std::size_t ib, ValidationReport& rep)
    {
        const Posed& A = *copies[ia];
        const Posed& B = *copies[ib];
        if (!boxes_overlap(A.lo, A.hi, B.lo, B.hi) && pair_clearance == 0.0) return verdict(Validity::valid, "VALID");
        if (bounds_certify(A, B)) return verdict(Validity::valid, "VALID");
        ++rep.pair_tests;
        std::vector<std::uint32_t> ta, tb;
        const Contact c = surfaces(A, B, rep, k, &ta, &tb);
        if (c == Contact::cross) return verdict(Validity::invalid, "VALIDATION_PAIR", ia, ib);
        if (c == Contact::touch) {
            // contacts without a strict crossing: overlap is still certain when
            // a touching face's centroid lies strictly inside the other copy
            if (centroid_inside(A, ta, B) || centroid_inside(B, tb, A))
                return verdict(Validity::invalid, "VALIDATION_PAIR", ia, ib);
            return verdict(Validity::indeterminate, "VALIDATION_PAIR_UNRESOLVED", ia, ib);
        }
        // disjoint surfaces: the copies overlap only if one holds the other
        for (int side = 0; side != 2; ++side) {
            const Posed& In = side ? B : A;
            const Posed& Out = side ? A : B;
            for (std::uint32_t v : part_components) {
                const int inside = point_in_mesh({ &In, v }, Out, k);
                if (inside == 1) return verdict(Validity::invalid, "VALIDATION_PAIR", ia, ib);
                if (inside < 0) return verdict(Validity::indeterminate, "VALIDATION_PAIR_UNRESOLVED", ia, ib);
            }
        }
        if (pair_clearance > 0.0) return verdict(Validity::indeterminate, "VALIDATION_PAIR_UNRESOLVED", ia, ib);
        return verdict(Validity::valid, "VALID");
    }

    ValidationReport contain(std::size_t i, ValidationReport& rep)
    {
        const Posed& P = *copies[i];
        if (const auto* box = std::get_if<BoxContainer>(&container)) {
            // every vertex within [w, L - w] exactly: the extreme vertices decide
            for (int a = 0; a != 3; ++a) {
                const Ref lo { &P, P.min_v[a] }, hi { &P, P.max_v[a] };
                bool ok_lo = compare_value(lo, a, wall_clearance, k) >= 0;
                bool ok_hi;
                if (const int s = certain_sign(pt(box->dims_mm[a]) - pt(wall_clearance) - P.iv[hi.v][a]))
                    ok_hi = s >= 0;
                else {
                    ++k.exact;
                    const auto& X = exact_vertex(P, hi.v);
                    ok_hi = ((Exact::of(box->dims_mm[a]) - Exact::of(wall_clearance)) * X[3] - X[a]).sign() >= 0;
                }
                if (!ok_lo || !ok_hi) return verdict(Validity::invalid, "VALIDATION_CONTAINER", i);
            }
            return verdict(Validity::valid, "VALID");
        }
        const Posed& C = *container_mesh;
        if (wall_clearance > 0.0) return verdict(Validity::indeterminate, "VALIDATION_CONTAINER_UNRESOLVED", i);
        ++rep.pair_tests;
        std::vector<std::uint32_t> tp;
        const Contact c = surfaces(P, C, rep, k, &tp, nullptr);
        if (c == Contact::cross) return verdict(Validity::invalid, "VALIDATION_CONTAINER", i);
        if (c == Contact::touch) {
            // a touching face whose centroid lies strictly outside the container
            std::sort(tp.begin(), tp.end());
            tp.erase(std::unique(tp.begin(), tp.end()), tp.end());
            for (std::uint32_t t : tp) {
                const Posed cp = centroid_point(P, t);
                const Ref x { &cp, 0 };
                if (point_in_mesh(x, C, k) == 0) return verdict(Validity::invalid, "VALIDATION_CONTAINER", i);
            }
            return verdict(Validity::indeterminate, "VALIDATION_CONTAINER_UNRESOLVED", i);
        }
        for (std::uint32_t v : part_components) {
            const int inside = point_in_mesh({ &P, v }, C, k);
            if (inside == 0) return verdict(Validity::invalid, "VALIDATION_CONTAINER", i);
            if (inside < 0) return verdict(Validity::indeterminate, "VALIDATION_CONTAINER_UNRESOLVED", i);
        }
        return verdict(Validity::valid, "VALID");
    }

    std::size_t add(const Pose& pose)
    {
        auto p = std::make_unique<Posed>(pose_mesh(*part, pose));
        exact_bounds(*p, k);
        copies.push_back(std::move(p));
        return copies.size() - 1;
    }
};

Validator::Validator(const Mesh& part, Container container, double pair_clearance_mm, double wall_clearance_mm)
    : state_(new State { &part, container, pair_clearance_mm, wall_clearance_mm, {}, {}, {}, {} })
{
    state_->part_components = component_vertices(part);
    if (const auto* m = std::get_if<const Mesh*>(&container)) {
        state_->container_mesh = std::make_unique<Posed>(pose_mesh(**m, Pose {}));
        exact_bounds(*state_->container_mesh, state_->k);
    }
}

Validator::~Validator() { delete state_; }

std::size_t Validator::accepted() const { return state_->copies.size(); }

ValidationReport Validator::try_add(const Pose& pose)
{
    State& s = *state_;
    const std::uint64_t exact0 = s.k.exact;
    ValidationReport rep;
    const std::size_t i = s.add(pose);
    ValidationReport v = s.contain(i, rep);
    for (std::size_t j = 0; j != i && v.validity == Validity::valid; ++j) v = s.pair(j, i, rep);
    v.pair_tests = rep.pair_tests;
    v.triangle_tests = rep.triangle_tests;
    v.exact_evaluations = s.k.exact - exact0;
    if (v.validity != Validity::valid) s.copies.pop_back();
    return v;
}

ValidationReport validate_placements(const Mesh& part, std::span<const Pose> poses, const Container& container,
                                     double pair_clearance_mm, double wall_clearance_mm)
{
    for (const auto& p : poses) {
        const double n = std::hypot(std::hypot(p.rotation_xyzw[0], p.rotation_xyzw[1]),
                                    std::hypot(p.rotation_xyzw[2], p.rotation_xyzw[3]));
        bool finite = std::isfinite(n) && n > 0;
        for (double t : p.translation_mm) finite = finite && std::isfinite(t);
        if (!finite) return verdict(Validity::indeterminate, "KERNEL_POSE_INVALID");
    }
    Validator val(part, container, pair_clearance_mm, wall_clearance_mm);
    Validator::State& s = *val.state_;
    ValidationReport rep;
    for (const auto& p : poses) s.add(p);
    // pairs first (validation.cpp:395-466), then containment (:467-499)
    std::vector<std::size_t> order(poses.size());
    std::iota(order.begin(), order.end(), std::size_t(0));
    std::sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) {
        return s.copies[a]->lo[0].lo < s.copies[b]->lo[0].lo;
    });
    for (std::size_t x = 0; x != order.size(); ++x)
        for (std::size_t y = x + 1; y != order.size(); ++y) {
            const std::size_t a = order[x], b = order[y];
            if (s.copies[b]->lo[0].lo > s.copies[a]->hi[0].hi + pair_clearance_mm) break;  // sweep on x
            ValidationReport v = s.pair(std::min(a, b), std::max(a, b), rep);
            if (v.validity != Validity::valid) {
                v.pair_tests = rep.pair_tests;
                v.triangle_tests = rep.triangle_tests;
                v.exact_evaluations = s.k.exact;
                return v;
            }
        }
    for (std::size_t i = 0; i != poses.size(); ++i) {
        ValidationReport v = s.contain(i, rep);
        if (v.validity != Validity::valid) {
            v.pair_tests = rep.pair_tests;
            v.triangle_tests = rep.triangle_tests;
            v.exact_evaluations = s.k.exact;
            return v;
        }
    }
    ValidationReport ok = verdict(Validity::valid, "VALID");
    ok.pair_tests = rep.pair_tests;
    ok.triangle_tests = rep.triangle_tests;
    ok.exact_evaluations = s.k.exact;
    return ok;
}

}  // namespace spectrapack::geometry::exact
