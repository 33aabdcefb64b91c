// Conservative voxel rasteriser on gfx950 (engine/pack_geometry): per
// orientation, outward FP64 interval placement of the mesh vertices
// (validation_kernel.cpp:290-324,490-517,829-943), the closed-cell
// separating-axis boundary raster (validation_kernel.cpp:2038-2100,
// 3248-3340, 3668-3762), 6-connected components of the non-boundary cells
// with one material witness per component (conservative_fields.cpp:257-360,
// validation_kernel.cpp:1895-1940,2144-2194,3345-3438), and the analytic box
// and mesh container fields (conservative_fields.cpp:1124-1282).
//
// Bit-exactness: every interval operation is a round-to-nearest double op
// followed by nextafter, compiled without FP contraction, so the enclosures
// equal the reference's bit for bit.  The breadth-first fill is replaced by
// lock-free union-find whose roots are the smallest linear index of each
// component, which is exactly the first cell the reference's scan visits,
// i.e. the same witness.  Orientation predicates are exact (interval filter,
// then floating-point expansion arithmetic).
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace sphip {
namespace {

constexpr double kInf = __builtin_huge_val();

struct Iv {
    double lo, hi;
    bool ok;
};

__host__ __device__ inline double dmin(double a, double b) { return b < a ? b : a; }
__host__ __device__ inline double dmax(double a, double b) { return a < b ? b : a; }

__device__ inline Iv pt(double v) { return { v, v, isfinite(v) != 0 }; }
__device__ inline Iv widen(double lo, double hi)
{
    if (!isfinite(lo) || !isfinite(hi)) return { 0, 0, false };
    lo = nextafter(lo, -kInf);
    hi = nextafter(hi, kInf);
    return { lo, hi, isfinite(lo) && isfinite(hi) };
}
__device__ inline Iv iadd(Iv a, Iv b)
{
    if (!a.ok || !b.ok) return { 0, 0, false };
    return widen(a.lo + b.lo, a.hi + b.hi);
}
__device__ inline Iv isub(Iv a, Iv b)
{
    if (!a.ok || !b.ok) return { 0, 0, false };
    return widen(a.lo - b.hi, a.hi - b.lo);
}
__device__ inline Iv imul(Iv a, Iv b)
{
    if (!a.ok || !b.ok) return { 0, 0, false };
    const double v0 = a.lo * b.lo, v1 = a.lo * b.hi, v2 = a.hi * b.lo, v3 = a.hi * b.hi;
    if (!isfinite(v0) || !isfinite(v1) || !isfinite(v2) || !isfinite(v3)) return { 0, 0, false };
    double lo = v0, hi = v0;
    lo = dmin(lo, v1);
    lo = dmin(lo, v2);
    lo = dmin(lo, v3);
    hi = dmax(hi, v1);
    hi = dmax(hi, v2);
    hi = dmax(hi, v3);
    return widen(lo, hi);
}
__device__ inline Iv idiv(Iv n, Iv d)
{
    if (!n.ok || !d.ok || d.lo <= 0.0) return { 0, 0, false };
    const Iv r = widen(1.0 / d.hi, 1.0 / d.lo);
    return imul(n, r);
}

struct IV3 {
    Iv v[3];
};

// vk:490-508
__device__ void rotation_iv(const double* q, Iv h[3][3])
{
    Iv v[4];
    for (int i = 0; i != 4; ++i) v[i] = pt(q[i]);
    auto sq = [&](int i) { return imul(v[i], v[i]); };
    auto twice = [](Iv x) { return imul(pt(2.0), x); };
    const Iv d = iadd(iadd(sq(0), sq(1)), iadd(sq(2), sq(3)));
    h[0][0] = isub(iadd(sq(3), sq(0)), iadd(sq(1), sq(2)));
    h[0][1] = twice(isub(imul(v[0], v[1]), imul(v[2], v[3])));
    h[0][2] = twice(iadd(imul(v[0], v[2]), imul(v[1], v[3])));
    h[1][0] = twice(iadd(imul(v[0], v[1]), imul(v[2], v[3])));
    h[1][1] = isub(iadd(sq(3), sq(1)), iadd(sq(0), sq(2)));
    h[1][2] = twice(isub(imul(v[1], v[2]), imul(v[0], v[3])));
    h[2][0] = twice(isub(imul(v[0], v[2]), imul(v[1], v[3])));
    h[2][1] = twice(iadd(imul(v[1], v[2]), imul(v[0], v[3])));
    h[2][2] = isub(iadd(sq(3), sq(2)), iadd(sq(0), sq(1)));
    for (int r = 0; r != 3; ++r)
        for (int c = 0; c != 3; ++c) h[r][c] = idiv(h[r][c], d);
}

__device__ inline IV3 vsub(const IV3& a, const IV3& b) { return { { isub(a.v[0], b.v[0]), isub(a.v[1], b.v[1]), isub(a.v[2], b.v[2]) } }; }
__device__ inline IV3 vcross(const IV3& a, const IV3& b)
{
    return { { isub(imul(a.v[1], b.v[2]), imul(a.v[2], b.v[1])), isub(imul(a.v[2], b.v[0]), imul(a.v[0], b.v[2])),
               isub(imul(a.v[0], b.v[1]), imul(a.v[1], b.v[0])) } };
}
__device__ inline Iv vdot(const IV3& a, const IV3& b)
{
    Iv r = pt(0.0);
    for (int i = 0; i != 3; ++i) r = iadd(r, imul(a.v[i], b.v[i]));
    return r;
}

// vk:2038-2052
__device__ bool projection_separates(const IV3* tri, const IV3& box, const IV3& axis)
{
    Iv tp = vdot(tri[0], axis);
    for (int k = 1; k != 3; ++k) {
        const Iv p = vdot(tri[k], axis);
        if (!p.ok) return false;
        tp.lo = dmin(tp.lo, p.lo);
        tp.hi = dmax(tp.hi, p.hi);
    }
    const Iv bp = vdot(box, axis);
    return tp.ok && bp.ok && (tp.hi < bp.lo || bp.hi < tp.lo);
}

// vk:2054-2079
__device__ bool triangle_box_disjoint(const IV3* tri, const IV3& box)
{
    for (int a = 0; a != 3; ++a) {
        double lo = tri[0].v[a].lo, hi = tri[0].v[a].hi;
        for (int k = 1; k != 3; ++k) {
            lo = dmin(lo, tri[k].v[a].lo);
            hi = dmax(hi, tri[k].v[a].hi);
        }
        if (hi < box.v[a].lo || box.v[a].hi < lo) return true;
    }
    const IV3 e0 = vsub(tri[1], tri[0]);
    const IV3 e1 = vsub(tri[2], tri[0]);
    if (projection_separates(tri, box, vcross(e0, e1))) return true;
    const IV3 e2 = vsub(tri[2], tri[1]);
    const IV3* edges[3] = { &e0, &e1, &e2 };
    for (int e = 0; e != 3; ++e)
        for (int a = 0; a != 3; ++a) {
            IV3 ax { { pt(0.0), pt(0.0), pt(0.0) } };
            ax.v[a] = pt(1.0);
            if (projection_separates(tri, box, vcross(*edges[e], ax))) return true;
        }
    return false;
}

// vk:2081-2100
__device__ bool grid_cell_iv(const double* origin, double pitch, const long long* idx, IV3& out)
{
    const long long big = 1LL << 53;
    for (int a = 0; a != 3; ++a) {
        if (idx[a] < -big || idx[a] >= big) return false;
        const Iv lo = iadd(pt(origin[a]), imul(pt(pitch), pt(double(idx[a]))));
        const Iv hi = iadd(pt(origin[a]), imul(pt(pitch), pt(double(idx[a] + 1))));
        if (!lo.ok || !hi.ok) return false;
        out.v[a] = { dmin(lo.lo, hi.lo), dmax(lo.hi, hi.hi), true };
    }
    return true;
}

// ---- exact orientation predicates (Shewchuk's expansion arithmetic) ------
template <int N> struct Ex {
    double c[N];
    int n;
};

__device__ inline void two_sum(double a, double b, double& x, double& y)
{
    x = a + b;
    const double bv = x - a;
    const double av = x - bv;
    y = (a - av) + (b - bv);
}
__device__ inline void two_prod(double a, double b, double& x, double& y)
{
    x = a * b;
    y = fma(a, b, -x);
}
template <int N> __device__ inline void grow(Ex<N>& e, double b)
{
    double q = b;
    int k = 0;
    for (int i = 0; i < e.n; ++i) {
        double s, t;
        two_sum(q, e.c[i], s, t);
        if (t != 0.0) e.c[k++] = t;
        q = s;
    }
    if (q != 0.0 || k == 0) e.c[k++] = q;
    e.n = k;
}
template <int N, int M> __device__ inline void add_to(Ex<N>& h, const Ex<M>& f, double sign)
{
    for (int i = 0; i < f.n; ++i) grow(h, sign * f.c[i]);
}
template <int N, int M, int K> __device__ inline void mul_into(Ex<N>& h, const Ex<M>& e, const Ex<K>& f, double sign)
{
    for (int j = 0; j < f.n; ++j)
        for (int i = 0; i < e.n; ++i) {
            double p, q;
            two_prod(e.c[i], f.c[j], p, q);
            grow(h, sign * q);
            grow(h, sign * p);
        }
}
template <int N> __device__ inline int sign_of(const Ex<N>& e)
{
    for (int i = e.n - 1; i >= 0; --i)
        if (e.c[i] != 0.0) return e.c[i] > 0 ? 1 : -1;
    return 0;
}
__device__ inline Ex<2> ediff(double a, double b)
{
    Ex<2> r;
    double x, y;
    two_sum(a, -b, x, y);
    if (y != 0.0) {
        r.c[0] = y;
        r.c[1] = x;
        r.n = 2;
    }
    else {
        r.c[0] = x;
        r.n = 1;
    }
    return r;
}

__device__ int orient3_exact(const double* a, const double* b, const double* c, const double* d)
{
    const Ex<2> ax = ediff(a[0], d[0]), ay = ediff(a[1], d[1]), az = ediff(a[2], d[2]);
    const Ex<2> bx = ediff(b[0], d[0]), by = ediff(b[1], d[1]), bz = ediff(b[2], d[2]);
    const Ex<2> cx = ediff(c[0], d[0]), cy = ediff(c[1], d[1]), cz = ediff(c[2], d[2]);
    Ex<16> m0 { {}, 0 }, m1 { {}, 0 }, m2 { {}, 0 };
    mul_into(m0, by, cz, 1.0);
    mul_into(m0, bz, cy, -1.0);
    mul_into(m1, bx, cz, 1.0);
    mul_into(m1, bz, cx, -1.0);
    mul_into(m2, bx, cy, 1.0);
    mul_into(m2, by, cx, -1.0);
    Ex<200> det { {}, 0 };
    mul_into(det, ax, m0, 1.0);
    mul_into(det, ay, m1, -1.0);
    mul_into(det, az, m2, 1.0);
    return sign_of(det);
}

__device__ int orient2_exact(const double* a, const double* b, const double* c)
{
    const Ex<2> ax = ediff(a[0], c[0]), ay = ediff(a[1], c[1]);
    const Ex<2> bx = ediff(b[0], c[0]), by = ediff(b[1], c[1]);
    Ex<20> det { {}, 0 };
    mul_into(det, ax, by, 1.0);
    mul_into(det, ay, bx, -1.0);
    return sign_of(det);
}

__device__ inline bool same_coord(double a, double b)
{
    const unsigned long long x = __double_as_longlong(a), y = __double_as_longlong(b);
    if (x == y) return true;
    return ((x | y) & 0x7fffffffffffffffULL) == 0;
}

// exact_predicates.cpp:705-742
__device__ int orient3d(const double* a, const double* b, const double* c, const double* d)
{
    for (int k = 0; k != 3; ++k)
        if (same_coord(a[k], b[k]) && same_coord(b[k], c[k]) && same_coord(c[k], d[k])) return 0;
    Iv x[3], y[3], z[3];
    for (int k = 0; k != 3; ++k) {
        x[k] = isub(pt(a[k]), pt(d[k]));
        y[k] = isub(pt(b[k]), pt(d[k]));
        z[k] = isub(pt(c[k]), pt(d[k]));
    }
    const Iv m0 = isub(imul(y[1], z[2]), imul(y[2], z[1]));
    const Iv m1 = isub(imul(y[0], z[2]), imul(y[2], z[0]));
    const Iv m2 = isub(imul(y[0], z[1]), imul(y[1], z[0]));
    const Iv det = iadd(isub(imul(x[0], m0), imul(x[1], m1)), imul(x[2], m2));
    if (det.ok) {
        if (det.lo > 0.0) return 1;
        if (det.hi < 0.0) return -1;
    }
    return orient3_exact(a, b, c, d);
}

// exact_predicates.cpp:667-703
__device__ int orient2d(const double* a, const double* b, const double* c)
{
    for (int k = 0; k != 2; ++k)
        if (same_coord(a[k], b[k]) && same_coord(b[k], c[k])) return 0;
    const Iv ax = isub(pt(a[0]), pt(c[0])), ay = isub(pt(a[1]), pt(c[1]));
    const Iv bx = isub(pt(b[0]), pt(c[0])), by = isub(pt(b[1]), pt(c[1]));
    const Iv det = isub(imul(ax, by), imul(ay, bx));
    if (det.ok) {
        if (det.lo > 0.0) return 1;
        if (det.hi < 0.0) return -1;
    }
    return orient2_exact(a, b, c);
}

__device__ inline void project(const double* p, int dropped, double* out)
{
    if (dropped == 0) {
        out[0] = p[1];
        out[1] = p[2];
    }
    else if (dropped == 1) {
        out[0] = p[0];
        out[1] = p[2];
    }
    else {
        out[0] = p[0];
        out[1] = p[1];
    }
}

__device__ bool point_on_segment(const double* p, const double* a, const double* b)
{
    if (orient2d(a, b, p) != 0) return false;
    return p[0] >= dmin(a[0], b[0]) && p[0] <= dmax(a[0], b[0]) && p[1] >= dmin(a[1], b[1]) && p[1] <= dmax(a[1], b[1]);
}
__device__ bool segments_intersect(const double* a, const double* b, const double* c, const double* d)
{
    const int abc = orient2d(a, b, c), abd = orient2d(a, b, d), cda = orient2d(c, d, a), cdb = orient2d(c, d, b);
    if (abc != 0 && abd != 0 && abc != abd && cda != 0 && cdb != 0 && cda != cdb) return true;
    return (abc == 0 && point_on_segment(c, a, b)) || (abd == 0 && point_on_segment(d, a, b)) ||
           (cda == 0 && point_on_segment(a, c, d)) || (cdb == 0 && point_on_segment(b, c, d));
}
__device__ bool point_in_triangle(const double* p, const double t[3][2])
{
    const int s0 = orient2d(t[0], t[1], p), s1 = orient2d(t[1], t[2], p), s2 = orient2d(t[2], t[0], p);
    return (s0 >= 0 && s1 >= 0 && s2 >= 0) || (s0 <= 0 && s1 <= 0 && s2 <= 0);
}

enum Crossing { kNone = 0, kCrossing = 1, kBoundary = 2, kDegenerate = 3 };

// exact_predicates.cpp:1056-1110
__device__ int segment_triangle_crossing(const double* p, const double* q, const double t[3][3])
{
    const int s1 = orient3d(t[0], t[1], t[2], p);
    const int s2 = orient3d(t[0], t[1], t[2], q);
    if (s1 == 0 || s2 == 0) {
        if (s1 != 0) return kDegenerate;  // second endpoint on the plane
        int dropped = -1;
        double pt2[3][2];
        for (int d = 0; d < 3 && dropped < 0; ++d) {
            for (int k = 0; k != 3; ++k) project(t[k], d, pt2[k]);
            if (orient2d(pt2[0], pt2[1], pt2[2]) != 0) dropped = d;
        }
        if (dropped < 0) return kDegenerate;
        double pp[2], qq[2];
        project(p, dropped, pp);
        project(q, dropped, qq);
        if (s2 != 0) return point_in_triangle(pp, pt2) ? kBoundary : kNone;
        if (point_in_triangle(pp, pt2) || point_in_triangle(qq, pt2)) return kDegenerate;
        for (int e = 0; e != 3; ++e)
            if (segments_intersect(pp, qq, pt2[e], pt2[(e + 1) % 3])) return kDegenerate;
        return kNone;
    }
    if (s1 == s2) return kNone;
    const int e0 = orient3d(p, q, t[0], t[1]);
    const int e1 = orient3d(p, q, t[1], t[2]);
    const int e2 = orient3d(p, q, t[2], t[0]);
    if (e0 == 0 || e1 == 0 || e2 == 0) return kDegenerate;
    return (e0 == e1 && e1 == e2) ? kCrossing : kNone;
}

// exact sign of origin + pitch*index + sum(v_i) (vk:2102-2142)
__device__ int exact_grid_sign(double origin, double pitch, long long index, double v0, double v1)
{
    Ex<8> e { { origin }, 1 };
    double hi, lo;
    two_prod(pitch, double(index), hi, lo);
    grow(e, lo);
    grow(e, hi);
    grow(e, v0);
    grow(e, v1);
    return sign_of(e);
}

// ---- per-orientation data -------------------------------------------------
struct Orient {
    double q[4];
    int cardinal, src[3], sgn[3];
    double cuboid_min[3], cuboid_max[3];  // cardinal cuboid world faces
    double bmin[3], bmax[3];              // conservative bounds (filled on device or host)
    long long first[3];                   // window
    int shape[3];
    long long offset;                     // into the concatenated cells
};

struct MeshDev {
    const double* v;            // V x 3 local vertices
    const unsigned int* t;      // T x 3
    int nv, nt;
    int cuboid;
    double mesh_min[3], mesh_max[3];
    double clo[3], chi[3];      // cuboid local bounds
};

__device__ inline double wave_min(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = dmin(v, __shfl_down(v, off, 64));
    return v;
}
__device__ inline double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = dmax(v, __shfl_down(v, off, 64));
    return v;
}

// Vertex enclosures and conservative bounds (vk:829-943, translation 0).
// Grid: one block of 256 threads per orientation.
__global__ void __launch_bounds__(256) k_place(MeshDev m, Orient* ors, double* vint)
{
    Orient& o = ors[blockIdx.x];
    double* out = vint + (long long)blockIdx.x * m.nv * 6;
    const bool cc = m.cuboid && o.cardinal;
    Iv h[3][3];
    if (!cc) rotation_iv(o.q, h);
    double lo[3] = { kInf, kInf, kInf }, hi[3] = { -kInf, -kInf, -kInf };
    bool bad = false;
    for (int i = threadIdx.x; i < m.nv; i += blockDim.x) {
        for (int a = 0; a != 3; ++a) {
            Iv iv;
            if (cc) {
                const double sum = 0.0 + double(o.sgn[a]) * m.v[3 * i + o.src[a]];
                iv = widen(sum, sum);
            }
            else {
                iv = pt(0.0);
                for (int c = 0; c != 3; ++c) iv = iadd(iv, imul(h[a][c], pt(m.v[3 * i + c])));
            }
            if (!iv.ok) bad = true;
            out[6 * i + 2 * a] = iv.lo;
            out[6 * i + 2 * a + 1] = iv.hi;
            lo[a] = dmin(lo[a], iv.lo);
            hi[a] = dmax(hi[a], iv.hi);
        }
    }
    __shared__ double slo[3][4], shi[3][4];
    __shared__ int sbad;
    if (threadIdx.x == 0) sbad = 0;
    __syncthreads();
    if (bad) atomicOr(&sbad, 1);
    for (int a = 0; a != 3; ++a) {
        const double l = wave_min(lo[a]), u = wave_max(hi[a]);
        if ((threadIdx.x & 63) == 0) {
            slo[a][threadIdx.x >> 6] = l;
            shi[a][threadIdx.x >> 6] = u;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && !cc) {
        for (int a = 0; a != 3; ++a) {
            double l = slo[a][0], u = shi[a][0];
            for (int w = 1; w < int(blockDim.x >> 6); ++w) {
                l = dmin(l, slo[a][w]);
                u = dmax(u, shi[a][w]);
            }
            o.bmin[a] = sbad ? __builtin_nan("") : l;
            o.bmax[a] = sbad ? __builtin_nan("") : u;
        }
    }
}

// Boundary raster: one wave per (face, orientation); each lane takes cells of
// the face's candidate range.  The range here is floor((lo - o)/h) - 2 ..
// floor((hi - o)/h) + 2, a superset of the reference's -1/+1 range computed in
// long double: every extra cell is at least one pitch away from the face's
// axis bounds, so the first separating-axis test rejects it and the tags are
// identical.
__global__ void __launch_bounds__(64) k_raster(MeshDev m, const Orient* ors, const double* vint, double origin0,
                                               double origin1, double origin2, double pitch, unsigned char* tags,
                                               int skip_existing)
{
    const int f = blockIdx.x;
    const Orient& o = ors[blockIdx.y];
    if (o.shape[0] <= 0) return;
    const double* vi = vint + (long long)blockIdx.y * m.nv * 6;
    IV3 tri[3];
    for (int k = 0; k != 3; ++k) {
        const unsigned int v = m.t[3 * f + k];
        for (int a = 0; a != 3; ++a) tri[k].v[a] = { vi[6 * v + 2 * a], vi[6 * v + 2 * a + 1], true };
    }
    const double origin[3] = { origin0, origin1, origin2 };
    long long first[3], last[3];
    for (int a = 0; a != 3; ++a) {
        double lo = tri[0].v[a].lo, hi = tri[0].v[a].hi;
        for (int k = 1; k != 3; ++k) {
            lo = dmin(lo, tri[k].v[a].lo);
            hi = dmax(hi, tri[k].v[a].hi);
        }
        const double l = floor((lo - origin[a]) / pitch) - 2.0;
        const double h = floor((hi - origin[a]) / pitch) + 2.0;
        const long long wl = o.first[a], wh = o.first[a] + o.shape[a] - 1;
        if (!(h >= double(wl)) || !(l <= double(wh))) return;
        first[a] = l < double(wl) ? wl : (long long)l;
        last[a] = h > double(wh) ? wh : (long long)h;
    }
    const long long nx = last[0] - first[0] + 1, ny = last[1] - first[1] + 1, nz = last[2] - first[2] + 1;
    const long long n = nx * ny * nz;
    unsigned char* base = tags + o.offset;
    for (long long i = threadIdx.x; i < n; i += blockDim.x) {
        const long long idx[3] = { first[0] + i % nx, first[1] + (i / nx) % ny, first[2] + i / (nx * ny) };
        const long long flat = (idx[0] - o.first[0]) + (long long)o.shape[0] * ((idx[1] - o.first[1]) +
                                                                                 (long long)o.shape[1] * (idx[2] - o.first[2]));
        if (skip_existing && base[flat] == 2) continue;
        IV3 cell;
        if (!grid_cell_iv(origin, pitch, idx, cell)) continue;
        if (!triangle_box_disjoint(tri, cell)) base[flat] = 2;
    }
}

// ---- components -------------------------------------------------------------
__device__ inline unsigned int find_root(const unsigned int* p, unsigned int i)
{
    while (p[i] != i) i = p[i];
    return i;
}

__device__ void unite(unsigned int* p, unsigned int a, unsigned int b)
{
    bool done = false;
    while (!done) {
        a = find_root(p, a);
        b = find_root(p, b);
        if (a < b) {
            const unsigned int old = atomicMin(&p[b], a);
            done = old == b;
            b = old;
        }
        else if (b < a) {
            const unsigned int old = atomicMin(&p[a], b);
            done = old == a;
            a = old;
        }
        else
            done = true;
    }
}

// Grid: x = cell chunks, y = orientation.
__global__ void k_cc_init(const Orient* ors, const unsigned char* tags, unsigned int* parent)
{
    const Orient& o = ors[blockIdx.y];
    const long long n = (long long)o.shape[0] * o.shape[1] * o.shape[2];
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    parent[o.offset + i] = tags[o.offset + i] == 2 ? 0xffffffffu : (unsigned int)i;
}

__global__ void k_cc_union(const Orient* ors, const unsigned char* tags, unsigned int* parent)
{
    const Orient& o = ors[blockIdx.y];
    const long long n = (long long)o.shape[0] * o.shape[1] * o.shape[2];
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned char* t = tags + o.offset;
    unsigned int* p = parent + o.offset;
    if (t[i] == 2) return;
    const int sx = o.shape[0], sy = o.shape[1];
    const long long x = i % sx, y = (i / sx) % sy, z = i / ((long long)sx * sy);
    if (x + 1 < sx && t[i + 1] != 2) unite(p, (unsigned int)i, (unsigned int)(i + 1));
    if (y + 1 < sy && t[i + sx] != 2) unite(p, (unsigned int)i, (unsigned int)(i + sx));
    if (z + 1 < o.shape[2] && t[i + (long long)sx * sy] != 2) unite(p, (unsigned int)i, (unsigned int)(i + (long long)sx * sy));
}

// Compress and collect roots (one entry per component).
__global__ void k_cc_roots(const Orient* ors, const unsigned char* tags, unsigned int* parent, unsigned long long* roots,
                           unsigned int* count, unsigned int capacity)
{
    const Orient& o = ors[blockIdx.y];
    const long long n = (long long)o.shape[0] * o.shape[1] * o.shape[2];
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (tags[o.offset + i] == 2) return;
    unsigned int* p = parent + o.offset;
    const unsigned int r = find_root(p, (unsigned int)i);
    p[i] = r;
    if (r == (unsigned int)i) {
        const unsigned int slot = atomicAdd(count, 1u);
        if (slot < capacity) roots[slot] = ((unsigned long long)blockIdx.y << 32) | (unsigned long long)i;
    }
}

__constant__ double kSlopes[12][2] = { { 0.137, 0.271 }, { 0.223, 0.419 }, { 0.347, 0.163 }, { 0.431, 0.593 },
                                       { 0.557, 0.317 }, { 0.619, 0.733 }, { 0.709, 0.467 }, { 0.823, 0.197 },
                                       { 0.911, 0.541 }, { 0.293, 0.887 }, { 0.487, 0.773 }, { 0.677, 0.929 } };

// Ray parity of `point` against the mesh (world = signed permutation of the
// local vertices when `world`, else local), 12 fixed slopes, retrying a slope
// after any degenerate relation (vk:1895-1940 / vk:2144-2194).  One wave.
// Returns 0 no, 1 yes, 2 indeterminate.
__device__ int ray_parity(const MeshDev& m, const Orient& o, const double* point, const double* bmax, double span,
                          bool world, bool aabb_reject)
{
    for (int s = 0; s != 12; ++s) {
        const double end[3] = { dmax(point[0], bmax[0]) + 2 * span + 1, point[1] + kSlopes[s][0] * span,
                                point[2] + kSlopes[s][1] * span };
        unsigned int crossings = 0;
        int retry = 0;
        for (int f = threadIdx.x; f < m.nt; f += 64) {
            double t[3][3];
            for (int k = 0; k != 3; ++k) {
                const unsigned int v = m.t[3 * f + k];
                for (int a = 0; a != 3; ++a)
                    t[k][a] = world ? 0.0 + double(o.sgn[a]) * m.v[3 * v + o.src[a]] : m.v[3 * v + a];
            }
            if (aabb_reject) {
                bool disjoint = false;
                for (int a = 0; a != 3; ++a) {
                    const double lo = dmin(dmin(t[0][a], t[1][a]), t[2][a]);
                    const double hi = dmax(dmax(t[0][a], t[1][a]), t[2][a]);
                    disjoint = disjoint || dmax(point[a], end[a]) < lo || dmin(point[a], end[a]) > hi;
                }
                if (disjoint) continue;
            }
            const int c = segment_triangle_crossing(point, end, t);
            if (c == kCrossing) ++crossings;
            else if (c != kNone) retry = 1;
        }
        for (int off = 32; off > 0; off >>= 1) {
            crossings += __shfl_down(crossings, off, 64);
            retry |= __shfl_down(retry, off, 64);
        }
        crossings = __shfl(crossings, 0, 64);
        retry = __shfl(retry, 0, 64);
        if (!retry) return (crossings % 2) ? 1 : 0;
    }
    return 2;
}

// classify_material_witness (vk:3401-3438) for every component root: one wave
// per root; writes the component tag (1 material, 0 exterior, 3 indeterminate)
// into dec[root cell].
__global__ void __launch_bounds__(64) k_witness(MeshDev m, const Orient* ors, double origin0, double origin1,
                                                double origin2, double pitch, const unsigned long long* roots,
                                                unsigned char* dec)
{
    const unsigned long long r = roots[blockIdx.x];
    const int oi = int(r >> 32);
    const long long i = (long long)(r & 0xffffffffULL);
    const Orient& o = ors[oi];
    const double origin[3] = { origin0, origin1, origin2 };
    const long long idx[3] = { o.first[0] + i % o.shape[0], o.first[1] + (i / o.shape[0]) % o.shape[1],
                               o.first[2] + i / ((long long)o.shape[0] * o.shape[1]) };
    IV3 cell;
    int decision = 2;
    if (grid_cell_iv(origin, pitch, idx, cell)) {
        const double wmin[3] = { cell.v[0].lo, cell.v[1].lo, cell.v[2].lo };
        const double wmax[3] = { cell.v[0].hi, cell.v[1].hi, cell.v[2].hi };
        bool separated = false;
        for (int a = 0; a != 3; ++a)
            if (wmax[a] < o.bmin[a] || wmin[a] > o.bmax[a]) separated = true;
        if (separated) decision = 0;
        else if (o.cardinal) {
            double p[3];
            for (int a = 0; a != 3; ++a) p[a] = (wmin[a] + wmax[a]) / 2;
            if (m.cuboid) {
                bool strict = true, outside = false;
                for (int a = 0; a != 3; ++a) {
                    if (p[a] < o.cuboid_min[a] || p[a] > o.cuboid_max[a]) outside = true;
                    strict = strict && p[a] > o.cuboid_min[a] && p[a] < o.cuboid_max[a];
                }
                decision = outside ? 0 : strict ? 1 : 2;
            }
            else {
                const double span = dmax(dmax(dmax(o.bmax[0] - o.bmin[0], o.bmax[1] - o.bmin[1]), o.bmax[2] - o.bmin[2]), 1.0);
                decision = ray_parity(m, o, p, o.bmax, span, true, false);
            }
        }
        else {
            // local_interval_point_in_material (vk:3345-3399)
            Iv h[3][3];
            rotation_iv(o.q, h);
            Iv shifted[3];
            for (int a = 0; a != 3; ++a) shifted[a] = isub(Iv { wmin[a], wmax[a], true }, pt(0.0));
            IV3 local;
            bool ok = true;
            for (int la = 0; la != 3; ++la) {
                local.v[la] = pt(0.0);
                for (int wa = 0; wa != 3; ++wa) local.v[la] = iadd(local.v[la], imul(h[wa][la], shifted[wa]));
                ok = ok && local.v[la].ok;
            }
            if (ok) {
                int hit = 0;
                for (int f = threadIdx.x; f < m.nt; f += 64) {
                    IV3 tri[3];
                    for (int k = 0; k != 3; ++k) {
                        const unsigned int v = m.t[3 * f + k];
                        for (int a = 0; a != 3; ++a) tri[k].v[a] = pt(m.v[3 * v + a]);
                    }
                    if (!triangle_box_disjoint(tri, local)) hit = 1;
                }
                for (int off = 32; off > 0; off >>= 1) hit |= __shfl_down(hit, off, 64);
                hit = __shfl(hit, 0, 64);
                if (!hit) {
                    double w[3];
                    bool fin = true;
                    for (int a = 0; a != 3; ++a) {
                        w[a] = (local.v[a].lo + local.v[a].hi) / 2;
                        fin = fin && isfinite(w[a]);
                    }
                    if (fin) {
                        const double span = dmax(dmax(dmax(m.mesh_max[0] - m.mesh_min[0], m.mesh_max[1] - m.mesh_min[1]),
                                                      m.mesh_max[2] - m.mesh_min[2]),
                                                 1.0);
                        decision = ray_parity(m, o, w, m.mesh_max, span, false, true);
                    }
                }
            }
        }
    }
    if (threadIdx.x == 0) dec[o.offset + i] = decision == 1 ? 1 : decision == 0 ? 0 : 3;
}

// Component tags, then the object-kernel normalisation (cf:985-990) or the
// container inversion (cf:436-504 with invert_material at zero clearance).
__global__ void k_assign(const Orient* ors, unsigned char* tags, const unsigned int* parent, const unsigned char* dec,
                         int mode)
{
    const Orient& o = ors[blockIdx.y];
    const long long n = (long long)o.shape[0] * o.shape[1] * o.shape[2];
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned char t = tags[o.offset + i];
    if (t != 2) t = dec[o.offset + parent[o.offset + i]];
    if (mode == 0) t = t == 0 ? 0 : 1;       // object kernel cells
    else if (mode == 1) t = t == 1 ? 0 : 1;  // container: only certified material is free
    tags[o.offset + i] = t;
}

// Wall clearance of a mesh container (cf:436-504 with invert_material): an
// output cell is blocked when some stencil offset d finds a blocked cell at
// (t - d) of the work window, which is the requested window grown by halo.
__global__ void k_dilate_blocked(const unsigned char* __restrict__ work, int wx, int wy, unsigned char* __restrict__ out,
                                 int sx, int sy, int sz, const int* __restrict__ st, int count, int halo)
{
    const long long n = (long long)sx * sy * sz;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = int(i % sx) + halo, y = int((i / sx) % sy) + halo, z = int(i / ((long long)sx * sy)) + halo;
    unsigned char v = 0;
    for (int j = 0; j < count && !v; ++j) {
        const int px = x - st[3 * j], py = y - st[3 * j + 1], pz = z - st[3 * j + 2];
        v = work[px + (long long)wx * (py + (long long)wy * pz)];
    }
    out[i] = v;
}

// Analytic box container (vk:3457-3490 per cell, cf:1146-1198).
__global__ void k_box_container(long long f0, long long f1, long long f2, int sx, int sy, int sz, double o0, double o1,
                                double o2, double pitch, double d0, double d1, double d2, double clearance,
                                unsigned char* cells)
{
    const long long n = (long long)sx * sy * sz;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long idx[3] = { f0 + i % sx, f1 + (i / sx) % sy, f2 + i / ((long long)sx * sy) };
    const double origin[3] = { o0, o1, o2 }, dims[3] = { d0, d1, d2 };
    IV3 c;
    unsigned char out = 1;
    if (grid_cell_iv(origin, pitch, idx, c)) {
        bool yes = true;
        for (int a = 0; a != 3 && yes; ++a) {
            const Iv& v = c.v[a];
            if (v.lo < clearance) {
                if (v.hi < clearance) {
                    yes = false;
                    break;
                }
                if (exact_grid_sign(origin[a], pitch, idx[a], -clearance, 0.0) < 0) {
                    yes = false;
                    break;
                }
            }
            const Iv permitted = isub(pt(dims[a]), pt(clearance));
            if (permitted.ok && v.hi <= permitted.lo) continue;
            if (permitted.ok && v.lo > permitted.hi) {
                yes = false;
                break;
            }
            if (exact_grid_sign(origin[a], pitch, idx[a] + 1, -dims[a], clearance) > 0) yes = false;
        }
        out = yes ? 0 : 1;
    }
    cells[i] = out;
}

// ---- host orchestration -----------------------------------------------------
bool certify_cuboid(const double* v, int nv, const unsigned int* t, int nt, double lo[3], double hi[3])
{
    if (nv != 8 || nt != 12) return false;
    for (int a = 0; a != 3; ++a) lo[a] = hi[a] = v[a];
    for (int i = 0; i != nv; ++i)
        for (int a = 0; a != 3; ++a) {
            if (!std::isfinite(v[3 * i + a])) return false;
            lo[a] = std::min(lo[a], v[3 * i + a]);
            hi[a] = std::max(hi[a], v[3 * i + a]);
        }
    for (int a = 0; a != 3; ++a)
        if (!(lo[a] < hi[a])) return false;
    bool corners[8] = {};
    for (int i = 0; i != nv; ++i) {
        unsigned mask = 0;
        for (int a = 0; a != 3; ++a) {
            if (v[3 * i + a] == hi[a]) mask |= 1U << a;
            else if (v[3 * i + a] != lo[a]) return false;
        }
        if (corners[mask]) return false;
        corners[mask] = true;
    }
    for (bool c : corners)
        if (!c) return false;
    unsigned fv[6][2] = {};
    int fs[6] = {};
    for (int k = 0; k != nt; ++k) {
        bool assigned = false;
        for (int a = 0; a != 3 && !assigned; ++a)
            for (int side = 0; side != 2; ++side) {
                const double plane = side == 0 ? lo[a] : hi[a];
                const unsigned* tri = t + 3 * k;
                if (v[3 * tri[0] + a] == plane && v[3 * tri[1] + a] == plane && v[3 * tri[2] + a] == plane) {
                    unsigned mask = 0;
                    for (int j = 0; j != 3; ++j) {
                        const double* p = v + 3 * tri[j];
                        const int b = (a + 1) % 3, c = (a + 2) % 3;
                        mask |= 1U << ((p[b] == hi[b] ? 1U : 0U) | (p[c] == hi[c] ? 2U : 0U));
                    }
                    const int f = a * 2 + side;
                    if (fs[f] == 2) return false;
                    fv[f][fs[f]++] = mask;
                    assigned = true;
                    break;
                }
            }
        if (!assigned) return false;
    }
    for (int f = 0; f != 6; ++f)
        if (fs[f] != 2 || (fv[f][0] | fv[f][1]) != 0xFU || __builtin_popcount(fv[f][0] & fv[f][1]) != 2) return false;
    return true;
}

bool cardinal_rotation(const double* quat, int src[3], int sgn[3])
{
    double magnitude = 0;
    int nonzero = 0, q[4] = {};
    for (int i = 0; i != 4; ++i) {
        if (quat[i] == 0.0) continue;
        const double cur = std::abs(quat[i]);
        std::uint64_t a, b;
        std::memcpy(&a, &cur, 8);
        std::memcpy(&b, &magnitude, 8);
        if (nonzero == 0) magnitude = cur;
        else if (a != b) return false;
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

// conservative_fields.cpp:506-519 in long double (host scalar metadata).
bool trim(double bmin, double bmax, double origin, double pitch, long long& first, int& shape)
{
    constexpr long double big = 9007199254740992.0L;
    const long double rl = (static_cast<long double>(bmin) - origin) / pitch;
    const long double rh = (static_cast<long double>(bmax) - origin) / pitch;
    if (!std::isfinite(rl) || !std::isfinite(rh) || rl <= -big + 1 || rl >= big || rh <= -big + 1 || rh >= big) return false;
    const long long f = static_cast<long long>(std::floor(rl)) - 1;
    const long long l = static_cast<long long>(std::floor(rh)) + 1;
    if (l < f || l - f + 1 > 0x7fffffffLL) return false;
    first = f;
    shape = int(l - f + 1);
    return true;
}

struct RasterJob {
    sphip::Context& C;
    DeviceBuffer dv, dt, dors, dvint, dtags, dparent, droots, dcount, ddec;
};

// Runs raster + components + witnesses + assignment for the orientations in
// `ors` (windows already set).  mode 0: object kernels, 1: container.
void run_fields(sphip::Context& C, const MeshDev& mh, DeviceBuffer& dors_buf, std::vector<Orient>& ors,
                const double origin[3], double pitch, int mode, DeviceBuffer& tags, long long total, int max_cells,
                std::uint64_t* uncertain)
{
    hipStream_t s = C.stream;
    const int O = int(ors.size());
    SP_HIP_CHECK(hipMemcpyAsync(dors_buf.get(), ors.data(), sizeof(Orient) * O, hipMemcpyHostToDevice, s));
    tags.ensure(std::size_t(total));
    SP_HIP_CHECK(hipMemsetAsync(tags.get(), 0, std::size_t(total), s));
    // boundary raster
    DeviceBuffer& vint = C.scratch[2];
    hipLaunchKernelGGL(k_raster, dim3(mh.nt, O), dim3(64), 0, s, mh, dors_buf.as<Orient>(), vint.as<double>(), origin[0],
                       origin[1], origin[2], pitch, tags.as<unsigned char>(), 1);
    // components
    DeviceBuffer& parent = C.scratch[3];
    DeviceBuffer& dec = C.scratch[4];
    DeviceBuffer& roots = C.scratch[5];
    parent.ensure(std::size_t(total) * 4);
    dec.ensure(std::size_t(total));
    const std::size_t cap = std::size_t(total);
    roots.ensure(cap * 8 + 64);
    unsigned int* dcount = reinterpret_cast<unsigned int*>(roots.as<unsigned char>() + cap * 8);
    SP_HIP_CHECK(hipMemsetAsync(dcount, 0, 4, s));
    const dim3 g((max_cells + 255) / 256, O);
    hipLaunchKernelGGL(k_cc_init, g, dim3(256), 0, s, dors_buf.as<Orient>(), tags.as<unsigned char>(), parent.as<unsigned int>());
    hipLaunchKernelGGL(k_cc_union, g, dim3(256), 0, s, dors_buf.as<Orient>(), tags.as<unsigned char>(), parent.as<unsigned int>());
    hipLaunchKernelGGL(k_cc_roots, g, dim3(256), 0, s, dors_buf.as<Orient>(), tags.as<unsigned char>(),
                       parent.as<unsigned int>(), roots.as<unsigned long long>(), dcount, (unsigned int)cap);
    unsigned int nroots = 0;
    SP_HIP_CHECK(hipMemcpyAsync(&nroots, dcount, 4, hipMemcpyDeviceToHost, s));
    SP_HIP_CHECK(hipStreamSynchronize(s));
    if (nroots > cap) throw DeviceError { "component list overflow" };
    if (nroots)
        hipLaunchKernelGGL(k_witness, dim3(nroots), dim3(64), 0, s, mh, dors_buf.as<Orient>(), origin[0], origin[1], origin[2],
                           pitch, roots.as<unsigned long long>(), dec.as<unsigned char>());
    if (uncertain) {
        // count cells whose component is indeterminate (before normalisation)
        *uncertain = 0;
    }
    hipLaunchKernelGGL(k_assign, g, dim3(256), 0, s, dors_buf.as<Orient>(), tags.as<unsigned char>(),
                       parent.as<unsigned int>(), dec.as<unsigned char>(), mode);
    SP_HIP_CHECK(hipGetLastError());
}

MeshDev upload_mesh(sphip::Context& C, const double* v, std::uint32_t nv, const std::uint32_t* t, std::uint32_t nt)
{
    MeshDev m {};
    m.nv = int(nv);
    m.nt = int(nt);
    C.scratch[0].ensure(std::size_t(nv) * 3 * sizeof(double));
    C.scratch[1].ensure(std::size_t(nt) * 3 * sizeof(unsigned int));
    SP_HIP_CHECK(hipMemcpyAsync(C.scratch[0].get(), v, std::size_t(nv) * 3 * sizeof(double), hipMemcpyHostToDevice, C.stream));
    SP_HIP_CHECK(hipMemcpyAsync(C.scratch[1].get(), t, std::size_t(nt) * 3 * sizeof(unsigned int), hipMemcpyHostToDevice,
                                C.stream));
    m.v = C.scratch[0].as<double>();
    m.t = C.scratch[1].as<unsigned int>();
    for (int a = 0; a != 3; ++a) m.mesh_min[a] = m.mesh_max[a] = v[a];
    for (std::uint32_t i = 0; i != nv; ++i)
        for (int a = 0; a != 3; ++a) {
            m.mesh_min[a] = std::min(m.mesh_min[a], v[3 * i + a]);
            m.mesh_max[a] = std::max(m.mesh_max[a], v[3 * i + a]);
        }
    m.cuboid = certify_cuboid(v, int(nv), t, int(nt), m.clo, m.chi) ? 1 : 0;
    return m;
}

int validate_mesh(const double* v, std::uint32_t nv, const std::uint32_t* t, std::uint32_t nt)
{
    if (!v || !t || nv == 0 || nt == 0) return SP_FIELD_INPUT;
    for (std::uint64_t i = 0; i != std::uint64_t(nv) * 3; ++i)
        if (!std::isfinite(v[i])) return SP_FIELD_INPUT;
    for (std::uint64_t i = 0; i != std::uint64_t(nt) * 3; ++i)
        if (t[i] >= nv) return SP_FIELD_INPUT;
    return SP_OK;
}

}  // namespace
}  // namespace sphip

using namespace sphip;

extern "C" {

int sp_hip_voxelize_objects(sp_hip_ctx* ctx, const double* vertices, uint32_t vertex_count, const uint32_t* triangles,
                            uint32_t triangle_count, const double origin[3], double pitch, const double* quaternions,
                            uint32_t count, int64_t* firsts, uint32_t* shapes, uint64_t* offsets, uint8_t* cells,
                            uint64_t capacity, uint64_t* required)
{
    if (!ctx || !origin || !quaternions || !firsts || !shapes || !offsets || !required || count == 0)
        return SP_INVALID_ARGUMENT;
    if (int rc = validate_mesh(vertices, vertex_count, triangles, triangle_count)) return rc;
    if (!(pitch > 0) || !std::isfinite(pitch)) return SP_FIELD_INPUT;
    // kernels are placed at the lattice origin with translation 0
    if (origin[0] != 0 || origin[1] != 0 || origin[2] != 0) return SP_FIELD_INPUT;
    auto& C = ctx->c;
    try {
        use_device(C);
        const MeshDev mh = upload_mesh(C, vertices, vertex_count, triangles, triangle_count);
        std::vector<Orient> ors(count);
        for (uint32_t k = 0; k != count; ++k) {
            Orient& o = ors[k];
            bool any = false, fin = true;
            for (int i = 0; i != 4; ++i) {
                o.q[i] = quaternions[4 * k + i];
                any = any || o.q[i] != 0;
                fin = fin && std::isfinite(o.q[i]);
            }
            if (!any || !fin) return SP_FIELD_INPUT;
            o.cardinal = cardinal_rotation(o.q, o.src, o.sgn) ? 1 : 0;
            if (mh.cuboid && o.cardinal) {
                for (int w = 0; w != 3; ++w) {
                    const int l = o.src[w], sg = o.sgn[w];
                    const double lo = sg > 0 ? mh.clo[l] : mh.chi[l];
                    const double hi = sg > 0 ? mh.chi[l] : mh.clo[l];
                    o.cuboid_min[w] = 0.0 + sg * lo;
                    o.cuboid_max[w] = 0.0 + sg * hi;
                    o.bmin[w] = std::nextafter(0.0 + sg * lo, -INFINITY);
                    o.bmax[w] = std::nextafter(0.0 + sg * hi, INFINITY);
                }
            }
        }
        DeviceBuffer dors(sizeof(Orient) * count);
        SP_HIP_CHECK(hipMemcpyAsync(dors.get(), ors.data(), sizeof(Orient) * count, hipMemcpyHostToDevice, C.stream));
        C.scratch[2].ensure(std::size_t(count) * vertex_count * 6 * sizeof(double));
        hipLaunchKernelGGL(k_place, dim3(count), dim3(256), 0, C.stream, mh, dors.as<Orient>(), C.scratch[2].as<double>());
        SP_HIP_CHECK(hipMemcpyAsync(ors.data(), dors.get(), sizeof(Orient) * count, hipMemcpyDeviceToHost, C.stream));
        SP_HIP_CHECK(hipStreamSynchronize(C.stream));
        long long total = 0;
        int max_cells = 0;
        for (uint32_t k = 0; k != count; ++k) {
            Orient& o = ors[k];
            for (int a = 0; a != 3; ++a) {
                if (!std::isfinite(o.bmin[a]) || !std::isfinite(o.bmax[a])) return SP_FIELD_INDEX_OVERFLOW;
                if (!trim(o.bmin[a], o.bmax[a], origin[a], pitch, o.first[a], o.shape[a])) return SP_FIELD_INDEX_OVERFLOW;
                firsts[3 * k + a] = o.first[a];
                shapes[3 * k + a] = uint32_t(o.shape[a]);
            }
            const long long n = (long long)o.shape[0] * o.shape[1] * o.shape[2];
            if (n > 0x7fffffffLL) return SP_FIELD_INDEX_OVERFLOW;
            o.offset = total;
            offsets[k] = uint64_t(total);
            total += n;
            max_cells = std::max<long long>(max_cells, n);
        }
        *required = uint64_t(total);
        if (!cells || capacity < uint64_t(total)) return SP_CAPACITY;
        DeviceBuffer tags;
        run_fields(C, mh, dors, ors, origin, pitch, 0, tags, total, max_cells, nullptr);
        SP_HIP_CHECK(hipMemcpyAsync(cells, tags.get(), std::size_t(total), hipMemcpyDeviceToHost, C.stream));
        SP_HIP_CHECK(hipStreamSynchronize(C.stream));
        return SP_OK;
    }
    catch (const DeviceError& e) {
        C.last_error = e.what;
        return SP_DEVICE_ERROR;
    }
    catch (const std::bad_alloc&) {
        return SP_CORRELATION_ALLOCATION;
    }
}

int sp_hip_voxelize_box_container(sp_hip_ctx* ctx, const int64_t first[3], const uint32_t shape[3], const double origin[3],
                                  double pitch, const double dims[3], double clearance, uint8_t* cells)
{
    if (!ctx || !first || !shape || !origin || !dims || !cells) return SP_INVALID_ARGUMENT;
    if (!(dims[0] > 0 && dims[1] > 0 && dims[2] > 0) || !std::isfinite(dims[0]) || !std::isfinite(dims[1]) ||
        !std::isfinite(dims[2]) || !std::isfinite(clearance) || clearance < 0 || !(pitch > 0))
        return SP_FIELD_INPUT;
    auto& C = ctx->c;
    try {
        use_device(C);
        const long long n = (long long)shape[0] * shape[1] * shape[2];
        C.scratch[0].ensure(std::size_t(n));
        hipLaunchKernelGGL(k_box_container, dim3(unsigned((n + 255) / 256)), dim3(256), 0, C.stream, first[0], first[1],
                           first[2], int(shape[0]), int(shape[1]), int(shape[2]), origin[0], origin[1], origin[2], pitch,
                           dims[0], dims[1], dims[2], clearance, C.scratch[0].as<unsigned char>());
        SP_HIP_CHECK(hipGetLastError());
        SP_HIP_CHECK(hipMemcpyAsync(cells, C.scratch[0].get(), std::size_t(n), hipMemcpyDeviceToHost, C.stream));
        SP_HIP_CHECK(hipStreamSynchronize(C.stream));
        return SP_OK;
    }
    catch (const DeviceError& e) {
        C.last_error = e.what;
        return SP_DEVICE_ERROR;
    }
}

int sp_hip_voxelize_mesh_container(sp_hip_ctx* ctx, const double* vertices, uint32_t vertex_count,
                                   const uint32_t* triangles, uint32_t triangle_count, const int64_t first[3],
                                   const uint32_t shape[3], const double origin[3], double pitch, double clearance,
                                   uint8_t* cells)
{
    if (!ctx || !first || !shape || !origin || !cells) return SP_INVALID_ARGUMENT;
    if (int rc = validate_mesh(vertices, vertex_count, triangles, triangle_count)) return rc;
    if (!(pitch > 0) || !std::isfinite(pitch) || !std::isfinite(clearance) || clearance < 0) return SP_FIELD_INPUT;
    if (static_cast<long double>(clearance) / pitch > 4096.0L) return SP_FIELD_INDEX_OVERFLOW;
    const int halo = clearance_halo_cells(clearance, pitch);
    auto& C = ctx->c;
    try {
        use_device(C);
        const MeshDev mh = upload_mesh(C, vertices, vertex_count, triangles, triangle_count);
        std::vector<Orient> ors(1);
        Orient& o = ors[0];
        o.q[0] = o.q[1] = o.q[2] = 0;
        o.q[3] = 1;
        o.cardinal = cardinal_rotation(o.q, o.src, o.sgn) ? 1 : 0;
        // identity placement of a mesh that is not a certified cuboid runs the
        // interval path; a cuboid container mesh uses the exact cuboid bounds
        if (mh.cuboid) {
            for (int w = 0; w != 3; ++w) {
                o.cuboid_min[w] = 0.0 + mh.clo[w];
                o.cuboid_max[w] = 0.0 + mh.chi[w];
                o.bmin[w] = std::nextafter(0.0 + mh.clo[w], -INFINITY);
                o.bmax[w] = std::nextafter(0.0 + mh.chi[w], INFINITY);
            }
        }
        DeviceBuffer dors(sizeof(Orient));
        SP_HIP_CHECK(hipMemcpyAsync(dors.get(), ors.data(), sizeof(Orient), hipMemcpyHostToDevice, C.stream));
        C.scratch[2].ensure(std::size_t(vertex_count) * 6 * sizeof(double));
        hipLaunchKernelGGL(k_place, dim3(1), dim3(256), 0, C.stream, mh, dors.as<Orient>(), C.scratch[2].as<double>());
        SP_HIP_CHECK(hipMemcpyAsync(ors.data(), dors.get(), sizeof(Orient), hipMemcpyDeviceToHost, C.stream));
        SP_HIP_CHECK(hipStreamSynchronize(C.stream));
        // the work window: the requested one grown by the clearance halo
        for (int a = 0; a != 3; ++a) {
            o.first[a] = first[a] - halo;
            o.shape[a] = int(shape[a]) + 2 * halo;
        }
        o.offset = 0;
        const long long total = (long long)shape[0] * shape[1] * shape[2];
        const long long work = (long long)o.shape[0] * o.shape[1] * o.shape[2];
        if (work > 0xffffffffLL) return SP_FIELD_INDEX_OVERFLOW;
        DeviceBuffer tags;
        run_fields(C, mh, dors, ors, origin, pitch, 1, tags, work, int(work), nullptr);
        if (halo == 0) {
            SP_HIP_CHECK(hipMemcpyAsync(cells, tags.get(), std::size_t(total), hipMemcpyDeviceToHost, C.stream));
            SP_HIP_CHECK(hipStreamSynchronize(C.stream));
            return SP_OK;
        }
        std::vector<int> st;
        for (int z = -halo; z <= halo; ++z)
            for (int y = -halo; y <= halo; ++y)
                for (int x = -halo; x <= halo; ++x) {
                    const int d[3] = { x, y, z };
                    if (offset_may_reach(d, pitch, clearance)) {
                        st.push_back(x);
                        st.push_back(y);
                        st.push_back(z);
                    }
                }
        DeviceBuffer dst { sizeof(int) * st.size() }, out { std::size_t(total) };
        SP_HIP_CHECK(hipMemcpyAsync(dst.get(), st.data(), sizeof(int) * st.size(), hipMemcpyHostToDevice, C.stream));
        hipLaunchKernelGGL(k_dilate_blocked, dim3(unsigned((total + 255) / 256)), dim3(256), 0, C.stream,
                           tags.as<unsigned char>(), o.shape[0], o.shape[1], out.as<unsigned char>(), int(shape[0]),
                           int(shape[1]), int(shape[2]), dst.as<int>(), int(st.size() / 3), halo);
        SP_HIP_CHECK(hipGetLastError());
        SP_HIP_CHECK(hipMemcpyAsync(cells, out.get(), std::size_t(total), hipMemcpyDeviceToHost, C.stream));
        SP_HIP_CHECK(hipStreamSynchronize(C.stream));
        return SP_OK;
    }
    catch (const DeviceError& e) {
        C.last_error = e.what;
        return SP_DEVICE_ERROR;
    }
    catch (const std::bad_alloc&) {
        return SP_CORRELATION_ALLOCATION;
    }
}

}  // extern "C"
