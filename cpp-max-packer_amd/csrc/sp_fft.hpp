// Device-side FFT building blocks for gfx950: batched Stockham passes over a
// tile of pencils held in LDS.  A tile is W pencils of N complex points, stored
// [n][w] with a row pitch of SW elements so that consecutive lanes touch
// consecutive LDS words.  Passes ping-pong between two LDS buffers (one barrier
// per pass).  Radices 2, 3, 4, 5 and 8 cover every length 2^a 3^b 5^c.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sphip {

template <typename T> struct Cx;
template <> struct Cx<float> {
    using type = float2;
};
template <> struct Cx<double> {
    using type = double2;
};
template <typename T> using cx = typename Cx<T>::type;

// Streaming accesses: data read or written once per pass, not reused by the
// same kernel, bypass cache allocation.
template <typename T2> __device__ __forceinline__ T2 nt_load(const T2* p)
{
    using E = decltype(p->x);
    using V = E __attribute__((ext_vector_type(2)));
    const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
    return T2 { v.x, v.y };
}

template <typename T2> __device__ __forceinline__ void nt_store(T2* p, T2 v)
{
    using E = decltype(v.x);
    using V = E __attribute__((ext_vector_type(2)));
    const V w = { v.x, v.y };
    __builtin_nontemporal_store(w, reinterpret_cast<V*>(p));
}


template <typename T2> __device__ __forceinline__ T2 cmul(T2 a, T2 b)
{
    return { a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x };
}
// a * conj(b)
template <typename T2> __device__ __forceinline__ T2 cmulc(T2 a, T2 b)
{
    return { a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y };
}
template <typename T2> __device__ __forceinline__ T2 cadd(T2 a, T2 b) { return { a.x + b.x, a.y + b.y }; }
template <typename T2> __device__ __forceinline__ T2 csub(T2 a, T2 b) { return { a.x - b.x, a.y - b.y }; }
// multiply by -i (forward) or +i (inverse)
template <typename T2> __device__ __forceinline__ T2 mul_mi(T2 a, bool forward)
{
    return forward ? T2 { a.y, -a.x } : T2 { -a.y, a.x };
}
template <typename T2> __device__ __forceinline__ T2 cconj(T2 a) { return { a.x, -a.y }; }

// Maximum factors per length (2^11 needs 4 radix-8 + ... ; 12 is ample).
constexpr int kMaxFactors = 12;

// A length plan: radices in pass order, product N.  Lives in kernel arguments.
struct FftPlan {
    int n;
    int count;
    int radix[kMaxFactors];
};

// In-register DFT of size R (R in {2,3,4,5,8}); v in, v out.
template <typename T, int R> struct Dft;

template <typename T> struct Dft<T, 2> {
    using T2 = cx<T>;
    __device__ __forceinline__ static void run(T2* v, bool)
    {
        const T2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <typename T> struct Dft<T, 4> {
    using T2 = cx<T>;
    __device__ __forceinline__ static void run(T2* v, bool fwd)
    {
        const T2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        const T2 c = cadd(v[1], v[3]);
        const T2 d = mul_mi(csub(v[1], v[3]), fwd);
        v[0] = cadd(a, c);
        v[1] = cadd(b, d);
        v[2] = csub(a, c);
        v[3] = csub(b, d);
    }
};

template <typename T> struct Dft<T, 8> {
    using T2 = cx<T>;
    __device__ __forceinline__ static void run(T2* v, bool fwd)
    {
        // radix-2 over two radix-4 DFTs of the even and odd samples
        T2 e[4] = { v[0], v[2], v[4], v[6] };
        T2 o[4] = { v[1], v[3], v[5], v[7] };
        Dft<T, 4>::run(e, fwd);
        Dft<T, 4>::run(o, fwd);
        const T r = T(0.70710678118654752440084436210484904);
        // o[k] *= w8^k, w8 = exp(-+ i pi/4)
        const T2 o1 = fwd ? T2 { (o[1].x + o[1].y) * r, (o[1].y - o[1].x) * r }
                          : T2 { (o[1].x - o[1].y) * r, (o[1].y + o[1].x) * r };
        const T2 o2 = mul_mi(o[2], fwd);
        const T2 o3 = fwd ? T2 { (o[3].y - o[3].x) * r, -(o[3].x + o[3].y) * r }
                          : T2 { -(o[3].x + o[3].y) * r, (o[3].x - o[3].y) * r };
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

template <typename T> struct Dft<T, 3> {
    using T2 = cx<T>;
    __device__ __forceinline__ static void run(T2* v, bool fwd)
    {
        const T c = T(-0.5);
        const T s = fwd ? T(-0.86602540378443864676372317075293618) : T(0.86602540378443864676372317075293618);
        const T2 t1 = cadd(v[1], v[2]);
        const T2 t2 = csub(v[1], v[2]);
        const T2 m = { v[0].x + c * t1.x, v[0].y + c * t1.y };
        const T2 j = { -s * t2.y, s * t2.x };  // i*s*t2
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m, j);
        v[2] = csub(m, j);
    }
};

template <typename T> struct Dft<T, 5> {
    using T2 = cx<T>;
    __device__ __forceinline__ static void run(T2* v, bool fwd)
    {
        const T c1 = T(0.30901699437494742410229341718281906);   // cos(2pi/5)
        const T c2 = T(-0.80901699437494742410229341718281906);  // cos(4pi/5)
        const T s1 = fwd ? T(-0.95105651629515357211643933337938214) : T(0.95105651629515357211643933337938214);
        const T s2 = fwd ? T(-0.58778525229247312916870595463907277) : T(0.58778525229247312916870595463907277);
        const T2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        const T2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        const T2 m1 = { v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y };
        const T2 m2 = { v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y };
        // i*(s1*b1 + s2*b2) and i*(s2*b1 - s1*b2)
        const T2 n1 = { -(s1 * b1.y + s2 * b2.y), s1 * b1.x + s2 * b2.x };
        const T2 n2 = { -(s2 * b1.y - s1 * b2.y), s2 * b1.x - s1 * b2.x };
        v[0] = cadd(v[0], cadd(a1, a2));
        v[1] = cadd(m1, n1);
        v[4] = csub(m1, n1);
        v[2] = cadd(m2, n2);
        v[3] = csub(m2, n2);
    }
};

// One Stockham pass of radix R over W pencils: src -> dst (both [n][w], pitch SW).
// tw: table of exp(-2 pi i k / N), k in [0, N), in LDS or global memory.
template <typename T, int R>
__device__ __forceinline__ void stockham_pass(const cx<T>* __restrict__ src, cx<T>* __restrict__ dst,
                                              const cx<T>* __restrict__ tw, int n, int w_count, int sw, int l,
                                              bool fwd)
{
    using T2 = cx<T>;
    const int m = n / (l * R);
    const int per_pencil = n / R;
    const int total = per_pencil * w_count;
    const int tw_step = n / (l * R);  // omega_{lR}^{jq} = omega_N^{jq * N/(lR)}
    for (int b = threadIdx.x; b < total; b += blockDim.x) {
        const int w = b % w_count;
        const int bf = b / w_count;  // butterfly within pencil: (k, j)
        const int j = bf % l;
        const int k = bf / l;
        T2 v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = src[((q * m + k) * l + j) * sw + w];
        if (l > 1) {
#pragma unroll
            for (int q = 1; q < R; ++q) {
                T2 t = tw[(j * q * tw_step) % n];
                if (!fwd) t.y = -t.y;
                v[q] = cmul(v[q], t);
            }
        }
        Dft<T, R>::run(v, fwd);
#pragma unroll
        for (int s = 0; s < R; ++s) dst[((k * R + s) * l + j) * sw + w] = v[s];
    }
}

// Full in-LDS transform of W pencils.  Returns the buffer holding the result
// (a or b).  Caller must __syncthreads() before (data loaded) — this function
// synchronises after every pass.
template <typename T>
__device__ cx<T>* lds_fft(cx<T>* a, cx<T>* b, const cx<T>* tw, const FftPlan& plan, int w_count, int sw, bool fwd)
{
    cx<T>* src = a;
    cx<T>* dst = b;
    int l = 1;
    for (int p = 0; p < plan.count; ++p) {
        const int r = plan.radix[p];
        switch (r) {
        case 8: stockham_pass<T, 8>(src, dst, tw, plan.n, w_count, sw, l, fwd); break;
        case 4: stockham_pass<T, 4>(src, dst, tw, plan.n, w_count, sw, l, fwd); break;
        case 2: stockham_pass<T, 2>(src, dst, tw, plan.n, w_count, sw, l, fwd); break;
        case 3: stockham_pass<T, 3>(src, dst, tw, plan.n, w_count, sw, l, fwd); break;
        default: stockham_pass<T, 5>(src, dst, tw, plan.n, w_count, sw, l, fwd); break;
        }
        __syncthreads();
        l *= r;
        cx<T>* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

}  // namespace sphip
