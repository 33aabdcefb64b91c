// Register-resident FFTs for lengths N = R^k (R = 8 or 16) on gfx950.
//
// A pencil of N points is owned by P = N/R threads, each holding R values in
// registers.  The transform is a Stockham sequence of k radix-R passes; the
// first pass reads x[t + P*q] straight from global memory (coalesced across t),
// the last pass leaves X[t + P*s] in thread t's registers (natural order,
// stride P), and passes exchange through LDS only between them — one exchange
// for N = 256 (16 x 16), two for N = 512 (8 x 8 x 8).  The output distribution
// of a forward transform equals the input distribution of the next one, so a
// product of spectra followed by an inverse transform needs no exchange.
#pragma once

#include "sp_fft.hpp"

namespace sphip {

// Radix-16 in registers as 4 x 4 with the w16 twiddles in between.
template <typename T> __device__ __forceinline__ void dft16(cx<T>* v, bool fwd)
{
    using T2 = cx<T>;
    const T c1 = T(0.92387953251128675613), s1 = T(0.38268343236508977173), h = T(0.70710678118654752440);
    T2 y[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        T2 a[4] = { v[n2], v[4 + n2], v[8 + n2], v[12 + n2] };
        Dft<T, 4>::run(a, fwd);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) y[n2 * 4 + k1] = a[k1];
    }
    // twiddle y[n2][k1] *= w16^(n2*k1)
    const T sg = fwd ? T(-1) : T(1);
    // w16^m = cos(2 pi m/16) + sg*i*sin(2 pi m/16)
    auto tw = [&](T2 a, T c, T s) { return T2 { a.x * c - a.y * sg * s, a.x * sg * s + a.y * c }; };
    y[1 * 4 + 1] = tw(y[1 * 4 + 1], c1, s1);   // m=1
    y[1 * 4 + 2] = tw(y[1 * 4 + 2], h, h);     // m=2
    y[1 * 4 + 3] = tw(y[1 * 4 + 3], s1, c1);   // m=3
    y[2 * 4 + 1] = tw(y[2 * 4 + 1], h, h);     // m=2
    y[2 * 4 + 2] = mul_mi(y[2 * 4 + 2], fwd);  // m=4
    y[2 * 4 + 3] = tw(y[2 * 4 + 3], -h, h);    // m=6
    y[3 * 4 + 1] = tw(y[3 * 4 + 1], s1, c1);   // m=3
    y[3 * 4 + 2] = tw(y[3 * 4 + 2], -h, h);    // m=6
    y[3 * 4 + 3] = tw(y[3 * 4 + 3], -c1, -s1); // m=9
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        T2 a[4] = { y[0 * 4 + k1], y[1 * 4 + k1], y[2 * 4 + k1], y[3 * 4 + k1] };
        Dft<T, 4>::run(a, fwd);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = a[k2];
    }
}

// dft16 of an input whose v[q] are zero for q >= NQ (NQ <= 3): the same 4 x 4
// decomposition with the zero terms left out.  The first radix-4 stage of
// column n2 then has the single input v[n2]; the second adds the NQ nonzero
// rows.  Every nonzero term is computed as in dft16, so the outputs equal
// dft16's.
template <typename T, int NQ> __device__ __forceinline__ void dft16_lead(cx<T>* v, bool fwd)
{
    static_assert(NQ >= 1 && NQ <= 3, "leading nonzero inputs");
    using T2 = cx<T>;
    const T c1 = T(0.92387953251128675613), s1 = T(0.38268343236508977173), h = T(0.70710678118654752440);
    const T sg = fwd ? T(-1) : T(1);
    auto tw = [&](T2 a, T c, T s) { return T2 { a.x * c - a.y * sg * s, a.x * sg * s + a.y * c }; };
    // y[n2][k1] = v[n2] w16^(n2 k1) for the nonzero rows n2 < NQ
    T2 y0 = v[0];
    T2 y1[4], y2[4];
    if constexpr (NQ >= 2) {
        y1[0] = v[1];
        y1[1] = tw(v[1], c1, s1);  // m=1
        y1[2] = tw(v[1], h, h);    // m=2
        y1[3] = tw(v[1], s1, c1);  // m=3
    }
    if constexpr (NQ >= 3) {
        y2[0] = v[2];
        y2[1] = tw(v[2], h, h);    // m=2
        y2[2] = mul_mi(v[2], fwd); // m=4
        y2[3] = tw(v[2], -h, h);   // m=6
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        // Dft<T, 4> of (y0, y1[k1], y2[k1], 0)
        if constexpr (NQ == 1) {
            v[k1] = v[k1 + 4] = v[k1 + 8] = v[k1 + 12] = y0;
        }
        else {
            const T2 a = NQ >= 3 ? cadd(y0, y2[k1]) : y0, b = NQ >= 3 ? csub(y0, y2[k1]) : y0;
            const T2 c = y1[k1];
            const T2 d = mul_mi(y1[k1], fwd);
            v[k1] = cadd(a, c);
            v[k1 + 4] = cadd(b, d);
            v[k1 + 8] = csub(a, c);
            v[k1 + 12] = csub(b, d);
        }
    }
}

template <typename T, int R> __device__ __forceinline__ void dft_reg(cx<T>* v, bool fwd)
{
    if constexpr (R == 16) dft16<T>(v, fwd);
    else Dft<T, R>::run(v, fwd);
}

template <int N, int R> struct RegShape {
    static constexpr int P = N / R;
    static constexpr int passes = (N == R) ? 1 : (N == R * R) ? 2 : (N == R * R * R) ? 3 : (N == R * R * R * R) ? 4 : -1;
    static_assert(passes > 0, "N must be a power of R");
    // LDS elements of one pencil: one pad slot every 16 plus one, so lanes of
    // consecutive pencils and stride-16 accesses land in distinct banks.
    static constexpr int stride = N + N / 16 + 1;
    __device__ static constexpr int pad(int i) { return i + (i >> 4); }
};

// Runs the transform for one pencil.  v: in = x[t + P*q], out = X[t + P*s].
// lds: this pencil's region (RegShape::stride elements); tw: exp(-2 pi i k/N)
// table (LDS).  Every thread of the block must call it (barriers inside).
// LEAD > 0 (radix 16): x[t + P*q] is zero for q >= LEAD, and the first pass
// leaves those terms out (dft16_lead).  WAVE: the pencil's P threads are P
// consecutive lanes of one wave (t = lane % P, P dividing 64) and nothing else
// uses its LDS region; LDS instructions of a wave complete in order, so the
// exchanges need no block barrier.  The caller synchronises the block once
// after filling the twiddle table.
template <typename T, int N, int R, int LEAD = 0, bool WAVE = false>
__device__ __forceinline__ void reg_fft(cx<T>* v, cx<T>* lds, const cx<T>* tw, int t, bool fwd)
{
    using T2 = cx<T>;
    using S = RegShape<N, R>;
    static_assert(!WAVE || 64 % S::P == 0, "a wave-local pencil must lie within one wave");
    auto sync = [] {
        if constexpr (WAVE) __builtin_amdgcn_wave_barrier();
        else __syncthreads();
    };
    // t < P lets the compiler drop the wrap of every twiddle index and fold
    // the padded LDS offsets into constants
    __builtin_assume(t >= 0 && t < S::P);
    int l = 1;
#pragma unroll
    for (int p = 0; p < S::passes; ++p) {
        const int m = N / (l * R);
        const int j = t % l, k = t / l;
        if (p > 0) {
            // gather this pass's inputs: src[(q*m + k)*l + j]
            sync();
#pragma unroll
            for (int q = 0; q < R; ++q) v[q] = lds[S::pad((q * m + k) * l + j)];
#pragma unroll
            for (int q = 1; q < R; ++q) {
                T2 w = tw[(j * q * m) & (N - 1)];
                if (!fwd) w.y = -w.y;
                v[q] = cmul(v[q], w);
            }
        }
        if constexpr (LEAD > 0 && R == 16) {
            if (p == 0) dft16_lead<T, LEAD>(v, fwd);
            else dft_reg<T, R>(v, fwd);
        }
        else
            dft_reg<T, R>(v, fwd);
        if (p + 1 < S::passes) {
            sync();
#pragma unroll
            for (int s = 0; s < R; ++s) lds[S::pad((k * R + s) * l + j)] = v[s];
        }
        l *= R;
    }
}

struct NoWork {
    __device__ void operator()() const {}
};

// reg_fft_lean (below) after its first radix-R pass and its barrier: the
// exchange and the second pass.  A caller that splits the transform around a
// block barrier, some waves taking only the barrier, runs dft_reg, the
// barrier, then this.
template <typename T, int N, int R, typename Mid = NoWork, int ROW = R + 1>
__device__ __forceinline__ void reg_fft_lean_second(cx<T>* v, cx<T>* lds, const cx<T>* tq, int t, bool fwd,
                                                    Mid mid = {})
{
    using S = RegShape<N, R>;
    static_assert(S::passes == 2 && R == 16, "one exchange of 16 x 16");
    static_assert(ROW == R + 1 || (ROW == R + 2 && sizeof(T) == 4), "row of 17 values, or of 18 FP32 values");
    asm volatile("" : "+v"(t));
    __builtin_assume(t >= 0 && t < R);
    // row t holds X1[16 t + s]; the second pass reads column t (l = R, m = 1,
    // j = t, k = 0)
    if constexpr (ROW == R + 2) {
#pragma unroll
        for (int s = 0; s < R; s += 2)
            *reinterpret_cast<float4*>(&lds[ROW * t + s]) = float4 { v[s].x, v[s].y, v[s + 1].x, v[s + 1].y };
    }
    else {
#pragma unroll
        for (int s = 0; s < R; ++s) lds[ROW * t + s] = v[s];
    }
    mid();
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int q0 = 0; q0 < R; q0 += 4) {
#pragma unroll
        for (int q = q0; q < q0 + 4; ++q) {
            v[q] = lds[ROW * q + t];
            if (q > 0) v[q] = cmul(v[q], tq[q * R + t]);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    dft_reg<T, R>(v, fwd);
}

// reg_fft for N = R * R (one exchange) with the exchange twiddles taken from
// a [q][t] table, tq[q * R + t] = w^(t q) in the transform's direction, so
// that every twiddle read is one base address plus a constant, and with the
// exchange reads and their twiddles issued four at a time (a kernel holding
// other data in registers keeps the transform's working set small).  Same
// arithmetic as reg_fft.
// The pencil's 16 threads are lanes of one wave (t = lane % 16 within a group
// of 16 consecutive lanes), and LDS instructions of a wave complete in order,
// so the exchange itself needs no block barrier.  SYNC: the pencil's region
// may still be read by other waves (it aliases a buffer they use), so the
// block synchronises before the first write.
// `mid` runs between the exchange's writes and its reads, when the pencil's
// registers hold nothing: a caller can issue loads for later work there.
// ROW: LDS values per exchange row (17: one pad value; 18 in FP32: rows
// start on 16-byte boundaries and each thread writes its row as eight
// 16-byte stores, the region itself being 16-byte aligned).
template <typename T, int N, int R, bool SYNC = true, typename Mid = NoWork, int ROW = R + 1>
__device__ __forceinline__ void reg_fft_lean(cx<T>* v, cx<T>* lds, const cx<T>* tq, int t, bool fwd, Mid mid = {})
{
    // a fresh copy of t per call: the twiddles are then read again here
    // rather than kept in registers from an earlier call
    asm volatile("" : "+v"(t));
    __builtin_assume(t >= 0 && t < R);
    dft_reg<T, R>(v, fwd);
    if (SYNC) __syncthreads();
    else __builtin_amdgcn_wave_barrier();
    reg_fft_lean_second<T, N, R, Mid, ROW>(v, lds, tq, t, fwd, mid);
}

}  // namespace sphip
