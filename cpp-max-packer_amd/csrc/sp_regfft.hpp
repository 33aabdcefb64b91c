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

// A + W B and A - W B for W = c + i s written as c (1 + i tau), tau = s / c:
// u = B (1 + i tau) takes two fused multiply-adds, and each output two more,
// six operations where a complex product and two complex additions take
// eight.
template <typename T>
__device__ __forceinline__ void bfly_tan(cx<T> A, cx<T> B, T c, T tau, cx<T>& plus, cx<T>& minus)
{
    const T ux = fma(-tau, B.y, B.x), uy = fma(tau, B.x, B.y);
    plus = cx<T> { fma(c, ux, A.x), fma(c, uy, A.y) };
    minus = cx<T> { fma(-c, ux, A.x), fma(-c, uy, A.y) };
}

// Radix-16 in registers as 4 x 4.  The first radix-4 stage runs on the four
// columns n2; the w16^(n2 k1) twiddles of the second are folded into its
// butterflies: with y_n the column k1 inputs and W = w16^k1, the radix-4
// DFT of (y0, W y1, W^2 y2, W^3 y3) is
//   b = y0 +- W^2 y2, c = y1 +- W^2 y3,
//   X0, X2 = b+ +- W c+,  X1, X3 = b- +- (w4 W) c-,
// each "+-" pair one bfly_tan (w4 = -+i, the radix-4 root).
template <typename T> __device__ __forceinline__ void dft16(cx<T>* v, bool fwd)
{
    using T2 = cx<T>;
    // cos and tan of m pi / 8; the sine carries the direction's sign sg
    const T c1 = T(0.92387953251128675613), s1 = T(0.38268343236508977173), h = T(0.70710678118654752440);
    const T t1 = T(0.41421356237309504880), t3 = T(2.41421356237309504880);
    const T sg = fwd ? T(-1) : T(1);
    T2 y[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        T2 a[4] = { v[n2], v[4 + n2], v[8 + n2], v[12 + n2] };
        Dft<T, 4>::run(a, fwd);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) y[n2 * 4 + k1] = a[k1];
    }
    {
        T2 a[4] = { y[0], y[4], y[8], y[12] };
        Dft<T, 4>::run(a, fwd);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[4 * k2] = a[k2];
    }
    // (c, tan) of w16^m = cos(m pi / 8) + i sg sin(m pi / 8)
    auto col = [&](int k1, T c2, T tau2, T c1_, T tau1, T cv, T tauv) {
        const T2 y0 = y[k1], y1 = y[4 + k1], y2 = y[8 + k1], y3 = y[12 + k1];
        T2 bp, bm, cp, cm;
        if (k1 == 2) {  // W^2 = w16^4 = w4
            const T2 r2 = mul_mi(y2, fwd), r3 = mul_mi(y3, fwd);
            bp = cadd(y0, r2);
            bm = csub(y0, r2);
            cp = cadd(y1, r3);
            cm = csub(y1, r3);
        }
        else {
            bfly_tan(y0, y2, c2, tau2, bp, bm);
            bfly_tan(y1, y3, c2, tau2, cp, cm);
        }
        bfly_tan(bp, cp, c1_, tau1, v[k1], v[k1 + 8]);
        bfly_tan(bm, cm, cv, tauv, v[k1 + 4], v[k1 + 12]);
    };
    // k1 = 1: W^2 = w16^2, W = w16^1, w4 W = w16^5
    col(1, h, sg, c1, sg * t1, -s1, -sg * t3);
    // k1 = 2: W = w16^2, w4 W = w16^6
    col(2, T(0), T(0), h, sg, -h, -sg);
    // k1 = 3: W^2 = w16^6, W = w16^3, w4 W = w16^7
    col(3, -h, -sg, s1, sg * t3, -c1, -sg * t1);
}

// Radix-16 in registers as 4 x 4 with the w16 twiddles applied between the
// two radix-4 stages (a complex product each): fewer values live at once
// than in dft16's folded butterflies, for kernels at the edge of their
// register budget.
template <typename T> __device__ __forceinline__ void dft16_twiddled(cx<T>* v, bool fwd)
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

// FOLD (radix 16): dft16's folded butterflies, else dft16_twiddled.
template <typename T, int R, bool FOLD = true> __device__ __forceinline__ void dft_reg(cx<T>* v, bool fwd)
{
    if constexpr (R == 16) {
        if constexpr (FOLD) dft16<T>(v, fwd);
        else dft16_twiddled<T>(v, fwd);
    }
    else Dft<T, R>::run(v, fwd);
}

template <int N, int R> struct RegShape {
    static constexpr int P = N / R;
    // N = 128 at R = 16 is the mixed 16 x 8 transform (reg_fft_16x8): eight
    // threads a pencil, sixteen values each, as for a power of R
    static constexpr bool mixed = N == 128 && R == 16;
    static constexpr int passes = (N == R) ? 1 : (N == R * R) ? 2 : (N == R * R * R) ? 3 : (N == R * R * R * R) ? 4
                                : mixed ? 2 : -1;
    static_assert(passes > 0, "N must be a power of R, or 128 at R = 16");
    // LDS elements of one pencil: one pad slot every 16 plus one, so lanes of
    // consecutive pencils and stride-16 accesses land in distinct banks.
    static constexpr int stride = N + N / 16 + 1;
    __device__ static constexpr int pad(int i) { return i + (i >> 4); }
};

// The 128-point transform as 16 x 8 (decimation in time, one exchange), in
// the distribution of a power-of-R transform: eight threads a pencil, thread
// t holding x[t + 8 q] (q < 16) in and X[t + 8 s] (s < 16) out.  With
// n = n2 + 8 n1 and k = k1 + 16 k2,
//   X[k1 + 16 k2] = sum_n2 w128^(n2 k1) w8^(n2 k2) Y[n2][k1],
//   Y[n2][k1]     = sum_n1 w16^(n1 k1) x[n2 + 8 n1]:
// thread t computes Y[t][.] (a radix 16 over its own values), writes it as
// row t of the exchange, and takes the two columns k1 = t and k1 = t + 8,
// whose twiddled radix 8s give X[t + 16 k2] and X[t + 8 + 16 k2], i.e.
// X[t + 8 s] for even and odd s.  LEAD > 0: x[t + 8 q] is zero for
// q >= LEAD (dft16_lead).  lds: the pencil's region (RegShape::stride
// elements); tw: exp(-2 pi i k / 128).
template <typename T, int LEAD, bool WAVE>
__device__ __forceinline__ void reg_fft_16x8(cx<T>* v, cx<T>* lds, const cx<T>* tw, int t, bool fwd)
{
    using T2 = cx<T>;
    using S = RegShape<128, 16>;
    static_assert(!WAVE || 64 % S::P == 0, "a wave-local pencil must lie within one wave");
    auto sync = [] {
        if constexpr (WAVE) __builtin_amdgcn_wave_barrier();
        else __syncthreads();
    };
    __builtin_assume(t >= 0 && t < 8);
    if constexpr (LEAD > 0) dft16_lead<T, LEAD>(v, fwd);
    else dft16<T>(v, fwd);
    sync();  // earlier reads of this region are done
#pragma unroll
    for (int s = 0; s < 16; ++s) lds[S::pad(16 * t + s)] = v[s];
    sync();
    T2 a[8], b[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        a[q] = lds[S::pad(16 * q + t)];
        b[q] = lds[S::pad(16 * q + t + 8)];
        if (q > 0) {
            T2 wa = tw[(q * t) & 127], wb = tw[(q * (t + 8)) & 127];
            if (!fwd) {
                wa.y = -wa.y;
                wb.y = -wb.y;
            }
            a[q] = cmul(a[q], wa);
            b[q] = cmul(b[q], wb);
        }
    }
    Dft<T, 8>::run(a, fwd);
    Dft<T, 8>::run(b, fwd);
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) {
        v[2 * k2] = a[k2];
        v[2 * k2 + 1] = b[k2];
    }
}

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
    if constexpr (S::mixed) {
        reg_fft_16x8<T, LEAD, WAVE>(v, lds, tw, t, fwd);
    }
    else {
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
}

struct NoWork {
    __device__ void operator()() const {}
};

// reg_fft_lean (below) after its first radix-R pass and its barrier: the
// exchange and the second pass.  A caller that splits the transform around a
// block barrier, some waves taking only the barrier, runs dft_reg, the
// barrier, then this.
template <typename T, int N, int R, typename Mid = NoWork, int ROW = R + 1, bool FOLD = true>
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
    dft_reg<T, R, FOLD>(v, fwd);
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
// 16-byte stores, the region itself being 16-byte aligned).  FOLD: as in
// dft_reg.
template <typename T, int N, int R, bool SYNC = true, typename Mid = NoWork, int ROW = R + 1, bool FOLD = true>
__device__ __forceinline__ void reg_fft_lean(cx<T>* v, cx<T>* lds, const cx<T>* tq, int t, bool fwd, Mid mid = {})
{
    // a fresh copy of t per call: the twiddles are then read again here
    // rather than kept in registers from an earlier call
    asm volatile("" : "+v"(t));
    __builtin_assume(t >= 0 && t < R);
    dft_reg<T, R, FOLD>(v, fwd);
    if (SYNC) __syncthreads();
    else __builtin_amdgcn_wave_barrier();
    reg_fft_lean_second<T, N, R, Mid, ROW, FOLD>(v, lds, tq, t, fwd, mid);
}

}  // namespace sphip
