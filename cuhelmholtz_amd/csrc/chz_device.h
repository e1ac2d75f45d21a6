// chz_device.h -- device building blocks of the line-transform passes.
//
// One workgroup owns a tile of L lines of one axis.  A line holds the N+2
// real slots of the padded layout (N cells, M = N/2 complex points).  In LDS
// the tile is stored as complex points, point-major: point k of line c sits
// at buf[k*LS + c] with LS = tile_ls(L, sizeof(T)) (L, plus one padding
// point for FP32).  Real slot s of line c is component (s & 1) of point s >> 1.
//
// Thread t of the workgroup works on line c = t % L as sub-thread
// u = t / L (TPL = blockDim / L sub-threads per line), so consecutive lanes
// touch consecutive lines: every LDS access of the fold, FFT and unfold
// phases is a contiguous run of L points.
//
// The FFT is a Stockham autosort over M points with each thread holding up to
// EPT points in registers (M <= TPL * EPT); a stage of radix R runs up to
// EPT/R butterflies per thread.  Only the forward transform (e^{-2 pi i/M}) is
// implemented; the inverse used by the C2R-type kinds is conj(FFT(conj(.))),
// folded into the fold and unfold phases.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace chz {

template <typename T> struct vec2_of;
template <> struct vec2_of<double> { typedef double2 type; };
template <> struct vec2_of<float> { typedef float2 type; };

template <typename V> __device__ __forceinline__ V mk(decltype(V().x) a, decltype(V().x) b) { V r; r.x = a; r.y = b; return r; }
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return mk<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return mk<V>(a.x - b.x, a.y - b.y); }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) { return mk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename V> __device__ __forceinline__ V cconj(V a) { return mk<V>(a.x, -a.y); }
template <typename V> __device__ __forceinline__ V cmul_mi(V a) { return mk<V>(a.y, -a.x); }   // a * (-i)
template <typename V> __device__ __forceinline__ V cmul_pi(V a) { return mk<V>(-a.y, a.x); }   // a * (+i)
template <typename V> __device__ __forceinline__ V cscale(V a, decltype(V().x) s) { return mk<V>(a.x * s, a.y * s); }

// cos(2 pi k / 48) for any integer k, from the exact values at multiples of
// pi/24 that the radices 2,3,4,6,8,12,16 need (k a multiple of 3 or 4).
__host__ __device__ constexpr double cos48(int k)
{
    k %= 48;
    if (k < 0) k += 48;
    if (k > 24) k = 48 - k;               // cos(2pi - a) = cos(a)
    bool neg = false;
    if (k > 12) { k = 24 - k; neg = true; } // cos(pi - a) = -cos(a)
    double v = k == 0 ? 1.0
             : k == 3 ? 0.92387953251128675613    // cos(pi/8)
             : k == 4 ? 0.86602540378443864676    // cos(pi/6)
             : k == 6 ? 0.70710678118654752440    // cos(pi/4)
             : k == 8 ? 0.5                       // cos(pi/3)
             : k == 9 ? 0.38268343236508977173    // cos(3pi/8)
             : 0.0;                               // k == 12
    return neg ? -v : v;
}
// cos(2 pi k / 20) for any integer k (multiples of 18 degrees, radices 5, 10, 20)
__host__ __device__ constexpr double cos20(int k)
{
    k %= 20;
    if (k < 0) k += 20;
    if (k > 10) k = 20 - k;
    bool neg = false;
    if (k > 5) { k = 10 - k; neg = true; }
    double v = k == 0 ? 1.0
             : k == 1 ? 0.95105651629515357212    // cos(pi/10)
             : k == 2 ? 0.80901699437494742410    // cos(pi/5)
             : k == 3 ? 0.58778525229247312917    // cos(3pi/10)
             : k == 4 ? 0.30901699437494742410    // cos(2pi/5)
             : 0.0;                               // k == 5
    return neg ? -v : v;
}
// W_R^e = e^{-2 pi i e / R}: (cos, -sin); -sin(a) = -cos(pi/2 - a)
template <int R, int E> struct Root {
    static constexpr bool b48 = 48 % R == 0;
    static constexpr int k = (E % R) * (b48 ? 48 / R : 20 / R);
    static constexpr double c = b48 ? cos48(k) : cos20(k);
    static constexpr double s = b48 ? -cos48(12 - k) : -cos20(5 - k);
};

template <int Begin, int End, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (Begin < End) {
        f(std::integral_constant<int, Begin>{});
        static_for<Begin + 1, End>(f);
    }
}

// a * W_R^E with the trivial cases resolved at compile time
template <int R, int E, typename V> __device__ __forceinline__ V rot(V a)
{
    constexpr int e = ((E % R) + R) % R;
    if constexpr (e == 0) return a;
    else if constexpr (2 * e == R) return mk<V>(-a.x, -a.y);
    else if constexpr (4 * e == R) return cmul_mi(a);
    else if constexpr (4 * e == 3 * R) return cmul_pi(a);
    else {
        typedef decltype(V().x) T;
        constexpr T c = (T)Root<R, e>::c, s = (T)Root<R, e>::s;
        return mk<V>(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

// ---- forward small DFTs in registers: x[o + st*r], r = 0..R-1 -------------
template <int R> struct Dft;

template <> struct Dft<2> {
    template <typename V> __device__ __forceinline__ static void run(V *x, int o, int st)
    {
        V a = x[o], b = x[o + st];
        x[o] = cadd(a, b);
        x[o + st] = csub(a, b);
    }
};
template <> struct Dft<3> {
    template <typename V> __device__ __forceinline__ static void run(V *x, int o, int st)
    {
        typedef decltype(V().x) T;
        const T h = (T)0.86602540378443864676;   // sin(2 pi/3)
        V a0 = x[o], a1 = x[o + st], a2 = x[o + 2 * st];
        V t1 = cadd(a1, a2), t2 = csub(a1, a2);
        V m = mk<V>(a0.x - (T)0.5 * t1.x, a0.y - (T)0.5 * t1.y);
        V n = mk<V>(h * t2.y, -h * t2.x);          // -i * sin(2pi/3) * t2
        x[o] = cadd(a0, t1);
        x[o + st] = cadd(m, n);
        x[o + 2 * st] = csub(m, n);
    }
};
template <> struct Dft<4> {
    template <typename V> __device__ __forceinline__ static void run(V *x, int o, int st)
    {
        V a0 = x[o], a1 = x[o + st], a2 = x[o + 2 * st], a3 = x[o + 3 * st];
        V t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = cmul_mi(csub(a1, a3));
        x[o] = cadd(t0, t2);
        x[o + 2 * st] = csub(t0, t2);
        x[o + st] = cadd(t1, t3);
        x[o + 3 * st] = csub(t1, t3);
    }
};
template <> struct Dft<5> {
    template <typename V> __device__ __forceinline__ static void run(V *x, int o, int st)
    {
        typedef decltype(V().x) T;
        const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;   // cos(2pi/5), cos(4pi/5)
        const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;    // sin(2pi/5), sin(4pi/5)
        V a0 = x[o], a1 = x[o + st], a2 = x[o + 2 * st], a3 = x[o + 3 * st], a4 = x[o + 4 * st];
        V t1 = cadd(a1, a4), t2 = cadd(a2, a3), t3 = csub(a1, a4), t4 = csub(a2, a3);
        V b1 = mk<V>(a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y);
        V b2 = mk<V>(a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y);
        V d1 = mk<V>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
        V d2 = mk<V>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
        x[o] = mk<V>(a0.x + t1.x + t2.x, a0.y + t1.y + t2.y);
        x[o + st] = cadd(b1, cmul_mi(d1));
        x[o + 4 * st] = csub(b1, cmul_mi(d1));
        x[o + 2 * st] = cadd(b2, cmul_mi(d2));
        x[o + 3 * st] = csub(b2, cmul_mi(d2));
    }
};
// R = P*Q: input index n = Q*n1 + n2, output index k = k1 + P*k2.
// P-point DFTs over n1 for each n2, twiddle W_R^{n2 k1}, Q-point DFTs over n2.
template <int P, int Q> struct DftPQ {
    template <typename V> __device__ __forceinline__ static void run(V *x, int o, int st)
    {
        constexpr int R = P * Q;
        V y[R];
        static_for<0, Q>([&](auto n2c) {
            constexpr int n2 = decltype(n2c)::value;
            V t[P];
            static_for<0, P>([&](auto n1c) { t[decltype(n1c)::value] = x[o + st * (Q * decltype(n1c)::value + n2)]; });
            Dft<P>::run(t, 0, 1);
            static_for<0, P>([&](auto k1c) {
                constexpr int k1 = decltype(k1c)::value;
                y[n2 * P + k1] = rot<R, n2 * k1>(t[k1]);
            });
        });
        static_for<0, P>([&](auto k1c) {
            constexpr int k1 = decltype(k1c)::value;
            V t[Q];
            static_for<0, Q>([&](auto n2c) { t[decltype(n2c)::value] = y[decltype(n2c)::value * P + k1]; });
            Dft<Q>::run(t, 0, 1);
            static_for<0, Q>([&](auto k2c) { x[o + st * (k1 + P * decltype(k2c)::value)] = t[decltype(k2c)::value]; });
        });
    }
};
template <> struct Dft<6> : DftPQ<2, 3> {};
template <> struct Dft<10> : DftPQ<2, 5> {};
template <> struct Dft<8> : DftPQ<2, 4> {};
template <> struct Dft<12> : DftPQ<4, 3> {};
template <> struct Dft<16> : DftPQ<4, 4> {};

// ---- tile addressing ---------------------------------------------------
template <typename T> struct Tile {
    typedef typename vec2_of<T>::type V;
    V *buf;     // (M+1) * LS points
    int LS;     // point stride (tile_ls(L))
    int c;      // line of this thread
    int L;      // lines of the tile
    __device__ __forceinline__ V &pt(int k) const { return buf[k * LS + c]; }
    __device__ __forceinline__ T &re(int s) const { return reinterpret_cast<T *>(buf)[(size_t)((s >> 1) * LS + c) * 2 + (s & 1)]; }
};

// One Stockham stage of radix R over M points, in place.  Thread u of a line
// runs the butterflies b = u + i*TPL (i < EPT/R, b < M/R), so the stage fits
// when ceil(M/(R*TPL)) <= EPT/R (the host planner guarantees it).  NSC > 0:
// the span Ns as a compile-time constant (b mod Ns and the twiddle stride
// then need no integer division).
template <int R, int EPT, int NSC = 0, typename TL>
__device__ __forceinline__ void fft_stage(const TL &tl, int u, int TPL, int M, int Ns_rt, const typename TL::V *__restrict__ tw)
{
    typedef typename TL::V V;
    constexpr int NB = EPT / R;
    const int Ns = NSC > 0 ? NSC : Ns_rt;
    const int MR = M / R;
    V v[NB * R];
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const int b = u + i * TPL;
        if (b < MR) {
#pragma unroll
            for (int r = 0; r < R; r++) v[i * R + r] = tl.pt(b + r * MR);
        }
    }
    __syncthreads();
    const int twstride = M / (Ns * R);
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const int b = u + i * TPL;
        if (b < MR) {
            const int j = b % Ns;
            V *x = v + i * R;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) x[r] = cmul(x[r], tw[r * j * twstride]);
            }
            Dft<R>::run(x, 0, 1);
            const int base = (b - j) * R + j;
#pragma unroll
            for (int q = 0; q < R; q++) tl.pt(base + q * Ns) = x[q];
        }
    }
    __syncthreads();
}

// One Stockham stage of any radix R as a direct DFT (the factors without a
// specialised butterfly: primes from 7 up, or the whole remaining factor).
// Output q of butterfly b is sum_r x_{b + r M/R} W_{Ns R}^{r j} W_R^{r q}
// (j = b mod Ns) = sum_r x_{b + r M/R} W_M^{r (j M/(Ns R) + q M/R)}: every
// factor comes from the M-point twiddle table.  A thread computes outputs
// o = u + i TPL (o = b + q M/R) straight from LDS into registers, so any R
// fits as long as M <= EPT * TPL; the work is M * R multiply-adds per line.
template <int EPT, typename V, typename PT>
__device__ __forceinline__ void stage_any_body(PT pt, int u, int TPL, int M, int Ns, int R, const V *__restrict__ tw)
{
    const int MR = M / R, twstride = MR / Ns;
    V out[EPT];
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        const int o = u + i * TPL;
        if (o < M) {
            const int b = o % MR, q = o / MR, j = b % Ns;
            const int step = j * twstride + q * MR;   // < M
            V acc = mk<V>(0, 0);
            int idx = 0;
            for (int r = 0; r < R; r++) {
                const V x = pt(b + r * MR), w = tw[idx];
                acc = mk<V>(acc.x + x.x * w.x - x.y * w.y, acc.y + x.x * w.y + x.y * w.x);
                idx += step;
                if (idx >= M) idx -= M;
            }
            out[i] = acc;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        const int o = u + i * TPL;
        if (o < M) {
            const int b = o % MR, q = o / MR, j = b % Ns;
            pt((b - j) * R + j + q * Ns) = out[i];
        }
    }
    __syncthreads();
}

// The direct stage on a plain tile (point k of line c at buf[k LS + c]), kept
// out of line: the kernels with run-time radix plans reach it from every
// inlined transform, and only sizes with large prime factors take it.
template <int EPT, typename V>
__device__ __noinline__ void fft_stage_any_plain(V *buf, int LS, int c, int u, int TPL, int M, int Ns, int R,
                                                 const V *__restrict__ tw)
{
    stage_any_body<EPT, V>([=](int k) -> V & { return buf[k * LS + c]; }, u, TPL, M, Ns, R, tw);
}

template <int EPT, typename TL>
__device__ __forceinline__ void fft_stage_any(const TL &tl, int u, int TPL, int M, int Ns, int R,
                                              const typename TL::V *__restrict__ tw)
{
    typedef typename TL::V V;
    if constexpr (std::is_same<TL, Tile<decltype(V().x)>>::value)
        fft_stage_any_plain<EPT, V>(tl.buf, tl.LS, tl.c, u, TPL, M, Ns, R, tw);
    else
        stage_any_body<EPT, V>([&](int k) -> V & { return tl.pt(k); }, u, TPL, M, Ns, R, tw);
}

// Full forward FFT: stages described by a packed radix list (one byte per
// stage, first stage in the low byte, 0 terminates).  Bytes 2, 3, 4, 5, 6, 8,
// 10, 12, 16 are the specialised butterflies; any other byte R > 1 is a
// direct stage of radix R, and byte 1 a direct stage of the whole remaining
// factor M / (product of the earlier stages).
template <int EPT, typename T, typename TL>
__device__ __forceinline__ void fft_forward(const TL &tl, int u, int TPL, int M, unsigned long long plan,
                                            const typename vec2_of<T>::type *__restrict__ tw)
{
    int Ns = 1;
    for (int st = 0; st < 8; st++) {
        int R = (int)((plan >> (8 * st)) & 255);
        if (R == 0) break;
        switch (R) {
            case 2: if constexpr (EPT >= 2) fft_stage<2, EPT>(tl, u, TPL, M, Ns, tw); break;
            case 3: if constexpr (EPT >= 3) fft_stage<3, EPT>(tl, u, TPL, M, Ns, tw); break;
            case 4: if constexpr (EPT >= 4) fft_stage<4, EPT>(tl, u, TPL, M, Ns, tw); break;
            case 5: if constexpr (EPT >= 5) fft_stage<5, EPT>(tl, u, TPL, M, Ns, tw); break;
            case 6: if constexpr (EPT >= 6) fft_stage<6, EPT>(tl, u, TPL, M, Ns, tw); break;
            case 8: if constexpr (EPT >= 8) fft_stage<8, EPT>(tl, u, TPL, M, Ns, tw); break;
            case 10: if constexpr (EPT >= 10) fft_stage<10, EPT>(tl, u, TPL, M, Ns, tw); break;
            case 12: if constexpr (EPT >= 12) fft_stage<12, EPT>(tl, u, TPL, M, Ns, tw); break;
            case 16: if constexpr (EPT >= 16) fft_stage<16, EPT>(tl, u, TPL, M, Ns, tw); break;
            default:
                if (R == 1) R = M / Ns;
                fft_stage_any<EPT>(tl, u, TPL, M, Ns, R, tw);
                break;
        }
        Ns *= R;
    }
}

// fft_forward with the radix plan a compile-time constant (the fixed-shape
// kernels): the stages unroll, so every stage's span Ns is a constant and
// the butterflies' index arithmetic needs no integer division (the loop
// above keeps Ns a run-time value and divides by it per butterfly).
template <int EPT, typename T, unsigned long long PLAN, int ST = 0, int NS = 1, typename TL>
__device__ __forceinline__ void fft_forward_c(const TL &tl, int u, int TPL, int M,
                                              const typename vec2_of<T>::type *__restrict__ tw)
{
    constexpr int R = ST < 8 ? (int)((PLAN >> (8 * ST)) & 255) : 0;
    if constexpr (R == 1) {
        // the whole remaining factor, one direct stage
        fft_stage_any<EPT>(tl, u, TPL, M, NS, M / NS, tw);
    } else if constexpr (R != 0) {
        constexpr bool SPEC = R == 2 || R == 3 || R == 4 || R == 5 || R == 6 || R == 8 || R == 10 || R == 12 || R == 16;
        if constexpr (SPEC && EPT >= R) fft_stage<R, EPT, NS>(tl, u, TPL, M, NS, tw);
        else fft_stage_any<EPT>(tl, u, TPL, M, NS, R, tw);
        fft_forward_c<EPT, T, PLAN, ST + 1, NS * R>(tl, u, TPL, M, tw);
    }
}

}  // namespace chz
