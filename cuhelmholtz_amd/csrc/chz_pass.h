// chz_pass.h -- the fused axis pass: stage a tile of lines into LDS, apply
// the per-kind fold, the M-point FFT and the per-kind unfold, store.
//
// Kind conventions are the reference's (all unnormalised, 1/2 of FFTW):
//   DST_1  p[1..N-1] -> p[1..N-1]   cudasymmfft/dst1funcinplace.cu:55-112
//   DCT_1  p[0..N]   -> p[0..N]     cudasymmfft/dct1funcinplace.cu:60-131
//   DCT_2  p[0..N-1] -> p[0..N-1]   cudasymmfft/dct2funcinplace.cu:49-112
//   DST_2  p[1..N]   -> p[1..N]     cudasymmfft/dst2funcinplace.cu:48-108
//   DCT_3  p[0..N-1] -> p[0..N-1]   cudasymmfft/dct3funcinplace.cu:47-95
//   DST_3  p[1..N]   -> p[1..N]     cudasymmfft/dst3funcinplace.cu:49-105 (p[0] taken as 0)
//   R2C    p[0..N-1] -> p[0..N+1]   cudasymmfft/dftfunc.cu:56-66
//   C2R    p[0..N+1] -> p[0..N-1]   cudasymmfft/dftfunc.cu:72-82
// The reference runs each as preOp kernel -> cuFFT Z2D/D2Z of length N ->
// postOp kernel.  Here the real-FFT split (N real <-> N/2 complex) is merged
// into the same fold/unfold so each kind costs exactly one M = N/2 point
// complex FFT on data that never leaves LDS.
#pragma once
#include "chz_device.h"
#include "chz_kernels.h"

namespace chz {

// The fold / unfold tables of a pass: w^k = (cos, sin)(2 pi k / N) for
// k <= M/2, and per n <= M either (cos, sin)(n pi / 2N) (DCT_2 / DST_2 /
// DCT_3 / DST_3) or (1 / sin(n pi / N), 0) (DST_1 / DCT_1).  A pointer to the
// pass's twiddle table in global memory reads them there (GTabs); LTabs reads
// copies staged in LDS, where the n-table holds whichever of the two the
// kind uses.
template <typename V> struct GTabs {
    const V *tw;
    int N;
    __device__ __forceinline__ V wk(int k) const { return tw[4 * k]; }
    __device__ __forceinline__ V cs(int n) const { return tw[n]; }
    __device__ __forceinline__ V rs(int n) const { return tw[TwOffsets::rsin(N) + n]; }
};
template <typename V> struct LTabs {
    const V *w, *n;
    __device__ __forceinline__ V wk(int k) const { return w[k]; }
    __device__ __forceinline__ V cs(int i) const { return n[i]; }
    __device__ __forceinline__ V rs(int i) const { return n[i]; }
};
template <typename V> __device__ __forceinline__ GTabs<V> tabs_of(const V *tw, int N) { return GTabs<V>{tw, N}; }
template <typename V> __device__ __forceinline__ LTabs<V> tabs_of(const LTabs<V> &t, int) { return t; }
// the n-table a kind uses: 1 = (cos, sin)(n pi / 2N), 2 = 1 / sin(n pi / N), 0 = none
__host__ __device__ constexpr int ntab_of(int k)
{
    return (k == K_DCT_2 || k == K_DST_2 || k == K_DCT_3 || k == K_DST_3) ? 1 : (k == K_DST_1 || k == K_DCT_1) ? 2 : 0;
}

// ---------------------------------------------------------------------------
// Fold: real line (slots) -> FFT input points 0..M-1, in place.
// ---------------------------------------------------------------------------

// Half-complex spectrum H_k (k = 0..M) that the reference's preOp builds for
// the Z2D-type kinds, read from the real line x.
template <int KIND, typename T, typename TL>
__device__ __forceinline__ typename vec2_of<T>::type hval(const TL &tl, int k, int M, int N)
{
    typedef typename vec2_of<T>::type V;
    const T h = (T)0.5;
    if constexpr (KIND == K_DST_1) {   // dst1funcinplace.cu:70-79
        if (k == 0) return mk<V>((T)2 * tl.re(1), 0);
        if (k == M) return mk<V>((T)-2 * tl.re(N - 1), 0);
        return mk<V>(tl.re(2 * k + 1) - tl.re(2 * k - 1), -tl.re(2 * k));
    } else if constexpr (KIND == K_DCT_1) {   // dct1funcinplace.cu:66-77 (even slots kept, odd differenced)
        if (k == 0) return mk<V>(tl.re(0), 0);
        if (k == M) return mk<V>(tl.re(N), 0);
        return mk<V>(tl.re(2 * k), tl.re(2 * k - 1) - tl.re(2 * k + 1));
    } else if constexpr (KIND == K_DCT_2) {   // dct2funcinplace.cu:63-71
        if (k == 0) return mk<V>(tl.re(0), 0);
        if (k == M) return mk<V>(tl.re(N - 1), 0);
        return mk<V>(h * (tl.re(2 * k) + tl.re(2 * k - 1)), h * (tl.re(2 * k) - tl.re(2 * k - 1)));
    } else if constexpr (KIND == K_DST_2) {   // dst2funcinplace.cu:55-69 (line read from p[1..N])
        if (k == 0) return mk<V>(tl.re(1), 0);
        if (k == M) return mk<V>(-tl.re(N), 0);
        return mk<V>(h * (tl.re(2 * k + 1) - tl.re(2 * k)), -h * (tl.re(2 * k + 1) + tl.re(2 * k)));
    } else {                                  // K_C2R: imaginary parts of DC and Nyquist do not enter
        if (k == 0) return mk<V>(tl.re(0), 0);
        if (k == M) return mk<V>(tl.re(N), 0);
        const V v = tl.pt(k);
        return mk<V>(v.x, v.y);
    }
}

template <int KIND, int EPT, typename T, typename TL, typename TB>
__device__ __forceinline__ void fold(const TL &tl, int u, int TPL, int M, int N, const TB &tb, T *red)
{
    typedef typename vec2_of<T>::type V;
    const auto tbs = tabs_of(tb, N);
    constexpr int PP = EPT / 2 + 1;        // >= ceil((M/2 + 1)/TPL)
    constexpr int PT = EPT + 1;            // >= ceil((M + 1)/TPL)
    if constexpr (kind_is_z2d(KIND)) {
        // Z2D via an M-point inverse FFT: C_k = E_k + i O_k,
        // E_k = H_k + conj H_{M-k}, O_k = (H_k - conj H_{M-k}) w^k, w = e^{2 pi i/N};
        // C_{M-k} = conj E_k + i conj O_k.  Stored conjugated (inverse = conj o FFT o conj).
        V ca[PP], cb[PP];
        if constexpr (KIND == K_DCT_1) {   // x1 = sum of odd samples (dct1funcinplace.cu:86-93)
            T s = 0;
            for (int j = u; j < M; j += TPL) s += tl.re(2 * j + 1);
            red[u * tl.L + tl.c] = s;
        }
        V wk[PP];
#pragma unroll
        for (int i = 0; i < PP; i++) {
            const int k = u + i * TPL;
            wk[i] = k <= M / 2 ? tbs.wk(k) : mk<V>(0, 0);
        }
#pragma unroll
        for (int i = 0; i < PP; i++) {
            const int k = u + i * TPL;
            if (k <= M / 2) {
                V Hk = hval<KIND, T>(tl, k, M, N), Hm = hval<KIND, T>(tl, M - k, M, N);
                V E = mk<V>(Hk.x + Hm.x, Hk.y - Hm.y);
                V D = mk<V>(Hk.x - Hm.x, Hk.y + Hm.y);
                V O = cmul(D, wk[i]);
                ca[i] = mk<V>(E.x - O.y, -(E.y + O.x));
                cb[i] = mk<V>(E.x + O.y, E.y - O.x);
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PP; i++) {
            const int k = u + i * TPL;
            if (k <= M / 2) {
                tl.pt(k) = ca[i];
                if (k != 0 && 2 * k != M) tl.pt(M - k) = cb[i];
            }
        }
    } else if constexpr (KIND != K_R2C) {
        // DCT_3 / DST_3 pre-rotation of the pairs (t, N-t), then the real line
        // itself is the FFT input (point m = x'_{2m} + i x'_{2m+1}).
        T va[PT], vb[PT];
        V csv[PT];
#pragma unroll
        for (int i = 0; i < PT; i++) {
            const int t = u + i * TPL;
            csv[i] = t <= M ? tbs.cs(t) : mk<V>(0, 0);   // (cos, sin)(t pi / 2N)
        }
#pragma unroll
        for (int i = 0; i < PT; i++) {
            const int t = u + i * TPL;
            if (t <= M) {
                const V cs = csv[i];
                if constexpr (KIND == K_DCT_3) {   // dct3funcinplace.cu:53-64 (its t+1 is t here)
                    if (t >= 1) {
                        const T a = tl.re(t), b = tl.re(N - t);
                        const T Ta = a + b, Tb = a - b;
                        va[i] = Ta * cs.y + Tb * cs.x;
                        vb[i] = Ta * cs.x - Tb * cs.y;
                    }
                } else {                   // DST_3, dst3funcinplace.cu:55-65 with p[0] = 0
                    const T a = t == 0 ? (T)0 : tl.re(t), b = tl.re(N - t);
                    const T Ta = a + b, Tb = a - b;
                    va[i] = Ta * cs.x + Tb * cs.y;
                    vb[i] = Ta * cs.y - Tb * cs.x;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PT; i++) {
            const int t = u + i * TPL;
            if (t <= M) {
                if constexpr (KIND == K_DCT_3) {
                    if (t >= 1) {
                        tl.re(t) = va[i];
                        if (2 * t != N) tl.re(N - t) = vb[i];
                    }
                } else {
                    tl.re(t) = va[i];
                    if (t != 0 && 2 * t != N) tl.re(N - t) = vb[i];
                }
            }
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Unfold: FFT output points -> real output slots, in place.
// ---------------------------------------------------------------------------
template <int KIND, int EPT, typename T, typename TL, typename TB>
__device__ __forceinline__ void unfold(const TL &tl, int u, int TPL, int M, int N, const TB &tb, const T *red)
{
    typedef typename vec2_of<T>::type V;
    const auto tbs = tabs_of(tb, N);
    constexpr int PP = EPT / 2 + 1;
    constexpr int PT = EPT + 1;
    const T h = (T)0.5;
    if constexpr (kind_is_z2d(KIND)) {
        // y_n (n = 0..N-1) is the real view of conj(FFT): y_{2m} = re, y_{2m+1} = -im.
        T oa[PT], ob[PT];
        T x1 = 0;
        if constexpr (KIND == K_DCT_1) {
            const int L = tl.L;
            for (int v = 0; v < TPL; v++) x1 += red[v * L + tl.c];
        }
        // per-n table value: (cos, sin)(n pi/2N) for DCT_2 / DST_2, (1/sin(n pi/N), 0) for DST_1 / DCT_1
        V tab[PT];
#pragma unroll
        for (int i = 0; i < PT; i++) {
            const int n = u + i * TPL;
            if constexpr (KIND == K_DCT_2 || KIND == K_DST_2) tab[i] = n <= M ? tbs.cs(n) : mk<V>(0, 0);
            else if constexpr (KIND != K_C2R) tab[i] = n <= M ? tbs.rs(n) : mk<V>(0, 0);
            else tab[i] = mk<V>(0, 0);
        }
#pragma unroll
        for (int i = 0; i < PT; i++) {
            const int n = u + i * TPL;
            if (n <= M) {
                const T a = (n & 1) ? -tl.re(n) : tl.re(n);
                const T b = n == 0 ? (T)0 : (((N - n) & 1) ? -tl.re(N - n) : tl.re(N - n));
                if constexpr (KIND == K_DST_1) {          // dst1funcinplace.cu:97-107
                    if (n == 0) { oa[i] = 0; ob[i] = 0; }
                    else {
                        const T Ta = (a + b) * tab[i].x * (T)0.25, Tb = (a - b) * h;
                        oa[i] = (Ta + Tb) * h; ob[i] = (Ta - Tb) * h;
                    }
                } else if constexpr (KIND == K_DCT_1) {   // dct1funcinplace.cu:116-127
                    if (n == 0) { oa[i] = a * h + x1; ob[i] = a * h - x1; }
                    else {
                        const T Ta = (a + b) * h, Tb = (a - b) * (T)0.25 * tab[i].x;
                        oa[i] = (Ta - Tb) * h; ob[i] = (Ta + Tb) * h;
                    }
                } else if constexpr (KIND == K_DCT_2) {   // dct2funcinplace.cu:93-105 (out[0] = y_0)
                    if (n == 0) { oa[i] = a; ob[i] = 0; }
                    else {
                        const V cs = tab[i];
                        const T Ta = a + b, Tb = a - b;
                        oa[i] = (Ta * cs.x + Tb * cs.y) * h; ob[i] = (Ta * cs.y - Tb * cs.x) * h;
                    }
                } else if constexpr (KIND == K_DST_2) {   // dst2funcinplace.cu:89-103 (out[0] = 0, out[N] = y_0)
                    if (n == 0) { oa[i] = 0; ob[i] = a; }
                    else {
                        const V cs = tab[i];
                        const T Ta = a + b, Tb = a - b;
                        oa[i] = (Ta * cs.y + Tb * cs.x) * h; ob[i] = (Ta * cs.x - Tb * cs.y) * h;
                    }
                } else {                                  // K_C2R
                    oa[i] = a; ob[i] = b;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PT; i++) {
            const int n = u + i * TPL;
            if (n <= M) {
                if (n == 0) {
                    tl.re(0) = oa[i];
                    if constexpr (KIND == K_DCT_1 || KIND == K_DST_2 || KIND == K_DST_1) tl.re(N) = ob[i];
                } else {
                    tl.re(n) = oa[i];
                    if (2 * n != N) tl.re(N - n) = ob[i];
                }
            }
        }
    } else {
        // D2Z split of the M-point FFT Z into the N-point spectrum X_k (k = 0..M):
        // E_k = (Z_k + conj Z_{M-k})/2, O_k = -i (Z_k - conj Z_{M-k})/2,
        // X_k = E_k + conj(w^k) O_k, X_{M-k} = conj(E_k - conj(w^k) O_k).
        V xa[PP], xb[PP], wk[PP];
#pragma unroll
        for (int i = 0; i < PP; i++) {
            const int k = u + i * TPL;
            wk[i] = k <= M / 2 ? tbs.wk(k) : mk<V>(0, 0);
        }
#pragma unroll
        for (int i = 0; i < PP; i++) {
            const int k = u + i * TPL;
            if (k <= M / 2) {
                const V Zk = tl.pt(k), Zm = tl.pt(k == 0 ? 0 : M - k);
                const V E = mk<V>(h * (Zk.x + Zm.x), h * (Zk.y - Zm.y));
                const V O = mk<V>(h * (Zk.y + Zm.y), -h * (Zk.x - Zm.x));
                const V w = wk[i];
                const V P = mk<V>(w.x * O.x + w.y * O.y, w.x * O.y - w.y * O.x);   // conj(w) O
                xa[i] = mk<V>(E.x + P.x, E.y + P.y);
                xb[i] = mk<V>(E.x - P.x, -(E.y - P.y));
            }
        }
        __syncthreads();
        auto put = [&](int t, V X) {
            if constexpr (KIND == K_R2C) {            // interleaved Re/Im, Im of DC and Nyquist exactly 0
                tl.re(2 * t) = X.x;
                tl.re(2 * t + 1) = (t == 0 || t == M) ? (T)0 : X.y;
            } else if constexpr (KIND == K_DCT_3) {   // dct3funcinplace.cu:83-91
                if (t == 0) { tl.re(0) = X.x * h; }
                else {
                    tl.re(2 * t - 1) = (X.x - X.y) * h;
                    if (2 * t < N) tl.re(2 * t) = (X.x + X.y) * h;
                }
            } else {                                  // K_DST_3, dst3funcinplace.cu:84-101
                if (t == 0) { tl.re(0) = 0; tl.re(1) = X.x * h; }
                else {
                    tl.re(2 * t) = -(X.y + X.x) * h;
                    if (2 * t + 1 <= N) tl.re(2 * t + 1) = (X.x - X.y) * h;
                }
            }
        };
#pragma unroll
        for (int i = 0; i < PP; i++) {
            const int k = u + i * TPL;
            if (k <= M / 2) {
                put(k, xa[i]);
                if (2 * k != M) put(M - k, xb[i]);
            }
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Pass description and the generic kernel body
// ---------------------------------------------------------------------------



// The pass helpers take the thread index as an optional argument (tid < 0:
// threadIdx.x) so a loop over tiles can hand them an index the compiler
// cannot hoist address arithmetic out of.
//
// Loads in flight per thread while staging a tile (the loads of a chunk are
// all issued before the first LDS write, so a thread waits once per chunk
// rather than once per element); fixed-shape kernels size it to the tile.
template <typename T> struct StageChunk { static constexpr int value = sizeof(T) == 8 ? 24 : 32; };

// Load the tile into LDS.  ROWS: lines are consecutive rows of a row-major
// volume (x-pass), the tile is one contiguous run of memory.  !ROWS: lines
// are L consecutive x-columns (y/z passes), slot stride s_sp.
template <typename T, bool ROWS, typename SRC, int CHX = 0>
__device__ __forceinline__ void load_tile(const PassArgs &p, const Tile<T> &tl, int nlines, long long lbase, const SRC *src,
                                          int tid = -1)
{
    constexpr int CH = CHX ? CHX : StageChunk<T>::value;
    const int t = tid < 0 ? (int)threadIdx.x : tid, nt = p.L * p.TPL;
    const int NS = p.N + 2;
    const int L = p.L;
    T *rbuf = reinterpret_cast<T *>(tl.buf);
    const int LS = tl.LS;
    if (ROWS) {
        // thread (row r, lane l) reads slots l, l + TPL, ... of row r: consecutive
        // lanes read consecutive doubles; every slot s < NS of every tile row is
        // written (0 outside the load window, beyond the row or the tile's lines).
        // The loads are unconditional at a clamped, valid index (the load window
        // lies inside the row) and masked afterwards.
        const int TPL = p.TPL;
        const int r = t / TPL, l = t - (t / TPL) * TPL;
        const int RL = p.s_rowlen;
        const bool rok = r < nlines;
        const SRC *rowp = src + lbase + (long long)(rok ? r : 0) * p.s_sa;
        const int lo = p.in_lo, hi = p.in_hi < RL - 1 ? p.in_hi : RL - 1;
        for (int s0 = l; s0 < NS; s0 += CH * TPL) {
            T v[CH];
#pragma unroll
            for (int i = 0; i < CH; i++) {
                const int s = s0 + i * TPL;
                const bool ok = rok && s >= lo && s <= hi;
                v[i] = (T)rowp[ok ? s : lo];
                if (!ok) v[i] = 0;
            }
#pragma unroll
            for (int i = 0; i < CH; i++) {
                const int s = s0 + i * TPL;
                if (s < NS) rbuf[(size_t)((s >> 1) * LS + r) * 2 + (s & 1)] = v[i];
            }
        }
        (void)nt;
        (void)L;
    } else {
        const int c = t % L, u0 = t / L, step = nt / L;
        const bool ok = c < nlines;
        const SRC *base = src + lbase + (long long)c * p.s_sa;
        for (int s0 = u0; s0 < NS; s0 += CH * step) {
            T v[CH];
#pragma unroll
            for (int i = 0; i < CH; i++) {
                const int s = s0 + i * step;
                v[i] = 0;
                if (ok && s < NS && s >= p.in_lo && s <= p.in_hi) v[i] = (T)base[(long long)s * p.s_sp];
            }
#pragma unroll
            for (int i = 0; i < CH; i++) {
                const int s = s0 + i * step;
                if (s < NS) rbuf[(size_t)((s >> 1) * LS + c) * 2 + (s & 1)] = v[i];
            }
        }
    }
}

template <typename T, bool ROWS, typename DST, int CHX = 0>
__device__ __forceinline__ void store_tile(const PassArgs &p, const Tile<T> &tl, int nlines, long long lbase, DST *dst,
                                           int tid = -1)
{
    constexpr int CH = CHX ? CHX : StageChunk<T>::value;
    const int t = tid < 0 ? (int)threadIdx.x : tid, nt = p.L * p.TPL;
    const T sc = (T)p.scale;
    const T *rbuf = reinterpret_cast<const T *>(tl.buf);
    if (ROWS) {
        // thread (row r, lane l) writes slots l, l + TPL, ... of row r in the store window
        const int TPL = p.TPL;
        const int r = t / TPL, l = t - (t / TPL) * TPL;
        if (r >= nlines) return;
        DST *rowp = dst + lbase + (long long)r * p.d_sa;
        const int lo = p.st_lo, hi = p.st_hi < p.d_rowlen - 1 ? p.st_hi : p.d_rowlen - 1;
        for (int s0 = lo + l; s0 <= hi; s0 += CH * TPL) {
            T v[CH];
#pragma unroll
            for (int i = 0; i < CH; i++) {
                const int s = s0 + i * TPL;
                const int sc_ = s <= hi ? s : hi;
                v[i] = rbuf[(size_t)((sc_ >> 1) * tl.LS + r) * 2 + (sc_ & 1)] * sc;
                if (s < p.out_lo || s > p.out_hi) v[i] = 0;
            }
#pragma unroll
            for (int i = 0; i < CH; i++) {
                const int s = s0 + i * TPL;
                if (s <= hi) rowp[s] = (DST)v[i];
            }
        }
        (void)nt;
    } else {
        const int L = p.L;
        const int c = t % L, u0 = t / L, step = nt / L;
        if (c >= nlines) return;
        DST *base = dst + lbase + (long long)c * p.d_sa;
        for (int s0 = p.st_lo + u0; s0 <= p.st_hi; s0 += CH * step) {
            T v[CH];
#pragma unroll
            for (int i = 0; i < CH; i++) {
                const int s = s0 + i * step;
                v[i] = 0;
                if (s <= p.st_hi && s >= p.out_lo && s <= p.out_hi) v[i] = rbuf[(size_t)((s >> 1) * tl.LS + c) * 2 + (s & 1)] * sc;
            }
#pragma unroll
            for (int i = 0; i < CH; i++) {
                const int s = s0 + i * step;
                if (s <= p.st_hi) base[(long long)s * p.d_sp] = (DST)v[i];
            }
        }
    }
}

// Tile of this workgroup.  The grid is 1-D over (tiles along a) x (lines b);
// workgroups are dispatched round-robin over the XCDs, each with its own L2,
// so the linear id is swizzled to hand every XCD a contiguous range of
// logical tiles: neighbouring tiles, which share cache lines at their edges,
// then run on the same XCD at about the same time.
__device__ __forceinline__ void tile_of(const PassArgs &p, int &a_tile, int &b)
{
    const int ntiles = (p.na + p.L - 1) / p.L;
    const int total = (int)gridDim.x;
    const int w = (int)blockIdx.x;
    int logical = w;
    const int X = p.xcds;
    if (X > 1 && total >= X) {
        const int per = (total + X - 1) / X;          // tiles of the first `tall` XCDs
        const int tall = total - (per - 1) * X;       // XCDs that get `per` tiles
        const int xcd = w % X, local = w / X;
        logical = xcd < tall ? xcd * per + local : tall * per + (xcd - tall) * (per - 1) + local;
    }
    if (p.bfast) {
        const int nb = total / ntiles;
        a_tile = p.a0 + (logical / nb) * p.L;
        b = p.b0 + logical % nb;
    } else {
        a_tile = p.a0 + (logical % ntiles) * p.L;
        b = p.b0 + logical / ntiles;
    }
}

// Copy the M-point FFT twiddles of the pass from global memory into LDS.
template <typename V>
__device__ __forceinline__ void stage_twiddles(V *twl, const V *gtw, int M, int nt, int tid = -1)
{
    for (int i = tid < 0 ? (int)threadIdx.x : tid; i < M; i += nt) twl[i] = gtw[i];
}

// Shared-memory carve-up of a pass: tile points, FFT twiddles, reduction scratch.
template <typename T> struct Smem {
    typedef typename vec2_of<T>::type V;
    V *buf;
    V *twl;
    T *red;
    __device__ __forceinline__ Smem(unsigned char *smem, int M, int L)
    {
        buf = reinterpret_cast<V *>(smem);
        twl = buf + lds_points(M, L, (int)sizeof(T));
        red = reinterpret_cast<T *>(twl + M);
    }
};

// One line transform of the whole tile, data in LDS: the runtime kind's fold,
// the FFT, the kind's unfold (the branches are uniform over the workgroup).
// The FFT, the bulk of the code, is emitted once for all the kinds rather than
// once per kind.  PLANC != 0: the radix plan as a compile-time constant
// (fixed-shape kernels).  C2R_AT_STORE: a C2R transform stops after the FFT,
// its unfold (a sign flip of the odd slots) left to the store.  tw: the pass's
// twiddle table in global memory, or the fold / unfold tables staged in LDS
// (LTabs); twl: the FFT twiddles staged in LDS.
template <int EPT, typename T, unsigned long long PLANC = 0, bool C2R_AT_STORE = false, typename TL, typename TB>
__device__ __forceinline__ void line_transform(int kind, const TL &tl, int u, int TPL, int M, int N,
                                               unsigned long long plan, const TB &tw,
                                               const typename vec2_of<T>::type *twl, T *red)
{
    switch (kind) {
        case K_DCT_1: fold<K_DCT_1, EPT, T>(tl, u, TPL, M, N, tw, red); break;
        case K_DCT_2: fold<K_DCT_2, EPT, T>(tl, u, TPL, M, N, tw, red); break;
        case K_DCT_3: fold<K_DCT_3, EPT, T>(tl, u, TPL, M, N, tw, red); break;
        case K_DST_1: fold<K_DST_1, EPT, T>(tl, u, TPL, M, N, tw, red); break;
        case K_DST_2: fold<K_DST_2, EPT, T>(tl, u, TPL, M, N, tw, red); break;
        case K_DST_3: fold<K_DST_3, EPT, T>(tl, u, TPL, M, N, tw, red); break;
        case K_R2C: fold<K_R2C, EPT, T>(tl, u, TPL, M, N, tw, red); break;
        default: fold<K_C2R, EPT, T>(tl, u, TPL, M, N, tw, red); break;
    }
    if constexpr (PLANC != 0) fft_forward_c<EPT, T, PLANC>(tl, u, TPL, M, twl);
    else fft_forward<EPT, T>(tl, u, TPL, M, plan, twl);
    switch (kind) {
        case K_DCT_1: unfold<K_DCT_1, EPT, T>(tl, u, TPL, M, N, tw, red); break;
        case K_DCT_2: unfold<K_DCT_2, EPT, T>(tl, u, TPL, M, N, tw, red); break;
        case K_DCT_3: unfold<K_DCT_3, EPT, T>(tl, u, TPL, M, N, tw, red); break;
        case K_DST_1: unfold<K_DST_1, EPT, T>(tl, u, TPL, M, N, tw, red); break;
        case K_DST_2: unfold<K_DST_2, EPT, T>(tl, u, TPL, M, N, tw, red); break;
        case K_DST_3: unfold<K_DST_3, EPT, T>(tl, u, TPL, M, N, tw, red); break;
        case K_R2C: unfold<K_R2C, EPT, T>(tl, u, TPL, M, N, tw, red); break;
        default:
            if constexpr (!C2R_AT_STORE) unfold<K_C2R, EPT, T>(tl, u, TPL, M, N, tw, red);
            break;
    }
}

}  // namespace chz
