// chz_line.h -- fixed-shape FP64 pass kernels with lean tile I/O (N = 512).
//
// The fold, FFT and unfold phases are those of the generic pass (chz_pass.h,
// chz_device.h).  What differs is how a tile moves between HBM and LDS:
//   * in the x passes the tile is point-major with the lines of each point
//     row permuted by an XOR swizzle (STile): the FFT phases (lanes = lines)
//     still read whole point rows, and the row staging (consecutive lanes =
//     consecutive slots of one line) spreads over all LDS banks instead of
//     hitting a few;
//   * every global access of a tile is a wave-uniform 64-bit base plus a
//     32-bit per-lane offset that is the same for all the elements a thread
//     moves, so the loads and stores need no per-element address arithmetic,
//     and the LDS side is a per-lane base plus a compile-time offset;
//   * the input / output windows only cut into the first and last slot chunks
//     of a line (every kind's window contains slots 2..N-2), so only those
//     chunks carry window checks; full tiles skip the line checks.
#pragma once
#include "chz_axis_kernels.h"

namespace chz {
namespace line {

// Line permutation of point row k for the row passes: L = 8 swaps by
// (k >> 1) & 7, so the 16 consecutive points of a row-staging instruction land
// in 16 distinct 16-byte bank groups; L = 16 rows are 256 bytes and permute by
// k & 15.  Column passes stage with lanes = lines, which is conflict-free
// unpermuted, and keep the plain layout (no index arithmetic in the FFT).
template <int L, bool SW> __device__ __forceinline__ constexpr int swz(int k)
{
    if constexpr (!SW) return 0;
    else if constexpr (L == 8) return (k >> 1) & 7;
    else return k & (L - 1);
}

template <typename T, int LINES, bool SW> struct STile {
    typedef typename vec2_of<T>::type V;
    static constexpr int L = LINES;
    V *buf;
    int c;
    __device__ __forceinline__ int idx(int k) const { return k * L + (c ^ swz<L, SW>(k)); }
    __device__ __forceinline__ V &pt(int k) const { return buf[idx(k)]; }
    __device__ __forceinline__ T &re(int s) const { return reinterpret_cast<T *>(buf)[2 * idx(s >> 1) + (s & 1)]; }
};

template <typename T, int N, int TPL, int LINES, bool SW> struct Io {
    typedef typename vec2_of<T>::type V;
    typedef STile<T, LINES, SW> TL;
    static constexpr int L = LINES, M = N / 2, NS = N + 2;
    static constexpr int NCH = (NS + TPL - 1) / TPL;   // slot chunks of TPL per line
    static_assert(!SW || TPL % 32 == 0, "swizzled tiles: threads per line a multiple of 32 (chunk steps keep the row swizzle)");
    static_assert(TPL % 2 == 0, "threads per line: even (a chunk step keeps the slot parity)");
    // chunk i (slots i*TPL .. i*TPL + TPL - 1) lies inside every window
    static constexpr bool interior(int i) { return i * TPL >= 2 && i * TPL + TPL - 1 <= N - 2; }

    // LDS element of slot s0 + i*TPL of line c is the per-lane base plus
    // i * lds_step: s0 < TPL, TPL is even (a multiple of 32 when swizzled), so
    // a chunk step keeps the slot parity and moves the point index by TPL/2
    // (a multiple of 16), which leaves the row swizzle unchanged
    __device__ __forceinline__ static T *lds_base(V *buf, int c, int s0)
    {
        return reinterpret_cast<T *>(buf) + 2 * ((s0 >> 1) * L + (c ^ swz<L, SW>(s0 >> 1))) + (s0 & 1);
    }
    static constexpr int lds_step = TPL * L;   // doubles per chunk

    // ---- staging -----------------------------------------------------------
    // rows: thread (row r, lane l) reads slots l, l + TPL, ... of row r
    template <bool FULL>
    __device__ __forceinline__ static void stage_rows(const PassArgs &p, V *buf, int nlines, const T *gbase, int t)
    {
        const int r = t / TPL, l = t % TPL;
        const bool rok = FULL || r < nlines;
        const int lo = p.in_lo, hi = p.in_hi < p.s_rowlen - 1 ? p.in_hi : p.s_rowlen - 1;
        const T *rowp = gbase + (unsigned)((rok ? r : 0) * (int)p.s_sa + l);
        T *lp = lds_base(buf, r, l);
        T v[NCH];
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            const int s = l + i * TPL;
            if (interior(i)) v[i] = rok ? rowp[i * TPL] : (T)0;
            else {
                const bool ok = rok && s >= lo && s <= hi;
                v[i] = ok ? rowp[i * TPL] : (T)0;
            }
        }
#pragma unroll
        for (int i = 0; i < NCH; i++)
            if (l + i * TPL < NS) lp[i * lds_step] = v[i];
    }
    // columns: thread (line c, u) reads slots u, u + TPL, ... of x-column c
    template <bool FULL>
    __device__ __forceinline__ static void stage_cols(const PassArgs &p, V *buf, int nlines, const T *gbase, int t)
    {
        const int c = t % L, u = t / L;
        const bool cok = FULL || c < nlines;
        const unsigned lane = (unsigned)(c * (int)p.s_sa) + (unsigned)u * (unsigned)p.s_sp;
        const long long step = (long long)TPL * p.s_sp;
        T *lp = lds_base(buf, c, u);
        T v[NCH];
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            const T *gi = gbase + i * step;   // wave-uniform
            const int s = u + i * TPL;
            if (interior(i)) v[i] = cok ? gi[lane] : (T)0;
            else {
                const bool ok = cok && s < NS && s >= p.in_lo && s <= p.in_hi;
                v[i] = ok ? gi[lane] : (T)0;
            }
        }
#pragma unroll
        for (int i = 0; i < NCH; i++)
            if (u + i * TPL < NS) lp[i * lds_step] = v[i];
    }
    // solve pass 1: rows (j, k) of f (node layout, x fastest) with the boundary
    // folds applied while staging (body_xfwd of the generic pass)
    __device__ __forceinline__ static void stage_xfwd(const SolveArgs &sa, V *buf, int nlines, int a_tile, int k, int t)
    {
        const int r = t / TPL, l = t % TPL;
        const bool rok = r < nlines;
        const int j = a_tile + (rok ? r : 0);
        const RowFolds rf(sa, j, k);
        const long long fx = sa.NX + 1;
        const double *fbase = sa.f + (long long)a_tile * fx + (long long)k * fx * (sa.NY + 1);   // wave-uniform
        const double *frow = fbase + (unsigned)((rok ? r : 0) * (int)fx + l);
        const int xs = sa.s[0], xe = sa.e[0];
        T *lp = lds_base(buf, r, l);
        double raw[NCH];
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            const int s = l + i * TPL;
            if (interior(i)) raw[i] = frow[i * TPL];
            else raw[i] = (s >= xs && s <= xe) ? frow[i * TPL] : 0.0;
        }
        if (rf.any) {
#pragma unroll
            for (int i = 0; i < NCH; i++) {
                const int s = l + i * TPL, sc_ = (s >= xs && s <= xe) ? s : xs;
#pragma unroll
                for (int m = 0; m < 4; m++)
                    if (rf.fc[m] != 0.0) raw[i] += rf.fc[m] * rf.fp[m][sc_];
            }
        }
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            const int s = l + i * TPL;
            double x = raw[i];
            if (!interior(i)) {
                x += (s == xs ? rf.xlo : 0.0) + (s == xe ? rf.xhi : 0.0);
                if (!(s >= xs && s <= xe)) x = 0.0;
            }
            if (!rok) x = 0.0;
            if (s < NS) lp[i * lds_step] = (T)x;
        }
    }

    // ---- store ---------------------------------------------------------------
    // slots [st_lo, st_hi] of every line, value * scale inside [out_lo, out_hi], 0 outside
    // negodd: the C2R unfold (odd slots negated: the real view of conj(FFT)),
    // applied here instead of in LDS
    template <bool FULL>
    __device__ __forceinline__ static void store_rows(const PassArgs &p, V *buf, int nlines, T *gbase, int t, bool negodd)
    {
        const int r = t / TPL, l = t % TPL;
        if (!FULL && r >= nlines) return;
        const T sc = (negodd && (l & 1)) ? -(T)p.scale : (T)p.scale;   // slot parity = l parity (TPL even)
        const int hi = p.st_hi < p.d_rowlen - 1 ? p.st_hi : p.d_rowlen - 1;
        T *rowp = gbase + (unsigned)(r * (int)p.d_sa + l);
        const T *lp = lds_base(buf, r, l);
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            const int s = l + i * TPL;
            if (s >= NS) continue;
            const T v = lp[i * lds_step] * sc;
            if (interior(i)) rowp[i * TPL] = v;
            else if (s >= p.st_lo && s <= hi) rowp[i * TPL] = (s >= p.out_lo && s <= p.out_hi) ? v : (T)0;
        }
    }
    template <bool FULL>
    __device__ __forceinline__ static void store_cols(const PassArgs &p, V *buf, int nlines, T *gbase, int t, bool negodd)
    {
        const int c = t % L, u = t / L;
        if (!FULL && c >= nlines) return;
        const T sc = (negodd && (u & 1)) ? -(T)p.scale : (T)p.scale;   // slot parity = u parity (TPL even)
        const unsigned lane = (unsigned)(c * (int)p.d_sa) + (unsigned)u * (unsigned)p.d_sp;
        const long long step = (long long)TPL * p.d_sp;
        const T *lp = lds_base(buf, c, u);
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            const int s = u + i * TPL;
            if (s >= NS) continue;
            T *gi = gbase + i * step;
            const T v = lp[i * lds_step] * sc;
            if (interior(i)) gi[lane] = v;
            else if (s >= p.st_lo && s <= p.st_hi) gi[lane] = (s >= p.out_lo && s <= p.out_hi) ? v : (T)0;
        }
    }
};

// the divide of the fused z pass (zmid_divide of the generic pass) on the swizzled tile
template <typename T, bool S19, typename TLT>
__device__ __forceinline__ void zdivide(const PassArgs &p, const SolveArgs &sa, const TLT &tl, int u, int TPL, int nlines,
                                        int a_tile, int y)
{
    const int c = tl.c;
    const int x = a_tile + c;
    const double *ex = sa.eig, *ey = sa.eig + (sa.NX + 2), *ez = ey + (sa.NY + 2);
    const LineDivide<S19> dv(sa, c < nlines ? ex[x] : 0.0, c < nlines ? ey[y] : 0.0);
    const bool m0xy = sa.singular && c < nlines && (sa.bc[0] == 0 ? x < 2 : x == 0) && (sa.bc[1] == 0 ? y < 2 : y == 0);
#pragma unroll 4
    for (int s = p.out_lo + u; s <= p.out_hi; s += TPL) {
        T &v = tl.re(s);
        const bool m0 = m0xy && (sa.bc[2] == 0 ? s < 2 : s == 0);
        const double r = m0 ? 0.0 : dv(ez[s]);
        v = (T)((double)v * r);
    }
}

// One pass of the solve or the transform on a tile (op as BatchOp: rows,
// columns, solve pass 1, fused z pass).
template <typename T, int OP, int N, int TPL, int L, unsigned long long PLAN, int EPT, bool FULL>
__device__ __forceinline__ void run(const PassArgs &p, const SolveArgs &sa, int a_tile, int b, int nlines, unsigned char *smem)
{
    typedef Io<T, N, TPL, L, OP == B_ROWS || OP == B_XFWD> IO;
    typedef typename IO::V V;
    typedef typename IO::TL TL;
    constexpr int M = N / 2;
    const int t = (int)threadIdx.x;
    const int c = t % L, u = t / L;
    V *buf = reinterpret_cast<V *>(smem);
    V *twl = buf + (M + 1) * L;
    T *red = reinterpret_cast<T *>(twl + M);
    const V *tw = reinterpret_cast<const V *>(p.tw);
    const TL tl{buf, c};
    stage_twiddles(twl, tw + TwOffsets::fftw(N), M, TPL * L, t);
    const T *src = reinterpret_cast<const T *>(p.src) + (long long)a_tile * p.s_sa + (long long)b * p.s_sb;
    if constexpr (OP == B_ROWS) IO::template stage_rows<FULL>(p, buf, nlines, src, t);
    if constexpr (OP == B_COLS || OP == B_ZMID) IO::template stage_cols<FULL>(p, buf, nlines, src, t);
    if constexpr (OP == B_XFWD) IO::stage_xfwd(sa, buf, nlines, a_tile, b, t);
    __syncthreads();
    PassArgs q = p;
    // the last transform of the pass; a C2R one leaves its unfold (a sign
    // flip of the odd slots) to the store
    int last = p.kind;
    if constexpr (OP == B_ZMID) {
        last = sa.kind2;
        if (p.kind == K_R2C && sa.kind2 == K_C2R) {
            fft_forward<EPT, T>(tl, u, TPL, M, PLAN, twl);
            if (sa.stencil19) periodic_split_divide<EPT, T, TL, true>(tl, u, TPL, M, nlines, a_tile, b, sa, tw);
            else periodic_split_divide<EPT, T, TL, false>(tl, u, TPL, M, nlines, a_tile, b, sa, tw);
            fft_forward<EPT, T>(tl, u, TPL, M, PLAN, twl);
        } else {
            line_transform<EPT, T>(p.kind, tl, u, TPL, M, N, PLAN, tw, twl, red);
            if (sa.stencil19) zdivide<T, true>(p, sa, tl, u, TPL, nlines, a_tile, b);
            else zdivide<T, false>(p, sa, tl, u, TPL, nlines, a_tile, b);
            __syncthreads();
            if (last == K_C2R) {
                fold<K_C2R, EPT, T>(tl, u, TPL, M, N, tw, red);
                fft_forward<EPT, T>(tl, u, TPL, M, PLAN, twl);
            } else {
                line_transform<EPT, T>(last, tl, u, TPL, M, N, PLAN, tw, twl, red);
            }
        }
        q.out_lo = sa.out2_lo;
        q.out_hi = sa.out2_hi;
    } else {
        if (last == K_C2R) {
            fold<K_C2R, EPT, T>(tl, u, TPL, M, N, tw, red);
            fft_forward<EPT, T>(tl, u, TPL, M, PLAN, twl);
        } else {
            line_transform<EPT, T>(last, tl, u, TPL, M, N, PLAN, tw, twl, red);
        }
    }
    T *dst = reinterpret_cast<T *>(p.dst) + (long long)a_tile * p.d_sa + (long long)b * p.d_sb;
    if constexpr (OP == B_ROWS || OP == B_XFWD) IO::template store_rows<FULL>(q, buf, nlines, dst, t, last == K_C2R);
    else IO::template store_cols<FULL>(q, buf, nlines, dst, t, last == K_C2R);
}

}  // namespace line

// Waves per SIMD that the LDS of one tile leaves room for (160 KB per CU, at
// most 4): the register budget of the kernel follows from it.
template <typename T, int N, int TPL, int L> constexpr int line_occupancy()
{
    constexpr long lds = ((long)(N / 2 + 1) * L + N / 2) * 2 * (long)sizeof(T) + (long)TPL * L * (long)sizeof(T);
    constexpr long wgs = 163840 / lds;
    constexpr long w = wgs * (TPL * L / 64) / 4;
    return w < 1 ? 1 : w > 4 ? 4 : (int)w;
}

// Fixed-shape pass kernel with lean tile I/O (same arguments as k_fixed).
// Precondition (checked by the host): every window of the pass contains
// slots 2..N-2, which holds for every kind and solve pass.
template <typename T, int OP, int N, int TPL, int L, unsigned long long PLAN, int EPT>
__global__ void __launch_bounds__(L *TPL, (line_occupancy<T, N, TPL, L>())) k_line(PassArgs p, SolveArgs sa,
                                                                                  const double *, double *)
{
    extern __shared__ __align__(16) unsigned char smem[];
    PassArgs q = p;
    q.N = N;
    q.M = N / 2;
    q.TPL = TPL;
    q.L = L;
    q.plan = PLAN;
    int a_tile, b;
    tile_of(q, a_tile, b);
    const int nlines = min(L, q.a0 + q.na - a_tile);
    if (nlines == L) line::run<T, OP, N, TPL, L, PLAN, EPT, true>(q, sa, a_tile, b, nlines, smem);
    else line::run<T, OP, N, TPL, L, PLAN, EPT, false>(q, sa, a_tile, b, nlines, smem);
}

}  // namespace chz
