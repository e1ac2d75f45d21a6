// chz_line.h -- fixed-shape pass kernels with lean tile I/O.
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

// LDS of a fixed-shape pass: the tile, the FFT twiddles and the reduction
// scratch, plus (fused z pass) the z eigenvalues and w^0..w^(M/2)
template <typename T, int N, int TPL, int L, int OP> constexpr long line_lds_core()
{
    return ((long)(N / 2 + 1) * L + N / 2) * 2 * (long)sizeof(T) + (long)TPL * L * (long)sizeof(T) +
           (OP == B_ZMID ? (long)((N + 2) * sizeof(double) + (N / 4 + 1) * 2 * sizeof(T)) : 0L);
}
// Waves per SIMD that LDS of lds bytes per tile leaves room for (160 KB per
// CU, at most 4): the register budget of the kernel follows from it.
template <int TPL, int L> constexpr int line_waves(long lds)
{
    const long wgs = 163840 / lds, w = wgs * (TPL * L / 64) / 4;
    return w < 1 ? 1 : w > 4 ? 4 : (int)w;
}
// LDS bytes of the fold / unfold tables (w^0..w^(M/2), unless the fused z
// pass holds them already, and the M + 1 entry n-table)
template <typename T, int N, int OP> constexpr long line_ftab_bytes()
{
    return (long)((OP == B_ZMID ? 0 : N / 4 + 1) + N / 2 + 1) * 2 * (long)sizeof(T);
}
// The fold / unfold tables are staged in LDS where that leaves the waves per
// SIMD unchanged and the tile within a CU's LDS (else they are read from the
// twiddle table in global memory)
template <typename T, int N, int TPL, int L, int OP> constexpr bool line_ftab_lds()
{
    return line_lds_core<T, N, TPL, L, OP>() + line_ftab_bytes<T, N, OP>() <= 163840 &&
           line_waves<TPL, L>(line_lds_core<T, N, TPL, L, OP>() + line_ftab_bytes<T, N, OP>()) ==
               line_waves<TPL, L>(line_lds_core<T, N, TPL, L, OP>());
}
// LDS bytes a kernel stages beyond the tile, FFT twiddles and reduction
// scratch (BT: the batched kernel, which reads the fold / unfold tables from
// global memory: its small tiles run more workgroups per CU than the wave
// bound above counts, and the tables measured slower there)
template <typename T, int N, int OP, int TPL, int L, bool BT = false> constexpr unsigned line_lds_extra()
{
    return (unsigned)(line_lds_core<T, N, TPL, L, OP>() - line_lds_core<T, N, TPL, L, B_ROWS>() +
                      (!BT && line_ftab_lds<T, N, TPL, L, OP>() ? line_ftab_bytes<T, N, OP>() : 0L));
}

namespace line {

// Line permutation of point row k for the row passes, chosen so that the 32
// lanes staging one row (16 consecutive points, both halves) hit 32 distinct
// bank slots and so that a chunk step (16 points) leaves it unchanged:
//   FP64 (16-byte points): L = 8 by (k >> 1) & 7, L = 16 by k & 15;
//   FP32 (8-byte points):  L = 8 by (k >> 2) & 3, L = 16 by (k >> 1) & 7.
// Column passes stage with lanes = lines, which is conflict-free unpermuted,
// and keep the plain layout (no index arithmetic in the FFT).
template <int L, bool SW, int ESZ> __device__ __forceinline__ constexpr int swz(int k)
{
    if constexpr (!SW) return 0;
    else if constexpr (ESZ == 8) return L == 8 ? (k >> 1) & 7 : k & (L - 1);
    else return L == 8 ? (k >> 2) & 3 : (k >> 1) & 7;
}

template <typename T, int LINES, bool SW> struct STile {
    typedef typename vec2_of<T>::type V;
    static constexpr int L = LINES;
    V *buf;
    int c;
    __device__ __forceinline__ int idx(int k) const { return k * L + (c ^ swz<L, SW, (int)sizeof(T)>(k)); }
    __device__ __forceinline__ V &pt(int k) const { return buf[idx(k)]; }
    __device__ __forceinline__ T &re(int s) const { return reinterpret_cast<T *>(buf)[2 * idx(s >> 1) + (s & 1)]; }
};

template <typename T, int N, int TPL, int LINES, bool SW> struct Io {
    typedef typename vec2_of<T>::type V;
    typedef STile<T, LINES, SW> TL;
    static constexpr int L = LINES, M = N / 2, NS = N + 2;
    static constexpr int NCH = (NS + TPL - 1) / TPL;   // slot chunks of TPL per line
    static_assert(!SW || TPL % 32 == 0, "swizzled tiles: threads per line a multiple of 32 (chunk steps keep the row swizzle)");
    static_assert(!SW || L == 8 || L == 16, "swizzled tiles: 8 or 16 lines");
    static_assert(TPL % 2 == 0, "threads per line: even (a chunk step keeps the slot parity)");
    // chunk i (slots i*TPL .. i*TPL + TPL - 1) lies inside every window
    static constexpr bool interior(int i) { return i * TPL >= 2 && i * TPL + TPL - 1 <= N - 2; }

    // LDS element of slot s0 + i*TPL of line c is the per-lane base plus
    // i * lds_step: s0 < TPL, TPL is even (a multiple of 32 when swizzled), so
    // a chunk step keeps the slot parity and moves the point index by TPL/2
    // (a multiple of 16), which leaves the row swizzle unchanged
    __device__ __forceinline__ static T *lds_base(V *buf, int c, int s0)
    {
        return reinterpret_cast<T *>(buf) + 2 * ((s0 >> 1) * L + (c ^ swz<L, SW, (int)sizeof(T)>(s0 >> 1))) + (s0 & 1);
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

    // refinement pass 1 of the mixed solve: FP64 residual r = f_fold - A U at
    // the slots of rows (j, k) of the unknown box (body_xres of the generic
    // pass), rounded to T.  Interior chunks load f, U and its four y / z
    // neighbour rows at per-lane row pointers plus constant offsets, and the x
    // neighbours at +-1 element; the x-boundary rules only enter the edge chunks.
    // U: the iterate, FP64 or (the first iterate of a mixed solve) FP32
    template <typename UT>
    __device__ __forceinline__ static void stage_xres(const SolveArgs &sa, const UT *U, V *buf, int nlines,
                                                      int a_tile, int k, int t)
    {
        // rows of 32+ lanes (a power of two): the x neighbours of the interior
        // chunks come from the adjacent lanes (below); narrower rows load them
        constexpr bool XSH = (TPL & (TPL - 1)) == 0 && TPL >= 32 && TPL <= 64;
        constexpr int RC = 2;   // chunks per load group (two groups in flight)
        const int r = t / TPL, l = t % TPL;
        const bool rok = r < nlines;
        const int j = a_tile + (rok ? r : 0);
        const long long fx = sa.NX + 1, fy = sa.NY + 1, fxy = fx * fy;
        const long long c0 = (long long)a_tile * fx + (long long)k * fxy;   // wave-uniform
        const unsigned rowoff = (unsigned)((rok ? r : 0) * (int)fx);
        const double *fr = sa.f + c0 + rowoff;                               // row starts
        const UT *ur = U + c0 + rowoff;
        const double *fl = fr + l;                                            // this lane's slot l
        const UT *ul = ur + l;
        // y / z neighbour rows of U: offsets and 0 / 1 weights (periodic wrap, Neumann mirror, Dirichlet 0)
        long long dm[2], dp[2];
        double wm[2], wp[2];
        const int q[2] = {j, k};
        const long long st[2] = {fx, fxy};
#pragma unroll
        for (int d = 0; d < 2; d++) {
            const int lo = sa.s[d + 1], hi = sa.e[d + 1], bc = sa.bc[d + 1];
            dm[d] = 0; dp[d] = 0; wm[d] = 0.0; wp[d] = 0.0;
            if (q[d] > lo) { dm[d] = -st[d]; wm[d] = 1.0; }
            else if (bc == 0) { dm[d] = (long long)(hi - lo) * st[d]; wm[d] = 1.0; }
            else if ((bc == 3 || bc == 4) && hi > lo) { dm[d] = st[d]; wm[d] = 1.0; }
            if (q[d] < hi) { dp[d] = st[d]; wp[d] = 1.0; }
            else if (bc == 0) { dp[d] = -(long long)(hi - lo) * st[d]; wp[d] = 1.0; }
            else if ((bc == 2 || bc == 3) && hi > lo) { dp[d] = -st[d]; wp[d] = 1.0; }
        }
        const UT *uym = ul + dm[0], *uyp = ul + dp[0], *uzm = ul + dm[1], *uzp = ul + dp[1];
        double xlo, xhi;
        RowFolds::xfolds(sa, j, k, xlo, xhi);
        const int xs = sa.s[0], xe = sa.e[0], bx = sa.bc[0];
        const int xm_end = bx == 0 ? xe : ((bx == 3 || bx == 4) && xe > xs ? xs + 1 : xs);
        const double xm_w = (bx == 0 || ((bx == 3 || bx == 4) && xe > xs)) ? 1.0 : 0.0;
        const int xp_end = bx == 0 ? xs : ((bx == 2 || bx == 3) && xe > xs ? xe - 1 : xe);
        const double xp_w = (bx == 0 || ((bx == 2 || bx == 3) && xe > xs)) ? 1.0 : 0.0;
        const double pert = sa.pertrb ? sa.pertrb[0] * sa.pert_scale : 0.0;
        const double lam = sa.lambda, ix = sa.ih2[0], iy = sa.ih2[1], iz = sa.ih2[2];
        T *lp = lds_base(buf, r, l);
        // XSH: U at this lane's slot of every chunk is loaded first (in U's
        // own type); in the interior chunks the x neighbours are the adjacent
        // lanes' values (the lanes of a row are one power-of-two segment of
        // the wave), across a chunk edge the neighbouring chunk's end lane
        // (every window contains slots 2..N-2 and xe >= N-1, so those values
        // are the plain U of the slot)
        UT u0a[XSH ? NCH : 1];
        if constexpr (XSH) {
            static_for<0, NCH>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                if constexpr (interior(i)) {
                    u0a[i] = ul[i * TPL];
                } else {
                    const int s = l + i * TPL;
                    u0a[i] = ur[(rok && s >= xs && s <= xe) ? s : xs];
                }
            });
        }
        // the residual of chunk i from its loaded values (f with folds, U and
        // its neighbours) into the tile
        auto finish = [&](auto ic, double fvi, double a, double m_, double p_, double uyi, double uzi) {
            constexpr int i = decltype(ic)::value;
            const int s = l + i * TPL;
            double acc = fvi - pert;
            bool ok = rok;
            if constexpr (!interior(i)) {
                acc += (s == xs ? xlo : 0.0) + (s == xe ? xhi : 0.0);
                m_ *= s > xs ? 1.0 : xm_w;
                p_ *= s < xe ? 1.0 : xp_w;
                ok = ok && s >= xs && s <= xe;
            }
            double au = lam * a + (m_ - 2 * a + p_) * ix;
            au += (uyi - 2 * a) * iy;
            au += (uzi - 2 * a) * iz;
            if (s < NS) lp[i * lds_step] = ok ? (T)(acc - au) : (T)0;
        };
        // every row without its y / z face folds: branch-free, software-
        // pipelined over chunk groups of RC (the loads of group g + 1 are
        // issued before the arithmetic of group g; a scheduling barrier after
        // each group keeps the compiler from hoisting every group's loads,
        // which would spill)
        {
            constexpr int NG = (NCH + RC - 1) / RC;
            double fq[2][RC];
            UT yq[2][RC][2], zq[2][RC][2], xq[2][RC][2], aq[2][RC];
            auto load = [&](auto gc) {
                constexpr int g = decltype(gc)::value, h = g & 1;
                static_for<0, RC>([&](auto ic) {
                    constexpr int ii = decltype(ic)::value, i = g * RC + ii;
                    if constexpr (i < NCH) {
                        if constexpr (interior(i)) {
                            fq[h][ii] = fl[i * TPL];
                            if constexpr (!XSH) {
                                aq[h][ii] = ul[i * TPL];
                                xq[h][ii][0] = ul[i * TPL - 1];
                                xq[h][ii][1] = ul[i * TPL + 1];
                            }
                            yq[h][ii][0] = uym[i * TPL];
                            yq[h][ii][1] = uyp[i * TPL];
                            zq[h][ii][0] = uzm[i * TPL];
                            zq[h][ii][1] = uzp[i * TPL];
                        } else {
                            const int s = l + i * TPL;
                            const int sc_ = (rok && s >= xs && s <= xe) ? s : xs;
                            const int sm = sc_ > xs ? sc_ - 1 : xm_end, sp = sc_ < xe ? sc_ + 1 : xp_end;
                            fq[h][ii] = fr[sc_];
                            if constexpr (!XSH) aq[h][ii] = ur[sc_];
                            xq[h][ii][0] = ur[sm];
                            xq[h][ii][1] = ur[sp];
                            yq[h][ii][0] = (ur + dm[0])[sc_];
                            yq[h][ii][1] = (ur + dp[0])[sc_];
                            zq[h][ii][0] = (ur + dm[1])[sc_];
                            zq[h][ii][1] = (ur + dp[1])[sc_];
                        }
                    }
                });
            };
            load(std::integral_constant<int, 0>());
            static_for<0, NG>([&](auto gc) {
                constexpr int g = decltype(gc)::value, h = g & 1;
                if constexpr (g + 1 < NG) load(std::integral_constant<int, g + 1>());
                static_for<0, RC>([&](auto ic) {
                    constexpr int ii = decltype(ic)::value, i = g * RC + ii;
                    if constexpr (i < NCH) {
                        double a, m_, p_;
                        if constexpr (interior(i) && XSH) {
                            a = (double)u0a[i];
                            const UT nx = __shfl(u0a[i], l + 1, TPL), nx1 = __shfl(u0a[i + 1], l + 1, TPL);
                            const UT pv = __shfl(u0a[i], l - 1, TPL), pv1 = __shfl(u0a[i - 1], l - 1, TPL);
                            p_ = (double)(l == TPL - 1 ? nx1 : nx);
                            m_ = (double)(l == 0 ? pv1 : pv);
                        } else {
                            if constexpr (XSH) a = (double)u0a[i];
                            else a = (double)aq[h][ii];
                            m_ = (double)xq[h][ii][0];
                            p_ = (double)xq[h][ii][1];
                        }
                        const double uyi = wm[0] * (double)yq[h][ii][0] + wp[0] * (double)yq[h][ii][1];
                        const double uzi = wm[1] * (double)zq[h][ii][0] + wp[1] * (double)zq[h][ii][1];
                        finish(std::integral_constant<int, i>(), fq[h][ii], a, m_, p_, uyi, uzi);
                    }
                });
                __builtin_amdgcn_sched_barrier(0);
            });
        }
        // rows on a y or z face of the unknown box (a few per plane, or the
        // whole plane at a z face): the residual again with the face folds
        // (loads from a face array or f), over the lane's own slots, in a
        // compact loop; the lane wrote these slots above, so no barrier
        if (RowFolds::has_faces(sa, j, k)) {
            const RowFolds rf(sa, j, k);
#pragma unroll 1
            for (int i = 0; i < NCH; i++) {
                const int s = l + i * TPL;
                if (s >= NS) break;
                const int sc_ = (rok && s >= xs && s <= xe) ? s : xs;
                const int sm = sc_ > xs ? sc_ - 1 : xm_end, sp = sc_ < xe ? sc_ + 1 : xp_end;
                double acc = fr[sc_] - pert;
#pragma unroll
                for (int m = 0; m < 4; m++)
                    if (rf.fc[m] != 0.0) acc += rf.fc[m] * rf.fp[m][sc_];
                const double a = ur[sc_];
                double m_ = ur[sm], p_ = ur[sp];
                acc += (s == xs ? xlo : 0.0) + (s == xe ? xhi : 0.0);
                m_ *= s > xs ? 1.0 : xm_w;
                p_ *= s < xe ? 1.0 : xp_w;
                const bool ok = rok && s >= xs && s <= xe;
                const double uyi = wm[0] * (double)(ur + dm[0])[sc_] + wp[0] * (double)(ur + dp[0])[sc_];
                const double uzi = wm[1] * (double)(ur + dm[1])[sc_] + wp[1] * (double)(ur + dp[1])[sc_];
                double au = lam * a + (m_ - 2 * a + p_) * ix;
                au += (uyi - 2 * a) * iy;
                au += (uzi - 2 * a) * iz;
                lp[i * lds_step] = ok ? (T)(acc - au) : (T)0;
            }
        }
    }

    // ---- store ---------------------------------------------------------------
    // last pass of a mixed-precision solve (body_xupd of the generic pass):
    // out = uin + scale * delta in FP64 over slots [st_lo, st_hi] of every row
    // (uin == nullptr: out = scale * delta), the iterates FP64 or FP32
    template <bool FULL, typename UIT, typename UOT>
    __device__ __forceinline__ static void store_rows_upd(const PassArgs &p, V *buf, int nlines, const UIT *uin,
                                                          UOT *out, long long obase, int t, bool negodd)
    {
        const int r = t / TPL, l = t % TPL;
        if (!FULL && r >= nlines) return;
        const double sc = (negodd && (l & 1)) ? -p.scale : p.scale;
        const int hi = p.st_hi < p.d_rowlen - 1 ? p.st_hi : p.d_rowlen - 1;
        const unsigned lane = (unsigned)(r * (int)p.d_sa + l);
        UOT *orow = out + obase + lane;
        const T *lp = lds_base(buf, r, l);
        // every slot of the iterate is loaded before the first store: out may
        // be uin (an in-place update; each lane reads and writes only its own
        // slots), and a store between the loads would make each chunk wait
        // for the previous one's store to complete
        double ui[NCH];
        if (uin) {
            const UIT *irow = uin + obase + lane;
#pragma unroll
            for (int i = 0; i < NCH; i++) {
                const int s = l + i * TPL;
                const bool ok = s < NS && (interior(i) || (s >= p.st_lo && s <= hi));
                ui[i] = ok ? (double)irow[i * TPL] : 0.0;
            }
        } else {
#pragma unroll
            for (int i = 0; i < NCH; i++) ui[i] = 0.0;
        }
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            const int s = l + i * TPL;
            if (s >= NS) continue;
            if (!interior(i) && !(s >= p.st_lo && s <= hi)) continue;
            orow[i * TPL] = (UOT)((double)lp[i * lds_step] * sc + ui[i]);
        }
    }
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
                                        int a_tile, int y, const double *__restrict__ ez)
{
    const int c = tl.c;
    const int x = a_tile + c;
    const double *ex = sa.eig, *ey = sa.eig + (sa.NX + 2);
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

// Exchange a double with the other lane of the pair (lanes 2a <-> 2a + 1):
// a DPP quad permutation [1, 0, 3, 2] of both 32-bit halves.
__device__ __forceinline__ double swap_pair(double v)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)bits, 0xB1, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(bits >> 32), 0xB1, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The fused z pass of a periodic z axis (R2C forward, divide, C2R inverse)
// for N = 512 in FP64 on tiles of 16 z-lines x 32 threads, radix plan
// 8 x 8 x 4.  The generic tile body moves every point through LDS between
// all phases; here the first stage of the forward FFT and the last stage of
// the inverse FFT stay in registers, which takes a quarter of the tile's LDS
// traffic (the staging write and its read back, the last stage's write and
// the store's read) and two barriers off the pass.  Arithmetic, operation
// order and results are those of the generic body.
//
// Thread t = l + 2 c + 32 h (l = t & 1) loads slots 2h + l + 32 i (i = 0..15)
// of line c: the lane pair (2h, 2h + 1) holds the real and imaginary parts
// of points h + 16 i, i.e. stage-1 butterflies h (points h + 32 r: i even) and
// h + 16 (i odd); one DPP swap of 8 values gives the even lane butterfly h
// and the odd lane butterfly h + 16, complex.  The last stage (radix 4,
// Ns = 64) runs butterflies h + 16 l and h + 16 l + 32 per lane, whose
// outputs are points h + 16 (l + 2 m); a second swap leaves each lane the
// slot parity it stores (the real parts on even slots, the imaginary parts,
// negated by the C2R unfold, on odd slots), at slots 2h + l + 32 i again.
template <bool FULL>
__device__ __forceinline__ void zmid_periodic_512(const PassArgs &p, const SolveArgs &sa, int a_tile, int b, int nlines,
                                                  unsigned char *smem, int t, bool tables)
{
    typedef double2 V;
    constexpr int N = 512, M = 256, L = 16, TPL = 32;
    V *buf = reinterpret_cast<V *>(smem);
    V *twl = buf + (M + 1) * L;
    double *red = reinterpret_cast<double *>(twl + M);
    double *ezl = red + TPL * L;
    V *wkl = reinterpret_cast<V *>(ezl + (N + 2));
    const V *tw = reinterpret_cast<const V *>(p.tw);
    if (tables) {
        stage_twiddles(twl, tw + TwOffsets::fftw(N), M, TPL * L, t);
        const double *ezg = sa.eig + (sa.NX + 2) + (sa.NY + 2);
        for (int i = t; i < N + 2; i += TPL * L) ezl[i] = ezg[i];
        for (int i = t; i <= M / 2; i += TPL * L) wkl[i] = tw[4 * i];
    }
    const int l = t & 1, cp = (t >> 1) & (L - 1), h = t >> 5;
    const bool cok = FULL || cp < nlines;
    // the R2C input window is slots 0..N-1: 16 chunks of 32 slots per line
    const double *src = reinterpret_cast<const double *>(p.src) + (long long)a_tile * p.s_sa + (long long)b * p.s_sb;
    const unsigned lane = (unsigned)(cp * (int)p.s_sa) + (unsigned)(2 * h + l) * (unsigned)p.s_sp;
    const long long step = (long long)TPL * p.s_sp;
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = cok ? (src + i * step)[lane] : 0.0;
    // forward stage 1 (radix 8, Ns = 1) on butterfly h + 16 l: outputs at 8 b + q
    {
        V x[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const double got = swap_pair(l ? v[2 * r] : v[2 * r + 1]);
            x[r] = l ? mk<V>(got, v[2 * r + 1]) : mk<V>(v[2 * r], got);
        }
        Dft<8>::run(x, 0, 1);
        const int bf = h + 16 * l;
#pragma unroll
        for (int q = 0; q < 8; q++) buf[(8 * bf + q) * L + cp] = x[q];
    }
    __syncthreads();
    // stages 2 and 3, the split / divide / combine and inverse stages 1 and 2
    // on the tile in LDS (line t % L, sub-thread t / L)
    const STile<double, L, false> tl{buf, t % L};
    const int u = t / L;
    fft_stage<8, 8>(tl, u, TPL, M, 8, twl);
    fft_stage<4, 8>(tl, u, TPL, M, 64, twl);
    periodic_split_divide<8, double, STile<double, L, false>, false, 1>(tl, u, TPL, M, nlines, a_tile, b, sa, wkl, ezl);
    fft_stage<8, 8>(tl, u, TPL, M, 1, twl);
    fft_stage<8, 8>(tl, u, TPL, M, 8, twl);
    // inverse stage 3 (radix 4, Ns = 64, twiddle stride 1) into registers
    V y[8];
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const int bb = h + 16 * l + 32 * e;   // j = bb (bb < Ns)
        V z[4];
#pragma unroll
        for (int r = 0; r < 4; r++) z[r] = buf[(bb + 64 * r) * L + cp];
#pragma unroll
        for (int r = 1; r < 4; r++) z[r] = cmul(z[r], twl[r * bb]);
        Dft<4>::run(z, 0, 1);
#pragma unroll
        for (int q = 0; q < 4; q++) y[e + 2 * q] = z[q];   // point bb + 64 q = h + 16 l + 32 (e + 2 q)
    }
    double o[16];
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const double got = swap_pair(l ? y[m].x : y[m].y);
        o[2 * m] = l ? got : y[m].x;
        o[2 * m + 1] = l ? y[m].y : got;
    }
    if (cok) {
        const double sc = l ? -p.scale : p.scale;   // C2R unfold: odd slots negated
        double *dst = reinterpret_cast<double *>(p.dst) + (long long)a_tile * p.d_sa + (long long)b * p.d_sb;
        const unsigned dl = (unsigned)(cp * (int)p.d_sa) + (unsigned)(2 * h + l) * (unsigned)p.d_sp;
        const long long dstep = (long long)TPL * p.d_sp;
#pragma unroll
        for (int i = 0; i < 16; i++) (dst + i * dstep)[dl] = o[i] * sc;
    }
}

// A wave-local fence for LDS data exchanged between the lanes of one wave:
// the wave executes its LDS instructions in order, so only the compiler has
// to be kept from moving the accesses across it.
__device__ __forceinline__ void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Lane (within a 16-lane line group) that holds the FFT outputs k = k1 + 16 k2
// of residue k1 after the first transpose: residues k1 and 16 - k1 (the
// pairs (k, M - k) of the real-FFT split) sit on lanes 2p and 2p + 1, and the
// self-paired residues 0 and 8 on lanes 0 and 1.
__device__ __forceinline__ int wz_lane_of(int k1) { return k1 == 0 ? 0 : k1 == 8 ? 1 : k1 < 8 ? 2 * k1 : 2 * (16 - k1) + 1; }
__device__ __forceinline__ int wz_res_of(int l) { return l == 0 ? 0 : l == 1 ? 8 : (l & 1) ? 16 - (l >> 1) : l >> 1; }

// The fused z pass of a periodic z axis (R2C forward, divide, C2R inverse)
// for N = 512 in FP64 with wave-resident transforms: a tile of 16 x-columns
// on 256 threads.  The tile is staged through LDS with whole 128-byte rows per
// slot (16 columns x 4 slots per load instruction), column-major (point k of
// column c at tb[257 c + k]); then each wave owns 4 columns, each z line
// (M = 256 complex points) held in 16 lanes x 16 registers, and both 256-point
// FFTs are four-step 16 x 16:
//   Z_{k1 + 16 k2} = sum_q W16^{q k2} W256^{q k1} sum_m z_{q + 16 m} W16^{m k1}
// (lane q: a 16-point DFT over its registers m, the twiddle, a transpose
// across the 16 lanes of the line through the line's own column of the tile,
// a second 16-point DFT).  The split / divide / combine pairs k with M - k,
// which the transpose puts on the lane pair (2p, 2p + 1) (wz_lane_of), so one
// DPP swap brings each lane its partner values.  The inverse FFT is the same
// four-step on the conj-combined values.  Two workgroup barriers in all
// (after staging, before the store); every other exchange stays in one wave.
// Arithmetic of the split / divide / combine as periodic_split_divide (for
// k > M/2 the same formulas with w^k = e^{2 pi i k / N} from the table).
template <bool FULL, bool S19>
__device__ __forceinline__ void zmid_periodic_512_wave(const PassArgs &p, const SolveArgs &sa, int a_tile, int b, int nlines,
                                                       unsigned char *smem, int t, bool tables)
{
    typedef double2 V;
    constexpr int N = 512, M = 256, CS = M + 1;
    // LDS: tile (16 columns x 257 points) | W256^j | w^k, k < M | z eigenvalues
    V *tb = reinterpret_cast<V *>(smem);
    V *twl = tb + 16 * CS;
    V *wkl = twl + M;
    double *ezl = reinterpret_cast<double *>(wkl + M);
    const V *tw = reinterpret_cast<const V *>(p.tw);
    if (tables) {
        stage_twiddles(twl, tw + TwOffsets::fftw(N), M, 256, t);
        for (int i = t; i < M; i += 256) wkl[i] = tw[4 * i];   // half[4k] = e^{2 pi i k / N}
        const double *ezg = sa.eig + (sa.NX + 2) + (sa.NY + 2);
        for (int i = t; i < N + 2; i += 256) ezl[i] = ezg[i];
    }
    // staging: thread (column sc, row sr) moves slots sr + 16 i, i = 0..31 (the R2C input window 0..N-1)
    const int sc = t & 15, sr = t >> 4;
    double *tr = reinterpret_cast<double *>(tb) + 2 * CS * sc;   // real slots of column sc
    {
        const bool ok = FULL || sc < nlines;
        const double *src = reinterpret_cast<const double *>(p.src) + (long long)a_tile * p.s_sa + (long long)b * p.s_sb;
        const unsigned off = (unsigned)(sc * (int)p.s_sa) + (unsigned)sr * (unsigned)p.s_sp;
        const long long st = 16LL * p.s_sp;
        double v[32];
#pragma unroll
        for (int i = 0; i < 32; i++) v[i] = ok ? (src + i * st)[off] : 0.0;
#pragma unroll
        for (int i = 0; i < 32; i++) tr[sr + 16 * i] = v[i];
    }
    __syncthreads();
    const int lane = t & 63, wv = t >> 6;
    const int cw = lane >> 4, l = lane & 15;   // line of the wave, lane within the line
    const int c = 4 * wv + cw;                 // column of the tile
    const bool cok = FULL || c < nlines;
    V *blk = tb + CS * c;                      // this line's column: also its transpose block
    V x[16];
    // lane l: points l + 16 m
#pragma unroll
    for (int m = 0; m < 16; m++) x[m] = blk[l + 16 * m];
    wave_lds_fence();
    // forward step 1: DFT over m, twiddle W256^{l k1}, transpose to lane wz_lane_of(k1), register l
    Dft<16>::run(x, 0, 1);
#pragma unroll
    for (int k1 = 1; k1 < 16; k1++) x[k1] = cmul(x[k1], twl[l * k1]);
#pragma unroll
    for (int k1 = 0; k1 < 16; k1++) {
        const int d = wz_lane_of(k1);
        blk[d * 16 + (l ^ d)] = x[k1];
    }
    wave_lds_fence();
#pragma unroll
    for (int j = 0; j < 16; j++) x[j] = blk[l * 16 + (j ^ l)];
    wave_lds_fence();
    // forward step 2: DFT over q -> Z_{k1 + 16 k2}, k1 = wz_res_of(l), in x[k2]
    Dft<16>::run(x, 0, 1);
    const int k1 = wz_res_of(l);
    // partner values Z_{M - k}: lanes 2..15 from the pair lane (register
    // 15 - k2 of the swapped copy), lane 1 (k1 = 8) from its own register
    // 15 - k2, lane 0 (k1 = 0) from its own register (16 - k2) mod 16
    // (selects on the scalar components: a select of the vector structs would
    // select their addresses and put the arrays in scratch memory)
    V zm[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const double sx = swap_pair(x[r].x), sy = swap_pair(x[r].y);
        zm[r].x = l >= 2 ? sx : x[r].x;
        zm[r].y = l >= 2 ? sy : x[r].y;
    }
    // split, divide, combine (periodic_split_divide's formulas for mode k = k1 + 16 k2)
    {
        const int xc = a_tile + c;
        const double *ex = sa.eig, *ey = sa.eig + (sa.NX + 2);
        const LineDivide<S19> dv(sa, cok ? ex[xc] : 0.0, cok ? ey[b] : 0.0);
        const bool m0xy = sa.singular && cok && (sa.bc[0] == 0 ? xc < 2 : xc == 0) && (sa.bc[1] == 0 ? b < 2 : b == 0);
        const double h = 0.5;
#pragma unroll
        for (int k2 = 0; k2 < 16; k2++) {
            const int k = k1 + 16 * k2;
            V Zm;
            Zm.x = l == 0 ? zm[(16 - k2) & 15].x : zm[15 - k2].x;
            Zm.y = l == 0 ? zm[(16 - k2) & 15].y : zm[15 - k2].y;
            const V Zk = x[k2], w = wkl[k];
            const V E = mk<V>(h * (Zk.x + Zm.x), h * (Zk.y - Zm.y));
            const V O = mk<V>(h * (Zk.y + Zm.y), -h * (Zk.x - Zm.x));
            const V P = mk<V>(w.x * O.x + w.y * O.y, w.x * O.y - w.y * O.x);
            V Xk = mk<V>(E.x + P.x, E.y + P.y), Xm = mk<V>(E.x - P.x, -(E.y - P.y));
            const double rk = (m0xy && k == 0) ? 0.0 : dv(ezl[2 * k]);
            const double rm = dv(ezl[2 * (M - k)]);
            Xk = mk<V>(Xk.x * rk, Xk.y * rk);
            Xm = mk<V>(Xm.x * rm, Xm.y * rm);
            if (k == 0) {
                Xk.y = 0;
                Xm.y = 0;
            }
            const V E2 = mk<V>(Xk.x + Xm.x, Xk.y - Xm.y);
            const V D2 = mk<V>(Xk.x - Xm.x, Xk.y + Xm.y);
            const V O2 = cmul(D2, w);
            x[k2] = mk<V>(E2.x - O2.y, -(E2.y + O2.x));
        }
    }
    // inverse step 1: DFT over k2, twiddle W256^{k1 n2}, transpose to lane n2, register k1
    Dft<16>::run(x, 0, 1);
#pragma unroll
    for (int n2 = 1; n2 < 16; n2++) x[n2] = cmul(x[n2], twl[k1 * n2]);
#pragma unroll
    for (int n2 = 0; n2 < 16; n2++) blk[n2 * 16 + (k1 ^ n2)] = x[n2];
    wave_lds_fence();
#pragma unroll
    for (int j = 0; j < 16; j++) x[j] = blk[l * 16 + (j ^ l)];
    wave_lds_fence();
    // inverse step 2: DFT over k1 -> Y_{n2 + 16 n1} (n2 = l) in x[n1], back into the column
    Dft<16>::run(x, 0, 1);
#pragma unroll
    for (int n1 = 0; n1 < 16; n1++) blk[l + 16 * n1] = x[n1];
    __syncthreads();
    // store: C2R unfold (x[2 n] = Re Y_n, x[2 n + 1] = -Im Y_n) times the ratio
    if (FULL || sc < nlines) {
        double *dst = reinterpret_cast<double *>(p.dst) + (long long)a_tile * p.d_sa + (long long)b * p.d_sb;
        const unsigned off = (unsigned)(sc * (int)p.d_sa) + (unsigned)sr * (unsigned)p.d_sp;
        const long long st = 16LL * p.d_sp;
        const double scl = (sr & 1) ? -p.scale : p.scale;   // slot parity = sr parity
#pragma unroll
        for (int i = 0; i < 32; i++) (dst + i * st)[off] = tr[sr + 16 * i] * scl;
    }
}

// LDS bytes of zmid_periodic_512_wave
constexpr unsigned zmid_wave_lds() { return (unsigned)((16 * 257 + 256 + 256) * 16 + 514 * 8); }

// Forward / inverse kind of the fused z pass for boundary code c (0..4).
__host__ __device__ constexpr int zmid_fwd_kind(int c) { return c == 0 ? K_R2C : c == 1 ? K_DST_1 : c == 2 ? K_DST_3 : c == 3 ? K_DCT_1 : K_DCT_3; }
__host__ __device__ constexpr int zmid_inv_kind(int c) { return c == 0 ? K_C2R : c == 1 ? K_DST_1 : c == 2 ? K_DST_2 : c == 3 ? K_DCT_1 : K_DCT_2; }

// One pass of the solve or the transform on a tile (op as BatchOp: rows,
// columns, solve pass 1, fused z pass).  KS < 0: the kind comes from the pass
// arguments (one kernel holds every kind's fold / unfold); KS >= 0: the kind
// is a compile-time constant (the pass kind, or for the fused z pass the z
// boundary code), so the kernel holds only that kind's code (a kernel with
// every kind inlined is several hundred KB, and where its hot path lands in
// the instruction cache changed its speed by half).
template <typename T, int OP, int N, int TPL, int L, unsigned long long PLAN, int EPT, bool FULL, int KS = -1,
          bool BT = false>
__device__ __forceinline__ void run(const PassArgs &p, const SolveArgs &sa, const double *u_in, double *u_out, int a_tile,
                                    int b, int nlines, unsigned char *smem, int t, bool tables = true)
{
    const int kind = KS < 0 ? p.kind : OP == B_ZMID ? zmid_fwd_kind(KS) : KS;
    const int kind2 = KS < 0 ? sa.kind2 : OP == B_ZMID ? zmid_inv_kind(KS) : sa.kind2;
    if constexpr (OP == B_ZMID && std::is_same<T, double>::value && N == 512 && TPL == 32 && L == 16 && EPT == 8 &&
                  PLAN == 0x040808ull) {
        // periodic z with the whole line in the windows: the register-staged form
        if (kind == K_R2C && kind2 == K_C2R && !sa.stencil19 && p.in_lo == 0 && p.in_hi == N - 1 &&
            p.st_lo == 0 && p.st_hi == N - 1 && sa.out2_lo == 0 && sa.out2_hi == N - 1) {
            zmid_periodic_512<FULL>(p, sa, a_tile, b, nlines, smem, t, tables);
            return;
        }
    }
    if constexpr (OP == B_ZMID && std::is_same<T, double>::value && N == 512 && TPL == 16 && L == 16 && EPT == 16 &&
                  PLAN == 0x1010ull) {
        // the tile's LDS (FP64 tile, twiddles, reduction scratch, z tables) holds the wave-resident form's
        static_assert(zmid_wave_lds() <= (unsigned)(((N / 2 + 1) * L + N / 2) * 16 + TPL * L * 8) +
                                             line_lds_extra<double, N, B_ZMID, TPL, L>(),
                      "LDS of the wave-resident fused z pass");
        if (kind == K_R2C && kind2 == K_C2R && p.in_lo == 0 && p.in_hi == N - 1 && p.st_lo == 0 &&
            p.st_hi == N - 1 && sa.out2_lo == 0 && sa.out2_hi == N - 1) {
            if (sa.stencil19) zmid_periodic_512_wave<FULL, true>(p, sa, a_tile, b, nlines, smem, t, tables);
            else zmid_periodic_512_wave<FULL, false>(p, sa, a_tile, b, nlines, smem, t, tables);
            return;
        }
    }
    // row tiles are swizzled where the chunk step allows it (32k threads per line, 8 or 16 lines)
    constexpr bool SW = OP != B_COLS && OP != B_ZMID && TPL % 32 == 0 && (L == 8 || L == 16);
    typedef Io<T, N, TPL, L, SW> IO;
    typedef typename IO::V V;
    typedef typename IO::TL TL;
    constexpr int M = N / 2;
    const int c = t % L, u = t / L;
    V *buf = reinterpret_cast<V *>(smem);
    V *twl = buf + (M + 1) * L;
    T *red = reinterpret_cast<T *>(twl + M);
    const V *tw = reinterpret_cast<const V *>(p.tw);
    const TL tl{buf, c};
    // the FFT twiddles and the z pass's tables (unchanged while a workgroup
    // stays on one problem: tables == false)
    if (tables) stage_twiddles(twl, tw + TwOffsets::fftw(N), M, TPL * L, t);
    // fused z pass: the z eigenvalues and w^0..w^(M/2) of the divide; with
    // FT the fold / unfold tables (w^0..w^(M/2) and the kind's n-table), all
    // staged with the tile (their loads overlap the tile's instead of
    // following a barrier in the middle of the transform)
    constexpr bool FT = !BT && line_ftab_lds<T, N, TPL, L, OP>();
    double *ezl = reinterpret_cast<double *>(red + TPL * L);
    V *wkl = reinterpret_cast<V *>(ezl + (OP == B_ZMID ? N + 2 : 0));
    V *ntl = wkl + (M / 2 + 1);
    if constexpr (OP == B_ZMID) if (tables) {
        const double *ezg = sa.eig + (sa.NX + 2) + (sa.NY + 2);
        for (int i = t; i < N + 2; i += TPL * L) ezl[i] = ezg[i];
    }
    if constexpr (OP == B_ZMID || FT) if (tables)
        for (int i = t; i <= M / 2; i += TPL * L) wkl[i] = tw[4 * i];
    // the n-table of the pass kind (the fused z pass: of its forward kind;
    // the inverse kind of every boundary code uses the same one)
    if constexpr (FT) if (tables && ntab_of(kind) != 0) {
        const V *src = ntab_of(kind) == 2 ? tw + TwOffsets::rsin(N) : tw;
        for (int i = t; i <= M; i += TPL * L) ntl[i] = src[i];
    }
    const LTabs<V> ltab{wkl, ntl};
    const T *src = reinterpret_cast<const T *>(p.src) + (long long)a_tile * p.s_sa + (long long)b * p.s_sb;
    if constexpr (OP == B_ROWS || OP == B_XUPD) IO::template stage_rows<FULL>(p, buf, nlines, src, t);
    if constexpr (OP == B_COLS || OP == B_ZMID) IO::template stage_cols<FULL>(p, buf, nlines, src, t);
    if constexpr (OP == B_XFWD) IO::stage_xfwd(sa, buf, nlines, a_tile, b, t);
    if constexpr (OP == B_XRES) {
        if (p.u_f32 & 1) IO::stage_xres(sa, reinterpret_cast<const float *>(u_in), buf, nlines, a_tile, b, t);
        else IO::stage_xres(sa, u_in, buf, nlines, a_tile, b, t);
    }
    __syncthreads();
    PassArgs q = p;
    // the radix plan as a compile-time constant for the FFT stages (PC), except
    // in fused z passes of 16 points per thread: unrolled there, the every-kind
    // kernels kept their stage arrays in scratch (616 bytes per lane) and ran
    // 2.5x slower, so those walk the plan at run time
    constexpr unsigned long long PC = (OP == B_ZMID && EPT >= 16) ? 0ull : PLAN;
    auto fwd = [&]() {
        if constexpr (PC != 0) fft_forward_c<EPT, T, PC>(tl, u, TPL, M, twl);
        else fft_forward<EPT, T>(tl, u, TPL, M, PLAN, twl);
    };
    // the last transform of the pass; a C2R one leaves its unfold (a sign
    // flip of the odd slots) to the store
    int last = kind;
    if constexpr (OP == B_ZMID) {
        last = kind2;
        if (kind == K_R2C && kind2 == K_C2R) {
            fwd();
            if (sa.stencil19) periodic_split_divide<EPT, T, TL, true, 1>(tl, u, TPL, M, nlines, a_tile, b, sa, wkl, ezl);
            else periodic_split_divide<EPT, T, TL, false, 1>(tl, u, TPL, M, nlines, a_tile, b, sa, wkl, ezl);
            fwd();
        } else {
            auto zt = [&](const auto &tb) {
                line_transform<EPT, T, PC>(kind, tl, u, TPL, M, N, PLAN, tb, twl, red);
                if (sa.stencil19) zdivide<T, true>(p, sa, tl, u, TPL, nlines, a_tile, b, ezl);
                else zdivide<T, false>(p, sa, tl, u, TPL, nlines, a_tile, b, ezl);
                __syncthreads();
                line_transform<EPT, T, PC, true>(last, tl, u, TPL, M, N, PLAN, tb, twl, red);
            };
            if constexpr (FT) zt(ltab);
            else zt(tw);
        }
        q.out_lo = sa.out2_lo;
        q.out_hi = sa.out2_hi;
    } else {
        auto xt = [&](const auto &tb) { line_transform<EPT, T, PC, true>(last, tl, u, TPL, M, N, PLAN, tb, twl, red); };
        if constexpr (FT) xt(ltab);
        else xt(tw);
    }
    const long long dbase = (long long)a_tile * p.d_sa + (long long)b * p.d_sb;
    T *dst = reinterpret_cast<T *>(p.dst) + dbase;
    if constexpr (OP == B_XUPD) {
        const bool neg = last == K_C2R;
        if (p.u_f32 & 2)   // the first iterate of a mixed solve, stored in FP32
            IO::template store_rows_upd<FULL>(q, buf, nlines, reinterpret_cast<const float *>(u_in),
                                              reinterpret_cast<float *>(u_out), dbase, t, neg);
        else if (p.u_f32 & 1)
            IO::template store_rows_upd<FULL>(q, buf, nlines, reinterpret_cast<const float *>(u_in), u_out, dbase, t, neg);
        else
            IO::template store_rows_upd<FULL>(q, buf, nlines, u_in, u_out, dbase, t, neg);
    }
    else if constexpr (OP == B_ROWS || OP == B_XFWD || OP == B_XRES)
        IO::template store_rows<FULL>(q, buf, nlines, dst, t, last == K_C2R);
    else IO::template store_cols<FULL>(q, buf, nlines, dst, t, last == K_C2R);
}

}  // namespace line

// Waves per SIMD that the LDS of one tile leaves room for
template <typename T, int N, int TPL, int L, int OP, bool BT = false> constexpr int line_occupancy()
{
    return line_waves<TPL, L>(line_lds_core<T, N, TPL, L, B_ROWS>() + (long)line_lds_extra<T, N, OP, TPL, L, BT>());
}

// Fixed-shape pass kernel with lean tile I/O (same arguments as k_fixed).
// Precondition (checked by the host): every window of the pass contains
// slots 2..N-2, which holds for every kind and solve pass.
template <typename T, int OP, int N, int TPL, int L, unsigned long long PLAN, int EPT, int KS = -1>
__global__ void __launch_bounds__(L *TPL, (line_occupancy<T, N, TPL, L, OP>())) k_line(PassArgs p, SolveArgs sa,
                                                                                  const double *u_in, double *u_out)
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
    const int t = (int)threadIdx.x;
    if (nlines == L) line::run<T, OP, N, TPL, L, PLAN, EPT, true, KS>(q, sa, u_in, u_out, a_tile, b, nlines, smem, t);
    else line::run<T, OP, N, TPL, L, PLAN, EPT, false, KS>(q, sa, u_in, u_out, a_tile, b, nlines, smem, t);
}

// Batched pass with lean tile I/O: the work-item loop of the batched passes
// (k_batch_fixed: items cut into 8 contiguous chunks, one per XCD; launched
// with one workgroup per item, a smaller grid walks several) around the tile
// body of k_line.  Same precondition on the windows of every problem of the
// launch.
template <typename T, int OP, int N, int TPL, int L, unsigned long long PLAN, int EPT, int KS = -1>
__global__ void __launch_bounds__(L *TPL, (line_occupancy<T, N, TPL, L, OP, true>()))
    k_line_batch(const BatchItem *__restrict__ items, int nitems, const BatchDesc *__restrict__ descs)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int G = (int)gridDim.x, bid = (int)blockIdx.x;
    int w, step, end;
    if (G >= 8 && G % 8 == 0) {
        const int per = (nitems + 7) / 8, x = bid % 8;
        w = x * per + bid / 8;
        step = G / 8;
        end = min(nitems, (x + 1) * per);
    } else {
        w = bid;
        step = G;
        end = nitems;
    }
    int prev = -1;   // problem of the previous item (its tables are in LDS)
    for (; w < end; w += step) {
        // a thread index the compiler cannot hoist address arithmetic out of
        // the item loop with (which would hold it in registers across items)
        int t = (int)threadIdx.x;
        asm volatile("" : "+v"(t));
        const BatchItem it = items[w];
        const bool tables = it.prob != prev;
        prev = it.prob;
        const BatchDesc *d = descs + it.prob;
        PassArgs q = d->p;
        q.N = N;
        q.M = N / 2;
        q.TPL = TPL;
        q.L = L;
        q.plan = PLAN;
        const int ntiles = (q.na + L - 1) / L;
        const int a_tile = q.a0 + (it.tile % ntiles) * L, b = q.b0 + it.tile / ntiles;
        const int nlines = min(L, q.a0 + q.na - a_tile);
        if (nlines == L)
            line::run<T, OP, N, TPL, L, PLAN, EPT, true, KS, true>(q, d->sa, d->u_in, d->u_out, a_tile, b, nlines, smem, t,
                                                                   tables);
        else
            line::run<T, OP, N, TPL, L, PLAN, EPT, false, KS, true>(q, d->sa, d->u_in, d->u_out, a_tile, b, nlines, smem, t,
                                                                    tables);
        __syncthreads();   // the next item reuses the tile memory
    }
}

}  // namespace chz
