// chz_axis_kernels.h -- __global__ kernel templates; instantiated per
// precision in the chz_k_*.hip translation units (compiled in parallel).
#pragma once
#include "chz_kernels.h"
#include "chz_pass.h"

namespace chz {

// One fused axis pass over a tile of L lines: load -> fold -> FFT -> unfold -> store.
template <typename T, typename SRC, typename DST, int EPT, bool ROWS, int CH = 0>
__device__ __forceinline__ void body_axis(const PassArgs &p, int a_tile, long long b, unsigned char *smem, int tid = -1)
{
    typedef typename vec2_of<T>::type V;
    const int L = p.L;
    const Smem<T> sm(smem, p.M, L);
    const int t0 = tid < 0 ? (int)threadIdx.x : tid;
    const Tile<T> tl{sm.buf, tile_ls(L, (int)sizeof(T)), t0 % L, L};
    const int u = t0 / L;
    const V *tw = reinterpret_cast<const V *>(p.tw);
    const int nlines = min(L, p.a0 + p.na - a_tile);
    stage_twiddles(sm.twl, tw + TwOffsets::fftw(p.N), p.M, p.L * p.TPL, t0);
    load_tile<T, ROWS, SRC, CH>(p, tl, nlines, (long long)a_tile * p.s_sa + b * p.s_sb, reinterpret_cast<const SRC *>(p.src), t0);
    __syncthreads();
    line_transform<EPT, T>(p.kind, tl, u, p.TPL, p.M, p.N, p.plan, tw, sm.twl, sm.red);
    store_tile<T, ROWS, DST, CH>(p, tl, nlines, (long long)a_tile * p.d_sa + b * p.d_sb, reinterpret_cast<DST *>(p.dst), t0);
}

template <typename T, typename SRC, typename DST, int EPT, bool ROWS>
__global__ void __launch_bounds__(CHZ_MAX_THREADS) k_axis_pass(PassArgs p)
{
    extern __shared__ __align__(16) unsigned char smem[];
    int a_tile, b;
    tile_of(p, a_tile, b);
    body_axis<T, SRC, DST, EPT, ROWS>(p, a_tile, b, smem);
}

// The boundary folds of one row (j, k) of the unknown box (the reference's
// folding, cudahelmholtz/_cuda_helmholtz_.inl:70-227, order x, y, z as in
// cuda_helmholtz_solver.cu:234-285), resolved once per row: the x folds are
// two constants (first / last unknown slot); the y and z folds are up to four
// (row pointer, coefficient) pairs read at the same slot s:
// Dirichlet  f[first] -= f[0]/h^2,   f[last] -= f[N]/h^2
// Neumann    f[first] += 2 g_l / h,  f[last] -= 2 g_r / h
struct RowFolds {
    const double *fp[4];
    double fc[4];
    double xlo, xhi;
    bool any;
    __device__ __forceinline__ RowFolds(const SolveArgs &sa, int j, int k)
    {
        const long long fx = sa.NX + 1, fy = sa.NY + 1, fxy = fx * fy;
        const double *frow = sa.f + j * fx + k * fxy;
        for (int m = 0; m < 4; m++) {
            fp[m] = frow;
            fc[m] = 0.0;
        }
        const int by = sa.bc[1], bz = sa.bc[2], bx = sa.bc[0];
        if (by != 0 && j == sa.s[1]) {
            if (by == 1 || by == 2) { fp[0] = sa.f + k * fxy; fc[0] = -sa.ih2[1]; }
            else { fp[0] = sa.bcb + k * fx; fc[0] = 2 * sa.ih[1]; }
        }
        if (by != 0 && j == sa.e[1]) {
            if (by == 1 || by == 4) { fp[1] = sa.f + sa.NY * fx + k * fxy; fc[1] = -sa.ih2[1]; }
            else { fp[1] = sa.bct + k * fx; fc[1] = -2 * sa.ih[1]; }
        }
        if (bz != 0 && k == sa.s[2]) {
            if (bz == 1 || bz == 2) { fp[2] = sa.f + j * fx; fc[2] = -sa.ih2[2]; }
            else { fp[2] = sa.bcf + j * fx; fc[2] = 2 * sa.ih[2]; }
        }
        if (bz != 0 && k == sa.e[2]) {
            if (bz == 1 || bz == 4) { fp[3] = sa.f + j * fx + sa.NZ * fxy; fc[3] = -sa.ih2[2]; }
            else { fp[3] = sa.bce + j * fx; fc[3] = -2 * sa.ih[2]; }
        }
        any = fc[0] != 0.0 || fc[1] != 0.0 || fc[2] != 0.0 || fc[3] != 0.0;
        xfolds(sa, j, k, xlo, xhi);
    }
    // the x-face terms alone (added at the first / last unknown slot of the row)
    __device__ __forceinline__ static void xfolds(const SolveArgs &sa, int j, int k, double &xlo, double &xhi)
    {
        const long long fx = sa.NX + 1, fy = sa.NY + 1;
        const double *frow = sa.f + j * fx + k * fx * fy;
        const int bx = sa.bc[0];
        xlo = 0.0;
        xhi = 0.0;
        if (bx == 1 || bx == 2) xlo = -frow[0] * sa.ih2[0];
        else if (bx == 3 || bx == 4) xlo = 2 * sa.bcl[k * fy + j] * sa.ih[0];
        if (bx == 1 || bx == 4) xhi = -frow[sa.NX] * sa.ih2[0];
        else if (bx == 2 || bx == 3) xhi = -2 * sa.bcr[k * fy + j] * sa.ih[0];
    }
    // does row (j, k) carry a y or z face fold (the value of any)
    __device__ __forceinline__ static bool has_faces(const SolveArgs &sa, int j, int k)
    {
        return (sa.bc[1] != 0 && (j == sa.s[1] || j == sa.e[1])) || (sa.bc[2] != 0 && (k == sa.s[2] || k == sa.e[2]));
    }
    // the folded rhs at slot s (xs, xe: the unknown x window) from the raw value v = f[row + s]
    __device__ __forceinline__ double apply(double v, int s, int xs, int xe) const
    {
        if (s == xs) v += xlo;
        if (s == xe) v += xhi;
        if (any) {
#pragma unroll
            for (int m = 0; m < 4; m++)
                if (fc[m] != 0.0) v += fc[m] * fp[m][s];
        }
        return v;
    }
};

// Solve pass 1: rows (j,k) of the unknown box of f (node layout), boundary
// folds, forward x transform, store into the padded spectral workspace.
template <typename T, int EPT, int CH = 0>
__device__ __forceinline__ void body_xfwd(const PassArgs &p, const SolveArgs &sa, int a_tile, int k, unsigned char *smem,
                                          int tid = -1)
{
    typedef typename vec2_of<T>::type V;
    const int L = p.L;
    const Smem<T> sm(smem, p.M, L);
    const int t0 = tid < 0 ? (int)threadIdx.x : tid;
    const Tile<T> tl{sm.buf, tile_ls(L, (int)sizeof(T)), t0 % L, L};
    const int u = t0 / L;
    const V *tw = reinterpret_cast<const V *>(p.tw);
    const int nlines = min(L, p.a0 + p.na - a_tile);
    stage_twiddles(sm.twl, tw + TwOffsets::fftw(p.N), p.M, p.L * p.TPL, t0);
    {
        // thread (row r, lane l) stages slots l, l + TPL, ... of row (a_tile + r, k)
        // of f with the boundary folds applied (0 outside the unknown x window)
        constexpr int FC = CH ? CH : StageChunk<T>::value;
        const int TPL = p.TPL, NS = p.N + 2;
        const int t = t0, r = t / TPL, l = t - (t / TPL) * TPL;
        const bool rok = r < nlines;
        const int j = a_tile + (rok ? r : 0);
        const RowFolds rf(sa, j, k);
        const double *frow = sa.f + (long long)j * (sa.NX + 1) + (long long)k * (sa.NX + 1) * (sa.NY + 1);
        const int xs = sa.s[0], xe = sa.e[0];
        T *rbuf = reinterpret_cast<T *>(tl.buf);
        for (int s0 = l; s0 < NS; s0 += FC * TPL) {
            // all raw loads first (one batch), then the folds: the y / z folds
            // exist on boundary rows only, the x folds at the window ends
            double raw[FC];
#pragma unroll
            for (int i = 0; i < FC; i++) {
                const int s = s0 + i * TPL;
                raw[i] = frow[(rok && s >= xs && s <= xe) ? s : xs];
            }
            if (rf.any) {
#pragma unroll
                for (int i = 0; i < FC; i++) {
                    const int s = s0 + i * TPL, sc_ = (s >= xs && s <= xe) ? s : xs;
#pragma unroll
                    for (int m = 0; m < 4; m++)
                        if (rf.fc[m] != 0.0) raw[i] += rf.fc[m] * rf.fp[m][sc_];
                }
            }
            T v[FC];
#pragma unroll
            for (int i = 0; i < FC; i++) {
                const int s = s0 + i * TPL;
                const bool ok = rok && s >= xs && s <= xe;
                const double x = raw[i] + (s == xs ? rf.xlo : 0.0) + (s == xe ? rf.xhi : 0.0);
                v[i] = ok ? (T)x : (T)0;
            }
#pragma unroll
            for (int i = 0; i < FC; i++) {
                const int s = s0 + i * TPL;
                if (s < NS) rbuf[(size_t)((s >> 1) * tl.LS + r) * 2 + (s & 1)] = v[i];
            }
        }
    }
    __syncthreads();
    line_transform<EPT, T>(p.kind, tl, u, p.TPL, p.M, p.N, p.plan, tw, sm.twl, sm.red);
    store_tile<T, true, T, CH>(p, tl, nlines, (long long)a_tile * p.d_sa + (long long)k * p.d_sb, reinterpret_cast<T *>(p.dst), t0);
}

template <typename T, int EPT>
__global__ void __launch_bounds__(CHZ_MAX_THREADS) k_solve_xfwd(PassArgs p, SolveArgs sa)
{
    extern __shared__ __align__(16) unsigned char smem[];
    int a_tile, k;
    tile_of(p, a_tile, k);
    body_xfwd<T, EPT>(p, sa, a_tile, k, smem);
}

// 1/d: hardware reciprocal estimate refined by two Newton steps
// r <- r + r (1 - d r) (about 1 ulp; the divide of the z pass)
__device__ __forceinline__ double recip(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    r = __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
    r = __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
    return r;
}

// Spectral divide factor of one z line (x, y): 1 / (e_x + e_y + e_z + lambda)
// for the 7-point operator; for the 19-point operator (the reference's
// _nonpoisson_devide_19points_, cudahelmholtz/_cuda_helmholtz_.inl:255-274,
// generalised to every boundary code through the axis eigenvalues e_a and to
// unequal spacings) B / (L19 + lambda) with
//   L19 = e_x + e_y + e_z + c_xy e_x e_y + c_xz e_x e_z + c_yz e_y e_z,
//   B   = 1 + b_x e_x + b_y e_y + b_z e_z,
// c_ab = (h_a^2 + h_b^2) / 12, b_a = h_a^2 / 12 (for h_a = h and
// e_a = (2 cos t_a - 2) / h^2 these are the reference's 2/3 (sum cos cos +
// sum cos) - 4 over h^2 and 1/2 + 1/6 sum cos).
// S19 is a compile-time choice (the callers branch once per tile), so the
// unrolled divide loops carry one reciprocal per value.
template <bool S19> struct LineDivide {
    double a, cz, b0, bz, lam;
    __device__ __forceinline__ LineDivide(const SolveArgs &sa, double ex, double ey)
    {
        lam = sa.lambda;
        if constexpr (!S19) {
            a = ex + ey;
            cz = 1.0;
            b0 = 1.0;
            bz = 0.0;
        } else {
            a = ex + ey + sa.c19[0] * ex * ey;
            cz = 1.0 + sa.c19[1] * ex + sa.c19[2] * ey;
            b0 = 1.0 + sa.c19[3] * ex + sa.c19[4] * ey;
            bz = sa.c19[5];
        }
    }
    __device__ __forceinline__ double operator()(double ez) const
    {
        if constexpr (!S19) return recip(a + ez + lam);
        else return (b0 + bz * ez) * recip(a + cz * ez + lam);
    }
};

// the divide of the fused z pass on the tile of z lines (x = a_tile + c, y)
template <typename T, bool S19 = false>
__device__ __forceinline__ void zmid_divide(const PassArgs &p, const SolveArgs &sa, const Tile<T> &tl, int u, int nlines,
                                            int a_tile, int y)
{
    const int c = tl.c;
    {
        const int x = a_tile + c;
        const double *ex = sa.eig, *ey = sa.eig + (sa.NX + 2), *ez = ey + (sa.NY + 2);
        const LineDivide<S19> dv(sa, c < nlines ? ex[x] : 0.0, c < nlines ? ey[y] : 0.0);
        const bool m0xy = sa.singular && c < nlines && (sa.bc[0] == 0 ? x < 2 : x == 0) && (sa.bc[1] == 0 ? y < 2 : y == 0);
        T *rbuf = reinterpret_cast<T *>(tl.buf);
#pragma unroll 4
        for (int s = p.out_lo + u; s <= p.out_hi; s += p.TPL) {
            T &v = rbuf[(size_t)((s >> 1) * tl.LS + c) * 2 + (s & 1)];
            const bool m0 = m0xy && (sa.bc[2] == 0 ? s < 2 : s == 0);
            const double r = m0 ? 0.0 : dv(ez[s]);
            v = (T)((double)v * r);
        }
    }
}

// Periodic z (R2C forward, C2R inverse): the split of the M-point FFT into
// the N-point half spectrum X_m (the R2C unfold), the divide of X_m by the
// eigenvalue sum of mode m, and the Z2D combine of the C2R fold all work on
// the point pairs (k, M - k), so they run as one pass: read Z_k, Z_{M-k},
// form X_k, X_{M-k}, scale, recombine into the (conjugated) inverse-FFT input.
// Same arithmetic as unfold<K_R2C> -> zmid_divide -> fold<K_C2R>.
// w^k is read as wsrc[WST * k] (the pass's global table: WST = 4, a per-tile
// copy of w^0..w^(M/2): WST = 1), the z eigenvalue of slot s as ezsrc[s]
// (null: the global table of sa).
template <int EPT, typename T, typename TL, bool S19 = false, int WST = 4>
__device__ __forceinline__ void periodic_split_divide(const TL &tl, int u, int TPL, int M, int nlines, int a_tile,
                                                      int y, const SolveArgs &sa,
                                                      const typename vec2_of<T>::type *__restrict__ wsrc,
                                                      const double *__restrict__ ezsrc = nullptr)
{
    typedef typename vec2_of<T>::type V;
    constexpr int PP = EPT / 2 + 1;
    const T h = (T)0.5;
    const int c = tl.c, x = a_tile + c;
    const double *ex = sa.eig, *ey = sa.eig + (sa.NX + 2);
    const double *ez = ezsrc ? ezsrc : ey + (sa.NY + 2);
    const LineDivide<S19> dv(sa, c < nlines ? ex[x] : 0.0, c < nlines ? ey[y] : 0.0);
    const bool m0xy = sa.singular && c < nlines && (sa.bc[0] == 0 ? x < 2 : x == 0) && (sa.bc[1] == 0 ? y < 2 : y == 0);
    V ca[PP], cb[PP], wk[PP];
#pragma unroll
    for (int i = 0; i < PP; i++) {
        const int k = u + i * TPL;
        wk[i] = k <= M / 2 ? wsrc[WST * k] : mk<V>(0, 0);   // w^k, w = e^{2 pi i / N}
    }
#pragma unroll
    for (int i = 0; i < PP; i++) {
        const int k = u + i * TPL;
        if (k <= M / 2) {
            const V Zk = tl.pt(k), Zm = tl.pt(k == 0 ? 0 : M - k);
            const V w = wk[i];
            // R2C split: X_k = E + conj(w^k) O, X_{M-k} = conj(E - conj(w^k) O)
            const V E = mk<V>(h * (Zk.x + Zm.x), h * (Zk.y - Zm.y));
            const V O = mk<V>(h * (Zk.y + Zm.y), -h * (Zk.x - Zm.x));
            const V P = mk<V>(w.x * O.x + w.y * O.y, w.x * O.y - w.y * O.x);
            V Xk = mk<V>(E.x + P.x, E.y + P.y), Xm = mk<V>(E.x - P.x, -(E.y - P.y));
            // divide by the eigenvalue sums of modes k and M - k (mode m = slots 2m, 2m + 1)
            const double rk = (m0xy && k == 0) ? 0.0 : dv(ez[2 * k]);
            const double rm = dv(ez[2 * (M - k)]);
            Xk = mk<V>((T)((double)Xk.x * rk), (T)((double)Xk.y * rk));
            Xm = mk<V>((T)((double)Xm.x * rm), (T)((double)Xm.y * rm));
            if (k == 0) {   // DC and Nyquist are real
                Xk.y = 0;
                Xm.y = 0;
            }
            // C2R combine: C_k = E' + i O', E' = H_k + conj H_{M-k}, O' = (H_k - conj H_{M-k}) w^k
            const V E2 = mk<V>(Xk.x + Xm.x, Xk.y - Xm.y);
            const V D2 = mk<V>(Xk.x - Xm.x, Xk.y + Xm.y);
            const V O2 = cmul(D2, w);
            ca[i] = mk<V>(E2.x - O2.y, -(E2.y + O2.x));
            cb[i] = mk<V>(E2.x + O2.y, E2.y - O2.x);
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
    __syncthreads();
}

// Solve pass 3: z lines (x, y) of the spectral box: forward z transform,
// multiply by 1/(eig_x + eig_y + eig_z + lambda) (the reference's
// _divide_ / _periodic_divide_, cudahelmholtz/_cuda_helmholtz_.inl:276-421),
// inverse z transform, all in LDS.
template <typename T, int EPT, int CH = 0>
__device__ __forceinline__ void body_zmid(const PassArgs &p, const SolveArgs &sa, int a_tile, int y, unsigned char *smem,
                                          int tid = -1)
{
    typedef typename vec2_of<T>::type V;
    const int L = p.L;
    const int t0 = tid < 0 ? (int)threadIdx.x : tid;
    const int c = t0 % L;
    const Smem<T> sm(smem, p.M, L);
    const Tile<T> tl{sm.buf, tile_ls(L, (int)sizeof(T)), c, L};
    const int u = t0 / L;
    const int nlines = min(L, p.a0 + p.na - a_tile);
    const V *tw = reinterpret_cast<const V *>(p.tw);
    stage_twiddles(sm.twl, tw + TwOffsets::fftw(p.N), p.M, p.L * p.TPL, t0);
    load_tile<T, false, T, CH>(p, tl, nlines, (long long)a_tile * p.s_sa + (long long)y * p.s_sb, reinterpret_cast<const T *>(p.src), t0);
    __syncthreads();
    if (p.kind == K_R2C && sa.kind2 == K_C2R) {
        fft_forward<EPT, T>(tl, u, p.TPL, p.M, p.plan, sm.twl);
        if (sa.stencil19) periodic_split_divide<EPT, T, Tile<T>, true>(tl, u, p.TPL, p.M, nlines, a_tile, y, sa, tw);
        else periodic_split_divide<EPT, T>(tl, u, p.TPL, p.M, nlines, a_tile, y, sa, tw);
        fft_forward<EPT, T>(tl, u, p.TPL, p.M, p.plan, sm.twl);
        unfold<K_C2R, EPT, T>(tl, u, p.TPL, p.M, p.N, tw, sm.red);
    } else {
        line_transform<EPT, T>(p.kind, tl, u, p.TPL, p.M, p.N, p.plan, tw, sm.twl, sm.red);
        if (sa.stencil19) zmid_divide<T, true>(p, sa, tl, u, nlines, a_tile, y);
        else zmid_divide<T>(p, sa, tl, u, nlines, a_tile, y);
        __syncthreads();
        line_transform<EPT, T>(sa.kind2, tl, u, p.TPL, p.M, p.N, p.plan, tw, sm.twl, sm.red);
    }
    PassArgs q = p;
    q.out_lo = sa.out2_lo;
    q.out_hi = sa.out2_hi;
    store_tile<T, false, T, CH>(q, tl, nlines, (long long)a_tile * p.d_sa + (long long)y * p.d_sb, reinterpret_cast<T *>(p.dst), t0);
}

template <typename T, int EPT>
__global__ void __launch_bounds__(CHZ_MAX_THREADS) k_solve_zmid(PassArgs p, SolveArgs sa)
{
    extern __shared__ __align__(16) unsigned char smem[];
    int a_tile, y;
    tile_of(p, a_tile, y);
    body_zmid<T, EPT>(p, sa, a_tile, y, smem);
}

// FP64 7-point residual r = f_fold - (Delta_h U + lambda U) at unknown node
// (s, j, k): the boundary folds of f (as in RowFolds) and the
// homogeneous operator (Dirichlet neighbours 0, Neumann ghost u_{-1} = u_1,
// periodic wrap), the discrete system of cuda_helmholtz_solver.
// U == nullptr: the folded rhs alone.  In the singular case (sa.pertrb set)
// the constant PERTRB (weighted mean of the folded rhs, the part of f outside
// the range of the operator) is removed, so the correction solves of a
// refinement only see a consistent rhs.
__device__ __forceinline__ double residual_at(const SolveArgs &sa, const double *__restrict__ U, int s, int j, int k)
{
    const long long fx = sa.NX + 1, fy = sa.NY + 1, fxy = fx * fy;
    const long long c = s + j * fx + k * fxy;
    double acc = sa.f[c];
    // folds, reference order x, y, z
    const double *row = sa.f + j * fx + k * fxy;
    if (sa.bc[0] != 0 && s == sa.s[0]) {
        if (sa.bc[0] == 1 || sa.bc[0] == 2) acc -= row[0] * sa.ih2[0];
        else acc += 2 * sa.bcl[k * fy + j] * sa.ih[0];
    }
    if (sa.bc[0] != 0 && s == sa.e[0]) {
        if (sa.bc[0] == 1 || sa.bc[0] == 4) acc -= row[sa.NX] * sa.ih2[0];
        else acc -= 2 * sa.bcr[k * fy + j] * sa.ih[0];
    }
    if (sa.bc[1] != 0 && j == sa.s[1]) {
        if (sa.bc[1] == 1 || sa.bc[1] == 2) acc -= sa.f[s + k * fxy] * sa.ih2[1];
        else acc += 2 * sa.bcb[k * fx + s] * sa.ih[1];
    }
    if (sa.bc[1] != 0 && j == sa.e[1]) {
        if (sa.bc[1] == 1 || sa.bc[1] == 4) acc -= sa.f[s + sa.NY * fx + k * fxy] * sa.ih2[1];
        else acc -= 2 * sa.bct[k * fx + s] * sa.ih[1];
    }
    if (sa.bc[2] != 0 && k == sa.s[2]) {
        if (sa.bc[2] == 1 || sa.bc[2] == 2) acc -= sa.f[s + j * fx] * sa.ih2[2];
        else acc += 2 * sa.bcf[j * fx + s] * sa.ih[2];
    }
    if (sa.bc[2] != 0 && k == sa.e[2]) {
        if (sa.bc[2] == 1 || sa.bc[2] == 4) acc -= sa.f[s + j * fx + sa.NZ * fxy] * sa.ih2[2];
        else acc -= 2 * sa.bce[j * fx + s] * sa.ih[2];
    }
    if (U == nullptr) return acc;
    if (sa.pertrb) acc -= sa.pertrb[0] * sa.pert_scale;
    // homogeneous operator
    const double u0 = U[c];
    const int q[3] = {s, j, k};
    const long long st[3] = {1, fx, fxy};
    double au = sa.lambda * u0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int lo = sa.s[a], hi = sa.e[a], bc = sa.bc[a];
        double um, up;
        if (q[a] > lo) um = U[c - st[a]];
        else um = bc == 0 ? U[c + (long long)(hi - lo) * st[a]] : (bc == 3 || bc == 4) ? (hi > lo ? U[c + st[a]] : 0.0) : 0.0;
        if (q[a] < hi) up = U[c + st[a]];
        else up = bc == 0 ? U[c - (long long)(hi - lo) * st[a]] : (bc == 2 || bc == 3) ? (hi > lo ? U[c - st[a]] : 0.0) : 0.0;
        au += (um - 2 * u0 + up) * sa.ih2[a];
    }
    return acc - au;
}

// Refinement pass 1: FP64 residual of the rows (j,k) of the unknown box for
// the current solution U, rounded to T, forward x transform -> workspace.
template <typename T, int EPT, int CH = 0>
__device__ __forceinline__ void body_xres(const PassArgs &p, const SolveArgs &sa, const double *U, int a_tile, int k,
                                          unsigned char *smem, int tid = -1)
{
    typedef typename vec2_of<T>::type V;
    const int L = p.L;
    const Smem<T> sm(smem, p.M, L);
    const int t0 = tid < 0 ? (int)threadIdx.x : tid;
    const Tile<T> tl{sm.buf, tile_ls(L, (int)sizeof(T)), t0 % L, L};
    const int u = t0 / L;
    const V *tw = reinterpret_cast<const V *>(p.tw);
    const int nlines = min(L, p.a0 + p.na - a_tile);
    stage_twiddles(sm.twl, tw + TwOffsets::fftw(p.N), p.M, p.L * p.TPL, t0);
    {
        // thread (row r, lane l) forms the residual at slots l, l + TPL, ... of
        // row (j, k) = (a_tile + r, k).  Everything that depends only on the row
        // -- the y / z neighbour rows of U with their boundary rules and the
        // y / z / x boundary folds -- is resolved once per thread; per slot only
        // the x direction remains.
        constexpr int RC = 4;   // slots in flight per thread (8 loads each)
        const int TPL = p.TPL;
        const int t = t0, r = t / TPL, l = t - (t / TPL) * TPL;
        const int NS = p.N + 2;
        T *rbuf = reinterpret_cast<T *>(tl.buf);
        const bool rok = r < nlines;
        const int j = a_tile + (rok ? r : 0);
        const long long fx = sa.NX + 1, fy = sa.NY + 1, fxy = fx * fy;
        const long long c0 = j * fx + k * fxy;
        const double *frow = sa.f + c0, *urow = U + c0;
        // neighbour rows along y (d = 1) and z (d = 2): offsets from the row and 0/1 weights
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
        const RowFolds rf(sa, j, k);
        const int xs = sa.s[0], xe = sa.e[0], bx = sa.bc[0];
        // x neighbours at the window ends: periodic wrap, Neumann reflection, Dirichlet 0
        const int xm_end = bx == 0 ? xe : ((bx == 3 || bx == 4) && xe > xs ? xs + 1 : xs);
        const double xm_w = (bx == 0 || ((bx == 3 || bx == 4) && xe > xs)) ? 1.0 : 0.0;
        const int xp_end = bx == 0 ? xs : ((bx == 2 || bx == 3) && xe > xs ? xe - 1 : xe);
        const double xp_w = (bx == 0 || ((bx == 2 || bx == 3) && xe > xs)) ? 1.0 : 0.0;
        const double pert = sa.pertrb ? sa.pertrb[0] * sa.pert_scale : 0.0;
        const double *uym = urow + dm[0], *uyp = urow + dp[0], *uzm = urow + dm[1], *uzp = urow + dp[1];
        for (int s0 = l; s0 < NS; s0 += RC * TPL) {
            // all loads of the chunk first (one batch), then the rare y / z
            // folds of boundary rows, then the arithmetic
            double fv[RC], u0[RC], um[RC], up[RC], uy[RC], uz[RC];
#pragma unroll
            for (int i = 0; i < RC; i++) {
                const int s = s0 + i * TPL;
                const int sc_ = (rok && s >= xs && s <= xe) ? s : xs;
                const int sm = sc_ > xs ? sc_ - 1 : xm_end, sp = sc_ < xe ? sc_ + 1 : xp_end;
                fv[i] = frow[sc_];
                u0[i] = urow[sc_];
                um[i] = urow[sm];
                up[i] = urow[sp];
                uy[i] = wm[0] * uym[sc_] + wp[0] * uyp[sc_];
                uz[i] = wm[1] * uzm[sc_] + wp[1] * uzp[sc_];
            }
            if (rf.any) {
#pragma unroll
                for (int i = 0; i < RC; i++) {
                    const int s = s0 + i * TPL, sc_ = (s >= xs && s <= xe) ? s : xs;
#pragma unroll
                    for (int m = 0; m < 4; m++)
                        if (rf.fc[m] != 0.0) fv[i] += rf.fc[m] * rf.fp[m][sc_];
                }
            }
            double v[RC];
#pragma unroll
            for (int i = 0; i < RC; i++) {
                const int s = s0 + i * TPL;
                const bool ok = rok && s >= xs && s <= xe;
                const double acc = fv[i] + (s == xs ? rf.xlo : 0.0) + (s == xe ? rf.xhi : 0.0) - pert;
                const double a = u0[i];
                const double m_ = um[i] * (s > xs ? 1.0 : xm_w), p_ = up[i] * (s < xe ? 1.0 : xp_w);
                double au = sa.lambda * a + (m_ - 2 * a + p_) * sa.ih2[0];
                au += (uy[i] - 2 * a) * sa.ih2[1];
                au += (uz[i] - 2 * a) * sa.ih2[2];
                v[i] = ok ? acc - au : 0.0;
            }
#pragma unroll
            for (int i = 0; i < RC; i++) {
                const int s = s0 + i * TPL;
                if (s < NS) rbuf[(size_t)((s >> 1) * tl.LS + r) * 2 + (s & 1)] = (T)v[i];
            }
        }
    }
    __syncthreads();
    line_transform<EPT, T>(p.kind, tl, u, p.TPL, p.M, p.N, p.plan, tw, sm.twl, sm.red);
    store_tile<T, true, T, CH>(p, tl, nlines, (long long)a_tile * p.d_sa + (long long)k * p.d_sb, reinterpret_cast<T *>(p.dst), t0);
}

template <typename T, int EPT>
__global__ void __launch_bounds__(CHZ_MAX_THREADS) k_solve_xres(PassArgs p, SolveArgs sa, const double *U)
{
    extern __shared__ __align__(16) unsigned char smem[];
    int a_tile, k;
    tile_of(p, a_tile, k);
    body_xres<T, EPT>(p, sa, U, a_tile, k, smem);
}

// Last pass of a mixed-precision solve: inverse x transform of the workspace
// rows, then out = Uin + ratio * delta in FP64 over the unknown x window
// (Uin == nullptr: out = ratio * delta).
template <typename T, int EPT, int CH = 0>
__device__ __forceinline__ void body_xupd(const PassArgs &p, const double *Uin, double *out, int a_tile, int k,
                                          unsigned char *smem, int tid = -1)
{
    typedef typename vec2_of<T>::type V;
    const int L = p.L;
    const Smem<T> sm(smem, p.M, L);
    const int t0 = tid < 0 ? (int)threadIdx.x : tid;
    const Tile<T> tl{sm.buf, tile_ls(L, (int)sizeof(T)), t0 % L, L};
    const int u = t0 / L;
    const V *tw = reinterpret_cast<const V *>(p.tw);
    const int nlines = min(L, p.a0 + p.na - a_tile);
    stage_twiddles(sm.twl, tw + TwOffsets::fftw(p.N), p.M, p.L * p.TPL, t0);
    load_tile<T, true, T, CH>(p, tl, nlines, (long long)a_tile * p.s_sa + (long long)k * p.s_sb, reinterpret_cast<const T *>(p.src), t0);
    __syncthreads();
    line_transform<EPT, T>(p.kind, tl, u, p.TPL, p.M, p.N, p.plan, tw, sm.twl, sm.red);
    {
        // thread (row r, lane l) updates slots st_lo + l, st_lo + l + TPL, ... of row r
        constexpr int UC = CH ? CH : 16;   // updates in flight per thread
        const int TPL = p.TPL;
        const int t = t0, r = t / TPL, l = t - (t / TPL) * TPL;
        if (r < nlines) {
            const long long rowb = (long long)(a_tile + r) * p.d_sa + (long long)k * p.d_sb;
            const int lo = p.st_lo, hi = p.st_hi < p.d_rowlen - 1 ? p.st_hi : p.d_rowlen - 1;
            const T *rbuf = reinterpret_cast<const T *>(tl.buf);
            const double sc = p.scale;
            for (int s0 = lo + l; s0 <= hi; s0 += UC * TPL) {
                double v[UC];
#pragma unroll
                for (int i = 0; i < UC; i++) {
                    const int s = s0 + i * TPL;
                    const int sc_ = s <= hi ? s : hi;
                    v[i] = (double)rbuf[(size_t)((sc_ >> 1) * tl.LS + r) * 2 + (sc_ & 1)] * sc;
                    if (Uin) v[i] += Uin[rowb + sc_];
                }
#pragma unroll
                for (int i = 0; i < UC; i++) {
                    const int s = s0 + i * TPL;
                    if (s <= hi) out[rowb + s] = v[i];
                }
            }
        }
    }
}

template <typename T, int EPT>
__global__ void __launch_bounds__(CHZ_MAX_THREADS) k_solve_xinv_upd(PassArgs p, const double *Uin, double *out)
{
    extern __shared__ __align__(16) unsigned char smem[];
    int a_tile, k;
    tile_of(p, a_tile, k);
    body_xupd<T, EPT>(p, Uin, out, a_tile, k, smem);
}

// Periodic duplicate node N <- node 0 along one axis over the whole face
// (cudahelmholtz/_cuda_helmholtz_.inl:58-68, 116-126, 173-183).
template <typename T>
__global__ void __launch_bounds__(256) k_periodic_copy(T *f, int NX, int NY, int NZ, int axis)
{
    const long long fx = NX + 1, fy = NY + 1;
    const int n1 = axis == 0 ? NY + 1 : NX + 1;
    const int n2 = axis == 2 ? NY + 1 : NZ + 1;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n1 * n2) return;
    const int a = (int)(idx % n1), b = (int)(idx / n1);
    if (axis == 0) f[NX + a * fx + b * fx * fy] = f[a * fx + b * fx * fy];
    else if (axis == 1) f[a + NY * fx + b * fx * fy] = f[a + b * fx * fy];
    else f[a + b * fx + NZ * fx * fy] = f[a + b * fx];
}

// ---------------------------------------------------------------------------
// Batched passes.  A batched solve runs, per pass of the solve, one launch per
// size class (problems sharing the cell count along the pass axis, hence the
// launch configuration).  The workgroups of a launch walk the class's list of
// work items (problem, tile), cut into 8 contiguous chunks, one per XCD
// (workgroup b runs on XCD b % 8), so the tiles of one problem share an L2;
// the host launches one workgroup per item (measured faster than a grid of
// resident workgroups walking several).  The pass arguments of every problem come
// from a descriptor table.
// ---------------------------------------------------------------------------
// Shape overrides of a batched pass: none (generic), or compile-time N,
// threads per line, lines and radix plan (fixed-shape batched kernels).
struct NoShape {
    static constexpr int CH = 0;
    __device__ __forceinline__ static void apply(PassArgs &) {}
};
template <int N, int TPL, int L, unsigned long long PLAN, typename T> struct FixShape {
    static constexpr int CH0 = (N + 16 + TPL - 1) / TPL;
    static constexpr int CH = CH0 < StageChunk<T>::value ? CH0 : StageChunk<T>::value;
    __device__ __forceinline__ static void apply(PassArgs &p)
    {
        p.N = N;
        p.M = N / 2;
        p.TPL = TPL;
        p.L = L;
        p.plan = PLAN;
    }
};

template <typename T, int EPT, int OP, typename S>
__device__ __forceinline__ void batch_loop(const BatchItem *__restrict__ items, int nitems,
                                           const BatchDesc *__restrict__ descs, unsigned char *smem)
{
    constexpr int CH = S::CH;
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
    for (; w < end; w += step) {
        // a thread index the compiler cannot hoist address arithmetic out of
        // the item loop with (which would hold it in registers across items)
        int tid = (int)threadIdx.x;
        asm volatile("" : "+v"(tid));
        const BatchItem it = items[w];
        const BatchDesc *d = descs + it.prob;
        PassArgs p = d->p;
        S::apply(p);
        const int ntiles = (p.na + p.L - 1) / p.L;
        const int a_tile = p.a0 + (it.tile % ntiles) * p.L, b = p.b0 + it.tile / ntiles;
        if constexpr (OP == B_ROWS) body_axis<T, T, T, EPT, true, CH>(p, a_tile, b, smem, tid);
        if constexpr (OP == B_COLS) body_axis<T, T, T, EPT, false, CH>(p, a_tile, b, smem, tid);
        if constexpr (OP == B_XFWD) body_xfwd<T, EPT, CH>(p, d->sa, a_tile, b, smem, tid);
        if constexpr (OP == B_ZMID) body_zmid<T, EPT, CH>(p, d->sa, a_tile, b, smem, tid);
        if constexpr (OP == B_XRES) body_xres<T, EPT, CH>(p, d->sa, d->u_in, a_tile, b, smem, tid);
        if constexpr (OP == B_XUPD) body_xupd<T, EPT, CH>(p, d->u_in, d->u_out, a_tile, b, smem, tid);
        __syncthreads();   // the next item reuses the tile memory
    }
}

template <typename T, int EPT, int OP>
__global__ void __launch_bounds__(CHZ_MAX_THREADS) k_batch(const BatchItem *__restrict__ items, int nitems,
                                                           const BatchDesc *__restrict__ descs)
{
    extern __shared__ __align__(16) unsigned char smem[];
    batch_loop<T, EPT, OP, NoShape>(items, nitems, descs, smem);
}

// fixed-shape batched pass (one size class with a compiled launch configuration)
template <typename T, int EPT, int OP, int N, int TPL, int L, unsigned long long PLAN>
__global__ void __launch_bounds__(L *TPL, 4) k_batch_fixed(const BatchItem *__restrict__ items, int nitems,
                                                           const BatchDesc *__restrict__ descs)
{
    extern __shared__ __align__(16) unsigned char smem[];
    batch_loop<T, EPT, OP, FixShape<N, TPL, L, PLAN, T>>(items, nitems, descs, smem);
}

// periodic duplicate copies of many problems: copy blockIdx.y, face element x
template <typename C>
__global__ void __launch_bounds__(256) k_batch_periodic(const C *__restrict__ tab)
{
    const C c = tab[blockIdx.y];
    const long long fx = c.NX + 1, fy = c.NY + 1;
    const int n1 = c.axis == 0 ? c.NY + 1 : c.NX + 1;
    const int n2 = c.axis == 2 ? c.NY + 1 : c.NZ + 1;
    double *f = c.f;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < (long long)n1 * n2;
         idx += (long long)gridDim.x * blockDim.x) {
        const int a = (int)(idx % n1), b = (int)(idx / n1);
        if (c.axis == 0) f[c.NX + a * fx + b * fx * fy] = f[a * fx + b * fx * fy];
        else if (c.axis == 1) f[a + c.NY * fx + b * fx * fy] = f[a + b * fx * fy];
        else f[a + b * fx + c.NZ * fx * fy] = f[a + b * fx];
    }
}

// ---------------------------------------------------------------------------
// Fixed-shape passes: the same bodies with the cell count, threads per line,
// lines per tile and radix plan as compile-time constants, so every index,
// stride and twiddle offset folds into immediates and the FFT stages unroll
// with constant spans (no integer division).  One kernel per (op, shape);
// the host uses it when a pass's launch configuration equals the shape.
// Registers are capped at 128 (at least 4 waves per SIMD, 16 per CU).
// ---------------------------------------------------------------------------
template <typename T, int EPT, int OP, int N, int TPL, int L, unsigned long long PLAN>
__global__ void __launch_bounds__(L *TPL, 4) k_fixed(PassArgs p, SolveArgs sa, const double *u_in,
                                                                     double *u_out)
{
    extern __shared__ __align__(16) unsigned char smem[];
    PassArgs q = p;
    q.N = N;
    q.M = N / 2;
    q.TPL = TPL;
    q.L = L;
    q.plan = PLAN;
    // values each thread stages per tile (rows up to 16 slots of padding)
    constexpr int CH0 = (N + 16 + TPL - 1) / TPL;
    constexpr int CH = CH0 < StageChunk<T>::value ? CH0 : StageChunk<T>::value;
    int a_tile, b;
    tile_of(q, a_tile, b);
    if constexpr (OP == B_ROWS) body_axis<T, T, T, EPT, true, CH>(q, a_tile, b, smem);
    if constexpr (OP == B_COLS) body_axis<T, T, T, EPT, false, CH>(q, a_tile, b, smem);
    if constexpr (OP == B_XFWD) body_xfwd<T, EPT, CH>(q, sa, a_tile, b, smem);
    if constexpr (OP == B_ZMID) body_zmid<T, EPT, CH>(q, sa, a_tile, b, smem);
    if constexpr (OP == B_XRES) body_xres<T, EPT, CH>(q, sa, u_in, a_tile, b, smem);
    if constexpr (OP == B_XUPD) body_xupd<T, EPT, CH>(q, u_in, u_out, a_tile, b, smem);
}

#define CHZ_EPT_SWITCH(ept, EXPR_TEMPLATE)  \
    switch (ept) {                          \
        case 4: return EXPR_TEMPLATE(4);    \
        case 5: return EXPR_TEMPLATE(5);    \
        case 6: return EXPR_TEMPLATE(6);    \
        case 8: return EXPR_TEMPLATE(8);    \
        case 10: return EXPR_TEMPLATE(10);  \
        case 12: return EXPR_TEMPLATE(12);  \
        case 16: return EXPR_TEMPLATE(16);  \
    }                                       \
    return nullptr;

}  // namespace chz
