// Folded line-operator passes for gfx950 (MI355X): the x/y spectral derivatives and the
// Fourier filter applied through the even/odd split of their circulant matrices.
//
// Reference: tevol_driven_cluster.m:60-61 (func_Lx / func_Ly: dense fourdif matrix times every
// line), src/numerics/ff.m:7-11 (filter).  Every line operator of the stepper is a circulant
// C[i][j] = c[(i - j) mod N] whose column is even (d_xx, the filter's real part: c[-k] = c[k])
// or odd (d_x, the filter's imaginary part: c[-k] = -c[k]).  With
//     f_j = x_j + x_{N-j},  g_j = x_j - x_{N-j}      (0 < j < N/2;  f_0 = x_0, f_{N/2} = x_{N/2}, g = 0 there)
// the line y = C x splits, for i = 0 .. N/2, into
//     A_i = sum_j (c[i-j] + c[i+j])/2 f_j,   B_i = sum_j (c[i-j] - c[i+j])/2 g_j      (j = 0 .. N/2)
//     even c:  y_i = A_i + B_i,  y_{N-i} = A_i - B_i;     odd c:  y_i = B_i + A_i,  y_{N-i} = B_i - A_i
// i.e. E (the even part of y) and O (its odd part) are each one product of a half-size matrix
// with half of the folded line: half the multiply-adds of the dense product.  The half-size
// matrices are Toeplitz + Hankel; their 16x4 fragments are formed in registers from two reads of
// one table tab[d + Z] = c[d mod N] / 2 (T part at d = i - j, H part at d = i + j).
//
// Each line enters the fold relative to its first value x_0 (f is formed from x - x_0; g does not
// change): every operator here annihilates constants except the filter's real part, whose row sum
// (1) times x_0 is added back at the output.  In FP32 this keeps the rounding of the summed
// Toeplitz + Hankel coefficients from acting on a line's mean (A_t is about mu along every line).
//
// Work decomposition (as the dense form): a work-group owns LB = 16 NT lines and one output
// group; each wave owns the position tiles role + 4 q (q < QF) of the half line and keeps the E
// and O accumulators of its tiles, so y_i and y_{N-i} are finished in registers.
#pragma once
#include <type_traits>

#include "hqt_line_core.h"

namespace hqt {

// t + s h with s = +-1 (exact), in the element type.
__device__ __forceinline__ double fadd_s(double s, double h, double t) { return __builtin_fma(s, h, t); }
__device__ __forceinline__ float fadd_s(float s, float h, float t) { return __builtin_fmaf(s, h, t); }

// Occupancy the folded kernels' register allocation targets (T: element type, R: type of the
// prefetch registers, double when a source of the pass is double).
template <typename T, typename R> constexpr int FP_WPE = sizeof(T) == 8 ? 2 : 3;

// Largest folded table (entries) for QF position tiles per wave: Z + 64 QF + Kp + 4 with
// Kp <= 64 QF + 8, Z = Kp + 4.
template <int QF> constexpr int fold_tlmax() { return 192 * QF + 32; }

template <typename T, int QF>
__device__ __forceinline__ void stage_ftables(const LPass& P, const LGroup& G, T* C, int TL, int tid) {
    constexpr int TPT = (fold_tlmax<QF>() + LP_THREADS - 1) / LP_THREADS;
    T tv[4][TPT];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (s < G.nslots) {
            const T* src = static_cast<const T*>(P.ops[G.slot_op[s]]);
#pragma unroll
            for (int r = 0; r < TPT; ++r) {
                const int t = tid + r * LP_THREADS;
                tv[s][r] = t < TL ? src[t] : (T)0;
            }
        }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (s < G.nslots) {
#pragma unroll
            for (int r = 0; r < TPT; ++r) {
                const int t = tid + r * LP_THREADS;
                if (t < TL) C[s * TL + t] = tv[s][r];
            }
        }
    }
}

// Rank-4 dots of the folded lines: D (lines x 4) = F Vf + G Vg on MFMA, the 4 waves splitting the
// k range; partial tiles are summed in wave order (independent of the lines per tile).
// Folded rank-4 table: Vf (Kp x 4) | Vg (Kp x 4) | UE (Mp x 4) | UO (Mp x 4), row-major.
template <typename T, int NT>
__device__ __forceinline__ void fold_lr_dots(const T* __restrict__ F, const T* __restrict__ G, int S, int Kp,
                                             const T* __restrict__ tab, T* DLp, T* dl, int wave, int lane) {
    typedef typename MF<T>::v4 v4;
    constexpr int LB = 16 * NT;
    const int ln = lane & 15, lk = lane >> 4;
    const T* vf = tab;
    const T* vg = tab + 4 * Kp;
#pragma unroll
    for (int u = 0; u < NT; ++u) {
        v4 acc = v4{0, 0, 0, 0};
        for (int j0 = 4 * wave; j0 < Kp; j0 += 16) {
            const T a = F[(16 * u + ln) * S + j0 + lk];
            const T b = ln < 4 ? vf[4 * (j0 + lk) + ln] : (T)0;
            acc = MF<T>::mma(a, b, acc);
        }
        for (int j0 = 4 * wave; j0 < Kp; j0 += 16) {
            const T a = G[(16 * u + ln) * S + j0 + lk];
            const T b = ln < 4 ? vg[4 * (j0 + lk) + ln] : (T)0;
            acc = MF<T>::mma(a, b, acc);
        }
        if (ln < 4) {
#pragma unroll
            for (int r = 0; r < 4; ++r) DLp[(wave * LB + 16 * u + MF<T>::row(lk, r)) * 4 + ln] = acc[r];
        }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < LB * 4; idx += LP_THREADS)
        dl[idx] = ((DLp[idx] + DLp[LB * 4 + idx]) + DLp[2 * LB * 4 + idx]) + DLp[3 * LB * 4 + idx];
}

// gmode 1 (x-lines): partials of G[r][s] = sum over the lines of a 16-line block in one z-slice of
// D[line][r] * V_y[y(line)][s], for the (at most two, ny >= 16) slices the block touches; written as
// doubles to gp[(block * 2 + slice offset) * 16 + 4 r + s].  Thread (4 r + s, line) forms one product,
// and each 16-lane group sums its lines in a fixed xor-butterfly order.
template <typename T, int NT>
__device__ __forceinline__ void fold_g_partials(const LPass& P, const LOut& O, const T* dl, int line0, int nl) {
    const T* vy = static_cast<const T*>(O.gtab);
    double* gp = const_cast<double*>(O.gp);
    const int tid = threadIdx.x, rs = tid >> 4, l = tid & 15, r = rs >> 2, s = rs & 3;
#pragma unroll
    for (int u = 0; u < NT; ++u) {
        const int blk = (line0 >> 4) + u;
        if (blk * 16 >= nl) break;
        const int z0 = (blk * 16) / P.ny, line = blk * 16 + l;
        double p0 = 0.0, p1 = 0.0;
        if (line < nl) {
            const int z = line / P.ny, y = line - z * P.ny;
            const double v = (double)dl[(16 * u + l) * 4 + r] * (double)vy[4 * y + s];
            if (z == z0) p0 = v;
            else p1 = v;
        }
#pragma unroll
        for (int off = 8; off >= 1; off >>= 1) {
            p0 += __shfl_xor(p0, off);
            p1 += __shfl_xor(p1, off);
        }
        if (l == 0) {
            gp[(blk * 2) * 16 + rs] = p0;
            gp[(blk * 2 + 1) * 16 + rs] = p1;
        }
    }
}

// gmode 2 (y-lines of slice z, x = x0 + l): G of the slice from the partials of the 16-line blocks
// that touch it, then dl[l][s] = sum_r U_x[x][r] G[r][s].  The partial sum is in a fixed order:
// thread (block slot b0, 4 r + s) adds blocks b0, b0 + 16, b0 + 32 of the slice, the 4 slots of a
// wave are combined by an xor butterfly and the 4 waves in wave order.  GL: 80 doubles of LDS.
// Ends with a barrier.
template <typename T, int NT>
__device__ __forceinline__ void fold_lr_from_g(const LPass& P, const LOut& O, T* dl, double* GL, int x0, int z) {
    constexpr int LB = 16 * NT;
    const int tid = threadIdx.x, rs = tid & 15, b0 = tid >> 4;
    const int blo = (z * P.ny) >> 4, bhi = ((z + 1) * P.ny - 1) >> 4;      // at most 33 blocks (ny <= 512)
    double q = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int blk = blo + b0 + 16 * k;
        if (blk <= bhi) q += O.gp[(blk * 2 + (z - (blk * 16) / P.ny)) * 16 + rs];
    }
    q += __shfl_xor(q, 16);
    q += __shfl_xor(q, 32);
    if ((tid & 63) < 16) GL[16 + (tid >> 6) * 16 + rs] = q;
    __syncthreads();
    if (tid < 16) GL[tid] = ((GL[16 + tid] + GL[32 + tid]) + GL[48 + tid]) + GL[64 + tid];
    __syncthreads();
    const T* ux = static_cast<const T*>(O.gtab);
    for (int idx = tid; idx < LB * 4; idx += LP_THREADS) {
        const int l = idx >> 2, s = idx & 3, x = x0 + l;
        double d = 0.0;
        if (x < P.nx) {
#pragma unroll
            for (int r = 0; r < 4; ++r) d += (double)ux[4 * x + r] * GL[4 * r + s];
        }
        dl[idx] = (T)d;
    }
    __syncthreads();
}

// Output of a finished line tile: y_i = E + O (+ addend) at i, then y_{N-i} = E - O (+ addend) at
// N - i, where E includes the line's constant term cst (row sums times reference values).  Each
// half loads all of its addends before its first store.  Returns whether a stored value is NaN;
// clears the accumulators.
template <typename T, int DIR, int QF, int NT>
__device__ __forceinline__ bool fold_out(const LPass& P, const LOut& O, typename MF<T>::v4 (&aE)[NT][QF],
                                         typename MF<T>::v4 (&aO)[NT][QF], const T (&cst)[NT][4], int line0, int x0,
                                         int z, int nl, int N, int role, int lane) {
    const int ln = lane & 15, lk = lane >> 4;
    const int nx = P.nx, rp = P.rp, pl = P.ps * z, Mh = N >> 1;
    // x-lines: element offsets of the 4 NT lines this lane stores (z = the z-slice of line0)
    int xb[NT][4];
#pragma unroll
    for (int u = 0; u < NT; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) xb[u][r] = DIR == 0 ? xline_base_at(P, line0, z, 16 * u + MF<T>::row(lk, r)) : 0;
    // element index of position i (direct) or N - i (mirror) of accumulator (u, q, r)
    auto idx = [&](int u, int q, int r, bool mirror) -> uint32_t {
        const int i0 = (role + 4 * q) * 16;
        int line, i;
        if (DIR == 0) {
            line = 16 * u + MF<T>::row(lk, r);      // offset from line0 until the bound check
            i = i0 + ln;
        } else {
            line = x0 + 16 * u + ln;
            i = i0 + MF<T>::row(lk, r);
        }
        const bool okl = DIR == 0 ? line0 + line < nl : line < nx;
        const bool ok = mirror ? (okl && i >= 1 && 2 * i < N) : (okl && i <= Mh);
        const int pos = mirror ? N - i : i;
        if (!ok) return Z_OOB;
        return DIR == 0 ? (uint32_t)(xb[u][r] + pos) : (uint32_t)(line + rp * pos + pl);
    };
    bool bad = false;
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        const bool mirror = half == 1;
        const T sO = mirror ? (T)-1 : (T)1;
        T add[NT][QF][4];
        auto load_add = [&](auto tag) {
            using S_ = decltype(tag);
            const Rs rs = rs_make(O.addend, fbytes<S_>(P));
#pragma unroll
            for (int u = 0; u < NT; ++u)
#pragma unroll
                for (int q = 0; q < QF; ++q)
#pragma unroll
                    for (int r = 0; r < 4; ++r) add[u][q][r] = (T)bldr<S_>(rs, idx(u, q, r, mirror));
        };
        if (!O.addend) {
#pragma unroll
            for (int u = 0; u < NT; ++u)
#pragma unroll
                for (int q = 0; q < QF; ++q)
#pragma unroll
                    for (int r = 0; r < 4; ++r) add[u][q][r] = 0;
        } else if (O.add_dt == DT_F64) {
            load_add(double{});
        } else {
            load_add(float{});
        }
        auto store = [&](auto tag) {
            using S_ = decltype(tag);
            const Rs rs = rs_make(O.out, fbytes<S_>(P));
#pragma unroll
            for (int u = 0; u < NT; ++u)
#pragma unroll
                for (int q = 0; q < QF; ++q)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const uint32_t e = idx(u, q, r, mirror);
                        const T v = ((aE[u][q][r] + cst[u][r]) + sO * aO[u][q][r]) + add[u][q][r];
                        bstr<S_>(rs, e, (S_)v);
                        bad |= e != Z_OOB && v != v;
                    }
        };
        if (O.out_dt == DT_F64) store(double{});
        else store(float{});
    }
#pragma unroll
    for (int u = 0; u < NT; ++u)
#pragma unroll
        for (int q = 0; q < QF; ++q) aE[u][q] = aO[u][q] = typename MF<T>::v4{0, 0, 0, 0};
    return bad;
}

// KIND tags the instantiation only (0: RHS stage, 1: constraint, 2: filter), as in k_line_pass.
// A work-group processes P.tpb consecutive line tiles.  Its work is a sequence of units, one per
// (line tile, source): a unit's lines are staged folded into LDS, then every output term on that
// source runs its k-loop; an output is written as soon as its last source is in.  The host only
// folds groups with one source, or two sources and one output, so one output's accumulators are
// live at a time.  The global loads of the next unit are issued right after a unit is staged and
// stay in flight in registers during the unit's MFMAs.
// The body of one folded pass for work-group (bx, by) of an nbx x nby grid.
template <typename T, int DIR, int QF, int KIND, int NT, typename R>
__device__ __forceinline__ void fold_pass_body(const LPass* __restrict__ Pp, const Members M, const int xcd_order,
                                               const int bx, const int by, const int nbx, const int nby) {
    constexpr int LB = 16 * NT;
    constexpr int JS = DIR == 0 ? 16 : LP_THREADS / LB;            // j step between a thread's elements
    constexpr int PFJ = ((64 * QF + 8) + JS - 1) / JS;             // folded positions per thread and line half
    constexpr int NH = DIR == 0 ? NT : 1;                          // line halves per thread (x-lines, LB = 32)
    // Block order (chosen per pass by the autotuner; it never changes a result).  xcd_order = 0:
    // the launch order, group-major.  xcd_order = 1: XCD-aware over the whole (tile, group) grid,
    // tile-major, so that one XCD's L2 serves a run of consecutive tiles (their lines share cache
    // lines at the edges, which matters for the strided y-lines) with all of their groups.
    const int lin0 = bx + nbx * by;
    int lin, bxl, gy;
    if (xcd_order) {
        lin = xcd_block(lin0, nbx * nby);
        bxl = lin / nby;
        gy = lin - bxl * nby;
    } else {
        lin = lin0;
        bxl = bx;
        gy = by;
    }
    const MemberBlock mb = member_block_at(M, bxl, nbx);
    const LPass& P = Pp[mb.m];
    if (gy >= P.ngroups) return;
    if (__builtin_nontemporal_load(&P.flags[0]) != 0) return;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int S = P.S, Kp = P.Kp, N = P.N, TL = P.TL, Z = P.Z, Mh = N >> 1;
    T* Vf = reinterpret_cast<T*>(smem_raw);
    T* Vg = Vf + LB * S;
    T* C = Vg + LB * S;
    const LGroup& G = P.g[gy];
    T* DL = C + G.nslots * TL;               // rank-4 dots: 2 outputs x LB lines x 4
    T* DLp = DL + 2 * LB * 4;                // rank-4 partial dots: 4 waves x LB lines x 4
    T* X0 = DLp + 4 * LB * 4;                // reference value x_0 of every line of each source: 2 x LB
    // the 4 x 4 cross-term matrix of the filter and its per-wave partials (gmode 2), 8-byte aligned after X0
    double* GL = reinterpret_cast<double*>(
        smem_raw + ((reinterpret_cast<unsigned char*>(X0 + 2 * LB) - smem_raw + 7) & ~(ptrdiff_t)7));
    const int tid = threadIdx.x;
    const int nl = P.ny * P.nz;
    const int tile0 = mb.bx * P.tpb;
    const int ntile = min(P.tpb, P.ntiles - tile0);
    const int nunit = ntile * G.ntiles;

    typedef typename MF<T>::v4 v4;
    v4 aE[NT][QF], aO[NT][QF];
#pragma unroll
    for (int u = 0; u < NT; ++u)
#pragma unroll
        for (int q = 0; q < QF; ++q) aE[u][q] = aO[u][q] = v4{0, 0, 0, 0};

    const int lane = tid & 63, ln = lane & 15, lk = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // Position tiles of this wave: role + 4 q, with the role rotated by the logical block index so
    // that the waves with one tile fewer (their last tile past the half line) fall on different
    // SIMDs in work-groups that share a CU (an XCD's 32 CUs take its blocks in turn, so blocks 1 and
    // 32 apart differ in rotation).  nq: tiles of the wave that hold output rows (QF or QF - 1).
    const int role = (wave + lin + (lin >> 5)) & 3;
    const int rt = (Mh + 1 + 15) >> 4;
    const int nq = role + 4 * (QF - 1) < rt ? QF : QF - 1;
    // staging map: x-lines: line tid/16 (+16 for the second half), positions tid%16 + 16 r;
    // y-lines: x offset tid%LB, positions tid/LB + (256/LB) r
    const int mrow = DIR == 0 ? (tid >> 4) : (tid % LB);
    const int jt = DIR == 0 ? (tid & 15) : (tid / LB);

    // Registers of the unit in flight: x_j, x_{N-j} and x_0 of the thread's lines in R (double when
    // any source of the pass is double, so the fold is formed in each source's own precision).
    R pa[NH][PFJ], pb[NH][PFJ], pref[NH];
    auto fetch = [&](int unit) {
        const int lt = tile0 + unit / G.ntiles, t = unit % G.ntiles;
        const void* src = G.src[t];
        const int dt = G.src_dt[t];
        int line0 = 0, z = 0, x0 = 0;    // x-lines: z = the z-slice of line0
        if (DIR == 0) {
            line0 = lt * LB;
            z = line0 / P.ny;
        } else {
            x0 = (lt % P.nxc) * LB;
            z = lt / P.nxc;
        }
        // x-lines: element offsets of this thread's NH lines
        int xb[NH];
#pragma unroll
        for (int h = 0; h < NH; ++h) xb[h] = DIR == 0 ? xline_base_at(P, line0, z, mrow + 16 * h) : 0;
        auto load = [&](auto tag) {
            using S_ = decltype(tag);
            const Rs rs = rs_make(src, fbytes<S_>(P));
#pragma unroll
            for (int h = 0; h < NH; ++h) {
                uint32_t e;
                if (DIR == 0) {
                    const int line = line0 + mrow + 16 * h;
                    e = line < nl ? (uint32_t)xb[h] : Z_OOB;
                } else {
                    const int x = x0 + mrow;
                    e = x < P.nx ? (uint32_t)(x + P.ps * z) : Z_OOB;
                }
                pref[h] = (R)bldr<S_>(rs, e);
            }
#pragma unroll
            for (int h = 0; h < NH; ++h)
#pragma unroll
                for (int r = 0; r < PFJ; ++r) {
                    const int j = jt + JS * r;
                    uint32_t e0, e1;
                    if (DIR == 0) {
                        const int line = line0 + mrow + 16 * h;
                        const bool okl = line < nl;
                        const int lb0 = xb[h];
                        e0 = (okl && j <= Mh) ? (uint32_t)(lb0 + j) : Z_OOB;
                        e1 = (okl && j >= 1 && 2 * j < N) ? (uint32_t)(lb0 + N - j) : Z_OOB;
                    } else {
                        const int x = x0 + mrow;
                        const bool okl = x < P.nx;
                        const int base = x + P.ps * z;
                        e0 = (okl && j <= Mh) ? (uint32_t)(base + P.rp * j) : Z_OOB;
                        e1 = (okl && j >= 1 && 2 * j < N) ? (uint32_t)(base + P.rp * (N - j)) : Z_OOB;
                    }
                    pa[h][r] = (R)bldr<S_>(rs, e0);
                    pb[h][r] = (R)bldr<S_>(rs, e1);
                }
        };
        if (dt == DT_F64) load(double{});
        else load(float{});
    };
    // f = (x_j - x_0) + (x_{N-j} - x_0), g = x_j - x_{N-j} into LDS (zero past N/2 and where
    // x_{N-j} is x_j itself), formed in R.
    auto deposit = [&](int t) {
#pragma unroll
        for (int h = 0; h < NH; ++h)
#pragma unroll
            for (int r = 0; r < PFJ; ++r) {
                const int j = jt + JS * r;
                const bool pair = j >= 1 && 2 * j < N;
                const R da = j <= Mh ? pa[h][r] - pref[h] : (R)0;
                const R db = pair ? pb[h][r] - pref[h] : (R)0;
                if (j < Kp) {
                    Vf[(mrow + 16 * h) * S + j] = (T)(da + db);
                    Vg[(mrow + 16 * h) * S + j] = pair ? (T)(pa[h][r] - pb[h][r]) : (T)0;
                }
            }
        if (jt == 0) {
#pragma unroll
            for (int h = 0; h < NH; ++h) X0[t * LB + mrow + 16 * h] = (T)pref[h];
        }
    };

    // the first unit's line loads and the operator tables are in flight together
    if (nunit > 0) fetch(0);
    stage_ftables<T, QF>(P, G, C, TL, tid);
    for (int unit = 0; unit < nunit; ++unit) {
        const int lt = tile0 + unit / G.ntiles, t = unit % G.ntiles;
        int line0 = 0, z = 0, x0 = 0;    // x-lines: z = the z-slice of line0
        if (DIR == 0) {
            line0 = lt * LB;
            z = line0 / P.ny;
        } else {
            x0 = (lt % P.nxc) * LB;
            z = lt / P.nxc;
        }
        __syncthreads();                 // the previous unit's MFMAs and outputs are done with LDS
        deposit(t);
        __syncthreads();
        if (unit + 1 < nunit) fetch(unit + 1);
        if constexpr (KIND == 2) {
            if (G.has_lr) {
                for (int o = 0; o < G.nout; ++o)
                    for (int e = 0; e < G.out[o].nterms; ++e) {
                        const LTerm& tm = G.out[o].term[e];
                        if (tm.lr && tm.tile == t) {
                            if (DIR == 1 && G.out[o].gmode == 2) {
                                fold_lr_from_g<T, NT>(P, G.out[o], DL + o * LB * 4, GL, x0, z);
                            } else {
                                fold_lr_dots<T, NT>(Vf, Vg, S, Kp, static_cast<const T*>(P.ops[tm.slot]), DLp,
                                                    DL + o * LB * 4, wave, lane);
                                __syncthreads();
                                // the x-pass's rank-4 output feeds the filter's cross term: partials of G
                                // (formed here, before any accumulator is live) instead of a field
                                if (DIR == 0 && G.out[o].gmode == 1) fold_g_partials<T, NT>(P, G.out[o], DL + o * LB * 4, line0, nl);
                            }
                        }
                    }
            }
        }
        for (int o = 0; o < G.nout; ++o) {
            const LOut& O = G.out[o];
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                if (e >= O.nterms || O.term[e].tile != t || O.term[e].lr) continue;
                const int sym = O.term[e].pad;              // +1: even column, -1: odd column
                // T part tab[i - j + Z], H part tab[i + j + Z]; lane (ln, lk) reads i = i0 + ln, j = j0 + lk
                const T* Ct = C + O.term[e].slot * TL + Z + ln - lk + role * 16;
                const T* Ch = C + O.term[e].slot * TL + Z + ln + lk + role * 16;
                // E gets (T + sg H) times f (even column) or g (odd column), O the other pair
                const T* Be = (sym > 0 ? Vf : Vg) + ln * S + lk;
                const T* Bo = (sym > 0 ? Vg : Vf) + ln * S + lk;
                const T sg = sym > 0 ? (T)1 : (T)-1;
                // the k-loop over the wave's first NQ position tiles
                auto kloop = [&](auto nq_c) {
                    constexpr int NQ = decltype(nq_c)::value;
                    if constexpr (NQ > 0) {
                        T ae0[NT], ao0[NT], ae1[NT], ao1[NT];
                        T t0[NQ], h0[NQ], t1[NQ], h1[NQ];
#pragma unroll
                        for (int u = 0; u < NT; ++u) { ae0[u] = Be[16 * u * S]; ao0[u] = Bo[16 * u * S]; }
#pragma unroll
                        for (int q = 0; q < NQ; ++q) { t0[q] = Ct[64 * q]; h0[q] = Ch[64 * q]; }
                        // two k-steps per iteration with ping-pong operand registers; the last loads
                        // run one k-step past Kp (values unused: rows are padded, the table is padded)
                        for (int j0 = 0; j0 < Kp; j0 += 8) {
#pragma unroll
                            for (int u = 0; u < NT; ++u) { ae1[u] = Be[16 * u * S + j0 + 4]; ao1[u] = Bo[16 * u * S + j0 + 4]; }
#pragma unroll
                            for (int q = 0; q < NQ; ++q) { t1[q] = Ct[64 * q - j0 - 4]; h1[q] = Ch[64 * q + j0 + 4]; }
                            LSB();
#pragma unroll
                            for (int q = 0; q < NQ; ++q) {
                                const T ke = fadd_s(sg, h0[q], t0[q]);
                                const T ko = fadd_s(-sg, h0[q], t0[q]);
#pragma unroll
                                for (int u = 0; u < NT; ++u) {
                                    aE[u][q] = DIR == 0 ? MF<T>::mma(ae0[u], ke, aE[u][q]) : MF<T>::mma(ke, ae0[u], aE[u][q]);
                                    aO[u][q] = DIR == 0 ? MF<T>::mma(ao0[u], ko, aO[u][q]) : MF<T>::mma(ko, ao0[u], aO[u][q]);
                                }
                            }
                            LSB();
#pragma unroll
                            for (int u = 0; u < NT; ++u) { ae0[u] = Be[16 * u * S + j0 + 8]; ao0[u] = Bo[16 * u * S + j0 + 8]; }
#pragma unroll
                            for (int q = 0; q < NQ; ++q) { t0[q] = Ct[64 * q - j0 - 8]; h0[q] = Ch[64 * q + j0 + 8]; }
                            LSB();
#pragma unroll
                            for (int q = 0; q < NQ; ++q) {
                                const T ke = fadd_s(sg, h1[q], t1[q]);
                                const T ko = fadd_s(-sg, h1[q], t1[q]);
#pragma unroll
                                for (int u = 0; u < NT; ++u) {
                                    aE[u][q] = DIR == 0 ? MF<T>::mma(ae1[u], ke, aE[u][q]) : MF<T>::mma(ke, ae1[u], aE[u][q]);
                                    aO[u][q] = DIR == 0 ? MF<T>::mma(ao1[u], ko, aO[u][q]) : MF<T>::mma(ko, ao1[u], aO[u][q]);
                                }
                            }
                            LSB();
                        }
                    }
                };
                // tiles past the half line hold no output rows: a wave whose last tile lies there skips it
                if (nq == QF) kloop(std::integral_constant<int, QF>{});
                else kloop(std::integral_constant<int, (QF > 1 ? QF - 1 : 0)>{});
            }
            if (t != G.ntiles - 1) continue;
            // all sources of output o are in: rank-4 expansion, constant term, store
            if constexpr (KIND == 2 && DIR == 0) {
                if (O.gmode == 1) continue;     // written as cross-term partials with the dots
            }
            if constexpr (KIND == 2) {
                if (G.has_lr) {
                    // E += UE D, O += UO D for the rank-4 term (one MFMA k-step per tile: the ranks are K)
                    for (int e = 0; e < O.nterms; ++e) {
                        if (!O.term[e].lr) continue;
                        const T* ue = static_cast<const T*>(P.ops[O.term[e].slot]) + 8 * Kp;
                        const T* uo = ue + 4 * P.Mp;
#pragma unroll
                        for (int u = 0; u < NT; ++u) {
                            const T d = DL[o * LB * 4 + (16 * u + ln) * 4 + lk];
#pragma unroll
                            for (int q = 0; q < QF; ++q) {
                                const int i0 = (role + 4 * q) * 16;
                                const T we = ue[4 * (i0 + ln) + lk], wo = uo[4 * (i0 + ln) + lk];
                                aE[u][q] = DIR == 0 ? MF<T>::mma(d, we, aE[u][q]) : MF<T>::mma(we, d, aE[u][q]);
                                aO[u][q] = DIR == 0 ? MF<T>::mma(d, wo, aO[u][q]) : MF<T>::mma(wo, d, aO[u][q]);
                            }
                        }
                    }
                }
            }
            // constant term of every line: sum over the terms of rsum times the source's x_0
            T cst[NT][4];
#pragma unroll
            for (int u = 0; u < NT; ++u)
#pragma unroll
                for (int r = 0; r < 4; ++r) cst[u][r] = 0;
            for (int e = 0; e < O.nterms; ++e) {
                if (O.term[e].lr || O.term[e].rsum == 0) continue;
                const T rsv = (T)O.term[e].rsum;
#pragma unroll
                for (int u = 0; u < NT; ++u)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int l = DIR == 0 ? 16 * u + MF<T>::row(lk, r) : 16 * u + ln;
                        cst[u][r] += rsv * X0[O.term[e].tile * LB + l];
                    }
            }
            const bool bad = fold_out<T, DIR, QF, NT>(P, O, aE, aO, cst, line0, x0, z, nl, N, role, lane);
            if (O.nan_check) {
                if (__any(bad) && lane == 0) {
                    const int step = __atomic_load_n(&P.flags[1], __ATOMIC_RELAXED);
                    atomicCAS(&P.flags[0], 0, step > 0 ? step : 1);
                }
            }
        }
    }
}

template <typename T, int DIR, int QF, int KIND, int NT, typename R>
__global__ __launch_bounds__(LP_THREADS) __attribute__((amdgpu_waves_per_eu(FP_WPE<T, R>))) void k_fold_pass(const LPass* __restrict__ Pp, const Members M, const int xcd_order) {
    fold_pass_body<T, DIR, QF, KIND, NT, R>(Pp, M, xcd_order, (int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x,
                                            (int)gridDim.y);
}

// An x-pass and a y-pass of one kind and shape in one launch (one context, one group each, 16 lines
// per work-group): work-groups [0, nb0) run the x-pass P0, the next nb1 the y-pass P1.  For two
// passes that are independent and each fill only part of the GPU (the constraint's).
template <typename T, int QF, int KIND, typename R>
__global__ __launch_bounds__(LP_THREADS) __attribute__((amdgpu_waves_per_eu(FP_WPE<T, R>))) void k_fold_pair(
    const LPass* __restrict__ P0, const LPass* __restrict__ P1, const int nb0, const int nb1, const int order0,
    const int order1) {
    const Members one{nullptr, 1};
    const int b = (int)blockIdx.x;
    // the y-pass starts at a multiple of 8, so its work-group i runs on XCD i mod 8 as in a launch of its
    // own and its XCD-aware order groups the same tiles on one L2; the padding work-groups do nothing
    const int nb0p = (nb0 + 7) & ~7;
    if (b < nb0) fold_pass_body<T, 0, QF, KIND, 1, R>(P0, one, order0, b, 0, nb0, 1);
    else if (b >= nb0p) fold_pass_body<T, 1, QF, KIND, 1, R>(P1, one, order1, b - nb0p, 0, nb1, 1);
}

template <typename T, int QF, int KIND, typename R>
hipError_t fold_pair_launch_one(const LPass* p0, const LPass* p1, int nb0, int nb1, size_t lds, hipStream_t s,
                                int order0, int order1) {
    const void* fn = reinterpret_cast<const void*>(&k_fold_pair<T, QF, KIND, R>);
    if (lds > 64 * 1024) {
        static bool done = false;
        if (!done) {
            const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            if (e != hipSuccess) return e;
            done = true;
        }
    }
    hipLaunchKernelGGL((k_fold_pair<T, QF, KIND, R>), dim3(((nb0 + 7) & ~7) + nb1), dim3(LP_THREADS), lds, s, p0, p1, nb0, nb1,
                       order0, order1);
    return hipGetLastError();
}

// The constraint pair (kind 1) of one precision, by QF (both passes of one QF).
template <typename T>
hipError_t fold_pair_launch_cons(const LPass* p0, const LPass* p1, int nb0, int nb1, int QF, size_t lds, hipStream_t s,
                                 int order0, int order1) {
    switch (QF) {
        case 1: return fold_pair_launch_one<T, 1, 1, T>(p0, p1, nb0, nb1, lds, s, order0, order1);
        case 2: return fold_pair_launch_one<T, 2, 1, T>(p0, p1, nb0, nb1, lds, s, order0, order1);
        case 3: return fold_pair_launch_one<T, 3, 1, T>(p0, p1, nb0, nb1, lds, s, order0, order1);
        case 4: return fold_pair_launch_one<T, 4, 1, T>(p0, p1, nb0, nb1, lds, s, order0, order1);
        case 5: return fold_pair_launch_one<T, 5, 1, T>(p0, p1, nb0, nb1, lds, s, order0, order1);
        default: return hipErrorInvalidValue;
    }
}

template <typename T, int DIR, int KIND, int NT, int QF, typename R>
hipError_t fold_launch_one(const LPass* d, Members M, dim3 grid, size_t lds, hipStream_t s, int order) {
    const void* fn = reinterpret_cast<const void*>(&k_fold_pass<T, DIR, QF, KIND, NT, R>);
    if (lds > 64 * 1024) {
        static bool done = false;
        if (!done) {
            const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            if (e != hipSuccess) return e;
            done = true;
        }
    }
    hipLaunchKernelGGL((k_fold_pass<T, DIR, QF, KIND, NT, R>), grid, dim3(LP_THREADS), lds, s, d, M, order);
    return hipGetLastError();
}

template <typename T, int DIR, int KIND, int NT, typename R>
hipError_t fold_launch_q(const LPass* d, Members M, dim3 grid, int QF, size_t lds, hipStream_t s, int order) {
    switch (QF) {
        case 1: return fold_launch_one<T, DIR, KIND, NT, 1, R>(d, M, grid, lds, s, order);
        case 2: return fold_launch_one<T, DIR, KIND, NT, 2, R>(d, M, grid, lds, s, order);
        case 3: return fold_launch_one<T, DIR, KIND, NT, 3, R>(d, M, grid, lds, s, order);
        case 4: return fold_launch_one<T, DIR, KIND, NT, 4, R>(d, M, grid, lds, s, order);
        case 5: return fold_launch_one<T, DIR, KIND, NT, 5, R>(d, M, grid, lds, s, order);
        default: return hipErrorInvalidValue;
    }
}

template <typename T, int DIR, typename R>
hipError_t fold_launch_tr(const LPass* d, Members M, dim3 grid, int QF, int nt, int kind, size_t lds, hipStream_t s,
                          int order) {
    if (nt == 2) {
        if (kind == 0) return fold_launch_q<T, DIR, 0, 2, R>(d, M, grid, QF, lds, s, order);
        if (kind == 1) return fold_launch_q<T, DIR, 1, 2, R>(d, M, grid, QF, lds, s, order);
        return fold_launch_q<T, DIR, 2, 2, R>(d, M, grid, QF, lds, s, order);
    }
    if (kind == 0) return fold_launch_q<T, DIR, 0, 1, R>(d, M, grid, QF, lds, s, order);
    if (kind == 1) return fold_launch_q<T, DIR, 1, 1, R>(d, M, grid, QF, lds, s, order);
    return fold_launch_q<T, DIR, 2, 1, R>(d, M, grid, QF, lds, s, order);
}

// Every folded instance of one (precision, direction) family; explicitly instantiated by the
// hqt_fold_*.hip translation units.  wide: prefetch registers in double (a double source in an
// FP32 pass); FP64 passes always use double.  order: the block order (k_fold_pass's xcd_order).
template <typename T, int DIR>
hipError_t fold_launch_td(const LPass* d, Members M, dim3 grid, int QF, int nt, int kind, size_t lds, hipStream_t s,
                          int wide, int order) {
    if constexpr (sizeof(T) == 8) {
        (void)wide;
        return fold_launch_tr<T, DIR, double>(d, M, grid, QF, nt, kind, lds, s, order);
    } else {
        if (wide) return fold_launch_tr<T, DIR, double>(d, M, grid, QF, nt, kind, lds, s, order);
        return fold_launch_tr<T, DIR, float>(d, M, grid, QF, nt, kind, lds, s, order);
    }
}

}  // namespace hqt
