// Line-operator passes for gfx950 (MI355X): the x/y spectral derivatives and the Fourier
// filter of the AdS4 stepper, applied as Toeplitz matrix products on MFMA.
//
// Reference: tevol_driven_cluster.m:60-61 (func_Lx / func_Ly: dense fourdif matrix times
// every line), src/numerics/ff.m:7-11 (filter).  A work-group owns 16 lines (the MFMA
// N/M dimension) of one output group and stages those lines in LDS once; each wave then
// sweeps its output position tiles over the full line with 16x16x4 MFMAs.  The operator
// matrix is never stored: its 16x4 fragments are read from a 2*Np-entry Toeplitz table
// (T[i][j] = tab[i - j + Np]) kept in LDS.
//
// Orientation: x-lines are contiguous in memory, so the output tile is (line x position)
// and every 16 lanes write 16 consecutive x values; y-lines are strided by nx, so the tile
// is (position x line) with lines = 16 consecutive x values, which again makes every 16
// lanes write 128 contiguous bytes.  The MFMA operand order is the only difference.
#pragma once
#include "hqt_buffer.h"
#include "hqt_kernels.h"

namespace hqt {

// 16x16x4 MFMA operand layouts (A[m][k] in lane m + 16k, B[k][n] in lane n + 16k).  The
// accumulator row held by lane group g = lane/16 in register r differs between the two
// instructions on gfx950 (measured): f64 rows are 4r + g, f32 rows are 4g + r.
template <typename T> struct MF;
template <> struct MF<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int r) { return 4 * r + g; }
};
template <> struct MF<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

template <typename T>
__device__ __forceinline__ T ldv(const void* p, int dt, int64_t i) {
    return dt == DT_F64 ? (T)(static_cast<const double*>(p)[i]) : (T)(static_cast<const float*>(p)[i]);
}
template <typename T>
__device__ __forceinline__ void stv(void* p, int dt, int64_t i, T v) {
    if (dt == DT_F64) static_cast<double*>(p)[i] = (double)v;
    else static_cast<float*>(p)[i] = (float)v;
}

// Rank-4 terms (the filter's imaginary part, U V^T with U, V of 4 columns; filter kernels only).
// Dots D^T (lines x 4) = X V on MFMA: the 4 waves split the k range (ranks padded to 16 MFMA
// columns), their partial tiles go to LDS and are summed in wave order, so the result does not
// depend on the lines per tile.  Table layout: V (Np x 4) then U (Mp x 4), row-major.
template <typename T, int NT>
__device__ __forceinline__ void lr_dots(const T* __restrict__ X, int S, int Np, const T* __restrict__ vt, T* DLp,
                                        T* dl, int wave, int lane) {
    typedef typename MF<T>::v4 v4;
    constexpr int LB = 16 * NT;
    const int ln = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int u = 0; u < NT; ++u) {
        v4 acc = v4{0, 0, 0, 0};
        for (int j0 = 4 * wave; j0 < Np; j0 += 16) {
            const T a = X[(16 * u + ln) * S + j0 + lk];
            const T b = ln < 4 ? vt[4 * (j0 + lk) + ln] : (T)0;
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

// Dots of every rank-4 term whose source is tile t (one per output at most); ends with a barrier.
template <typename T, int NT>
__device__ __forceinline__ void lr_terms(const LPass& P, const LGroup& G, int t, const T* X, int S, int Np, T* DLp,
                                         T* DL, int wave, int lane) {
    constexpr int LB = 16 * NT;
    for (int o = 0; o < G.nout; ++o)
        for (int e = 0; e < G.out[o].nterms; ++e) {
            const LTerm& tm = G.out[o].term[e];
            if (tm.lr && tm.tile == t) {
                lr_dots<T, NT>(X, S, Np, static_cast<const T*>(P.ops[tm.slot]), DLp, DL + o * LB * 4, wave, lane);
                __syncthreads();
            }
        }
}

// acc += U D for output o's rank-4 term: one MFMA k-step per output tile (the 4 ranks are K).
template <typename T, int DIR, int Q, int NT>
__device__ __forceinline__ void lr_expand(const LPass& P, const LOut& O, int o, int Np, const T* DL, int wave, int lane,
                                          typename MF<T>::v4 (&acc)[NT][Q]) {
    constexpr int LB = 16 * NT;
    const int ln = lane & 15, lk = lane >> 4;
    for (int e = 0; e < O.nterms; ++e) {
        if (!O.term[e].lr) continue;
        const T* ut = static_cast<const T*>(P.ops[O.term[e].slot]) + 4 * Np;
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            const T d = DL[o * LB * 4 + (16 * u + ln) * 4 + lk];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int i0 = (wave + 4 * q) * 16;
                const T uu = ut[4 * (i0 + ln) + lk];
                acc[u][q] = DIR == 0 ? MF<T>::mma(d, uu, acc[u][q]) : MF<T>::mma(uu, d, acc[u][q]);
            }
        }
    }
}

// Operator tables of a group into LDS: every thread issues all of its table loads (up to 4 slots
// of TL <= Np + Mp + 8 <= 128 Q + 8 entries) before the first LDS write.
template <typename T, int Q>
__device__ __forceinline__ void stage_tables(const LPass& P, const LGroup& G, T* C, int TL, int tid) {
    constexpr int TPT = (128 * Q + 8 + LP_THREADS - 1) / LP_THREADS;
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

// Keeps the operand loads of one k-step ahead of the MFMAs of the other (no instruction crosses it).
#define LSB() __builtin_amdgcn_sched_barrier(0)

// Occupancy the line kernels' register allocation targets (the branch-free loads would otherwise
// be hoisted until one wave per SIMD fits).
template <typename T> constexpr int LP_WPE = sizeof(T) == 8 ? 2 : 3;          // k_line_pass
template <typename T> constexpr int LP1_WPE = sizeof(T) == 8 ? 2 : 4;         // k_line_pass_one

// Bytes of one [nx, ny, nz] field of element type S (the buffer bound of every line-pass access).
template <typename S>
__device__ __forceinline__ int64_t fbytes(const LPass& P) {
    return (int64_t)P.ps * P.nz * (int64_t)sizeof(S);
}

// Outputs of a finished line tile: out = acc (+ addend).  The addends are all loaded before the
// first store and the data-type dispatch sits outside the unrolled loops, so a tile's output
// costs one memory round trip.  Returns whether any stored value is NaN; clears acc.
template <typename T, int DIR, int Q, int NT>
__device__ __forceinline__ bool tile_out(const LPass& P, const LOut& O, typename MF<T>::v4 (&acc)[NT][Q], int64_t line0,
                                         int x0, int z, int64_t nl, int N, int wave, int lane) {
    const int ln = lane & 15, lk = lane >> 4;
    // element index of accumulator (u, q, r) or Z_OOB outside the grid (32-bit: fields hold
    // fewer than 2^23 elements)
    const int l0 = (int)line0, nlv = (int)nl, nx = P.nx, rp = P.rp, pl = P.ps * z;
    // x-lines: element offsets of the 4 NT lines this lane stores
    int xb[NT][4];
    const int z0 = DIR == 0 ? l0 / P.ny : 0;
#pragma unroll
    for (int u = 0; u < NT; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) xb[u][r] = DIR == 0 ? xline_base_at(P, l0, z0, 16 * u + MF<T>::row(lk, r)) : 0;
    auto at = [&](int u, int q, int r) -> uint32_t {
        const int i0 = (wave + 4 * q) * 16;
        if (DIR == 0) {
            const int line = l0 + 16 * u + MF<T>::row(lk, r), i = i0 + ln;
            return (line < nlv && i < N) ? (uint32_t)(xb[u][r] + i) : Z_OOB;
        } else {
            const int x = x0 + 16 * u + ln, i = i0 + MF<T>::row(lk, r);
            return (x < nx && i < N) ? (uint32_t)(x + rp * i + pl) : Z_OOB;
        }
    };
    T add[NT][Q][4];
    auto load_add = [&](auto tag) {
        using S_ = decltype(tag);
        const Rs rs = rs_make(O.addend, fbytes<S_>(P));
#pragma unroll
        for (int u = 0; u < NT; ++u)
#pragma unroll
            for (int q = 0; q < Q; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    add[u][q][r] = (T)bldr<S_>(rs, at(u, q, r));
                }
    };
    if (!O.addend) {
#pragma unroll
        for (int u = 0; u < NT; ++u)
#pragma unroll
            for (int q = 0; q < Q; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r) add[u][q][r] = 0;
    } else if (O.add_dt == DT_F64) {
        load_add(double{});
    } else {
        load_add(float{});
    }
    bool bad = false;
    // out-of-grid elements are dropped by the buffer bounds check
    auto store = [&](auto tag) {
        using S_ = decltype(tag);
        const Rs rs = rs_make(O.out, fbytes<S_>(P));
#pragma unroll
        for (int u = 0; u < NT; ++u)
#pragma unroll
            for (int q = 0; q < Q; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const uint32_t e = at(u, q, r);
                    const T v = O.addend ? acc[u][q][r] + add[u][q][r] : acc[u][q][r];
                    bstr<S_>(rs, e, (S_)v);
                    bad |= e != Z_OOB && v != v;
                    acc[u][q][r] = 0;
                }
    };
    if (O.out_dt == DT_F64) store(double{});
    else store(float{});
    return bad;
}

// KIND only tags the instantiation (0: RHS stage, 1: constraint, 2: filter) so that
// profilers report the three uses separately; the code is identical.
// NT: 16-line tiles per work-group (LB = 16*NT lines).  With NT = 2 every operator fragment
// read from LDS feeds two MFMAs (one per line tile), halving the LDS operand traffic per MFMA
// at twice the accumulator registers and LDS line storage.
// This form serves the x-lines: a work-group processes P.tpb consecutive line tiles, the
// operator tables are staged once, and the lines of the next (tile, source) pair are loaded into
// registers while the MFMAs of the current pair run, so global-load latency overlaps the matrix
// work (k_line_pass_one below serves the y-lines).  Which NT and tpb are faster depends on the
// line length and precision; the host measures and picks per pass.  Every output is accumulated
// by the same MFMA sequence in all variants.
template <typename T, int DIR, int Q, int KIND, int NT>
__global__ __launch_bounds__(LP_THREADS) __attribute__((amdgpu_waves_per_eu(LP_WPE<T>))) void k_line_pass(const LPass* __restrict__ Pp, const Members M) {
    constexpr int LB = 16 * NT;
    constexpr int PF = 4 * Q * NT;            // staged elements per thread: LB*Np/256 <= LB*Mp/256 = 4*Q*NT
    const MemberBlock mb = member_block(M);
    const LPass& P = Pp[mb.m];
    if ((int)blockIdx.y >= P.ngroups) return;
    if (__builtin_nontemporal_load(&P.flags[0]) != 0) return;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int S = P.S, Np = P.Np, N = P.N, TL = P.TL;
    T* V = reinterpret_cast<T*>(smem_raw);
    T* C = V + LB * S;
    const LGroup& G = P.g[blockIdx.y];
    T* DL = C + G.nslots * TL;               // rank-4 dots: 2 outputs x LB lines x 4
    T* DLp = DL + 2 * LB * 4;                // rank-4 partial dots: 4 waves x LB lines x 4
    const int tid = threadIdx.x;

    stage_tables<T, Q>(P, G, C, TL, tid);

    const int64_t nl = (int64_t)P.ny * P.nz;
    const int tile0 = mb.bx * P.tpb;
    const int ntile = min(P.tpb, P.ntiles - tile0);
    const int npair = ntile * G.ntiles;

    typedef typename MF<T>::v4 v4;
    v4 acc[2][NT][Q];
#pragma unroll
    for (int o = 0; o < 2; ++o)
#pragma unroll
        for (int u = 0; u < NT; ++u)
#pragma unroll
            for (int q = 0; q < Q; ++q) acc[o][u][q] = v4{0, 0, 0, 0};

    const int lane = tid & 63, ln = lane & 15, lk = lane >> 4;
    // wave index and its tile count are wave-uniform: keep them in scalar registers so the
    // per-tile guards below are scalar branches, not exec-masked ones
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    // Lines of (tile, source) pair `pr` into registers.  x-lines: thread -> line m = tid/16, line
    // points j = tid%16 + 16 r (16 lanes read 128 contiguous bytes of one line); y-lines: thread
    // -> x offset m = tid%LB, line points j = tid/LB + (256/LB) r (LB lanes read LB contiguous
    // x values).  Either way a thread's elements are a fixed stride apart: one base offset, no
    // divisions, PF registers for the data.
    constexpr int JS = DIR == 0 ? 16 : LP_THREADS / LB;     // j step between a thread's elements
    const int mrow = DIR == 0 ? (tid >> 4) : (tid % LB);
    const int jt = DIR == 0 ? (tid & 15) : (tid / LB);
    T pf[PF];
    auto fetch = [&](int pr) {
        const int lt = tile0 + pr / G.ntiles, t = pr % G.ntiles;
        const void* src = G.src[t];
        const int dt = G.src_dt[t];
        bool okm;
        int64_t base, step;
        if (DIR == 0) {
            // LB lines per tile; with LB = 32 a thread covers lines mrow and mrow + 16
            const int64_t line = (int64_t)lt * LB + mrow;
            okm = line < nl;
            base = 0;                    // x-line offsets come from xbh below
            step = 0;
        } else {
            const int x = (lt % P.nxc) * LB + mrow, zz = lt / P.nxc;
            okm = x < P.nx;
            base = (int64_t)x + (int64_t)P.ps * zz;
            step = (int64_t)P.rp;
        }
        // x-lines: element offsets of this thread's NT lines
        int xbh[NT];
        if (DIR == 0) {
            const int l0 = lt * LB, z0 = l0 / P.ny;
#pragma unroll
            for (int u = 0; u < NT; ++u) xbh[u] = xline_base_at(P, l0, z0, mrow + 16 * u);
        }
        // out-of-grid elements read 0 through the buffer bounds check (no branches around the loads)
        auto load = [&](auto tag) {
            using S_ = decltype(tag);
            const Rs rs = rs_make(src, fbytes<S_>(P));
#pragma unroll
            for (int r = 0; r < PF; ++r) {
                uint32_t e;
                if (DIR == 0) {
                    const int half = r / (PF / NT), rr = r % (PF / NT);     // NT = 2: second line at +16
                    const int j = jt + JS * rr;
                    const int line = lt * LB + mrow + 16 * half;
                    e = (line < (int)nl && j < N) ? (uint32_t)(xbh[half] + j) : Z_OOB;
                } else {
                    const int j = jt + JS * r;
                    e = (okm && j < N) ? (uint32_t)((int)base + (int)step * j) : Z_OOB;
                }
                pf[r] = (T)bldr<S_>(rs, e);
            }
        };
        if (dt == DT_F64) load(double{});
        else load(float{});
    };
    auto deposit = [&]() {
#pragma unroll
        for (int r = 0; r < PF; ++r) {
            if (DIR == 0) {
                const int half = r / (PF / NT), rr = r % (PF / NT);
                const int j = jt + JS * rr;
                if (j < Np) V[(mrow + 16 * half) * S + j] = pf[r];
            } else {
                const int j = jt + JS * r;
                if (j < Np) V[mrow * S + j] = pf[r];
            }
        }
    };

    // x-lines prefetch the next pair into registers; y-lines use k_line_pass_one (launchers)
    if (npair > 0) fetch(0);
    for (int pr = 0; pr < npair; ++pr) {
        const int t = pr % G.ntiles;
        __syncthreads();                 // the previous pair's MFMAs are done with V
        deposit();
        __syncthreads();
        if (pr + 1 < npair) fetch(pr + 1);
        if constexpr (KIND == 2) {
            if (G.has_lr) lr_terms<T, NT>(P, G, t, V, S, Np, DLp, DL, wave, lane);
        }
#pragma unroll
        for (int o = 0; o < 2; ++o) {
            if (o >= G.nout) break;
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                if (e >= G.out[o].nterms || G.out[o].term[e].tile != t || G.out[o].term[e].lr) continue;
                // table entry of T[i][j] is tab[i - j + Np + 4]; lane (ln, lk) reads i = i0 + ln, j = j0 + lk
                const T* Cs = C + G.out[o].term[e].slot * TL + Np + 4 + ln - lk + wave * 16;
                const T* Vr = V + ln * S + lk;
                // two k-steps per iteration with ping-pong operand registers: the loads of
                // one k-step are in flight while the other's MFMAs issue.  The last loads run
                // one k-step past the tile (values unused; the table is padded for them and
                // the line rows are followed by the table in LDS).
                T va0[NT], va1[NT];
                T cb0[Q], cb1[Q];
#pragma unroll
                for (int u = 0; u < NT; ++u) va0[u] = Vr[16 * u * S];
#pragma unroll
                for (int q = 0; q < Q; ++q) cb0[q] = Cs[64 * q];
                for (int j0 = 0; j0 < Np; j0 += 8) {
#pragma unroll
                    for (int u = 0; u < NT; ++u) va1[u] = Vr[16 * u * S + j0 + 4];
#pragma unroll
                    for (int q = 0; q < Q; ++q) cb1[q] = Cs[64 * q - j0 - 4];
                    LSB();
#pragma unroll
                    for (int u = 0; u < NT; ++u)
#pragma unroll
                        for (int q = 0; q < Q; ++q)
                            acc[o][u][q] = DIR == 0 ? MF<T>::mma(va0[u], cb0[q], acc[o][u][q])
                                                    : MF<T>::mma(cb0[q], va0[u], acc[o][u][q]);
                    LSB();
#pragma unroll
                    for (int u = 0; u < NT; ++u) va0[u] = Vr[16 * u * S + j0 + 8];
#pragma unroll
                    for (int q = 0; q < Q; ++q) cb0[q] = Cs[64 * q - j0 - 8];
                    LSB();
#pragma unroll
                    for (int u = 0; u < NT; ++u)
#pragma unroll
                        for (int q = 0; q < Q; ++q)
                            acc[o][u][q] = DIR == 0 ? MF<T>::mma(va1[u], cb1[q], acc[o][u][q])
                                                    : MF<T>::mma(cb1[q], va1[u], acc[o][u][q]);
                    LSB();
                }
            }
        }
        if (t != G.ntiles - 1) continue;
        // all sources of this line tile are in: write its outputs and restart the accumulators
        const int lt = tile0 + pr / G.ntiles;
        int64_t line0 = 0;
        int z = 0, x0 = 0;
        if (DIR == 0) {
            line0 = (int64_t)lt * LB;
        } else {
            x0 = (lt % P.nxc) * LB;
            z = lt / P.nxc;
        }
#pragma unroll
        for (int o = 0; o < 2; ++o) {
            if (o >= G.nout) break;
            const LOut& O = G.out[o];
            if constexpr (KIND == 2) {
                if (G.has_lr) lr_expand<T, DIR, Q, NT>(P, O, o, Np, DL, wave, lane, acc[o]);
            }
            const bool bad = tile_out<T, DIR, Q, NT>(P, O, acc[o], line0, x0, z, nl, N, wave, lane);
            if (O.nan_check) {
                if (__any(bad) && lane == 0) {
                    const int step = __atomic_load_n(&P.flags[1], __ATOMIC_RELAXED);
                    atomicCAS(&P.flags[0], 0, step > 0 ? step : 1);
                }
            }
        }
    }
}

// One line tile per work-group, staged global to LDS inside the tile loop: the y-line form
// (lower register use keeps more waves resident, which the strided y accesses need).
template <typename T, int DIR, int Q, int KIND, int NT>
__global__ __launch_bounds__(LP_THREADS) __attribute__((amdgpu_waves_per_eu(LP1_WPE<T>))) void k_line_pass_one(const LPass* __restrict__ Pp, const Members M) {
    constexpr int LB = 16 * NT;
    const MemberBlock mb = member_block(M);
    const LPass& P = Pp[mb.m];
    if ((int)blockIdx.y >= P.ngroups) return;
    if (__builtin_nontemporal_load(&P.flags[0]) != 0) return;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int S = P.S, Np = P.Np, N = P.N, TL = P.TL;
    T* V = reinterpret_cast<T*>(smem_raw);
    T* C = V + LB * S;
    const LGroup& G = P.g[blockIdx.y];
    T* DL = C + G.nslots * TL;               // rank-4 dots: 2 outputs x LB lines x 4
    T* DLp = DL + 2 * LB * 4;                // rank-4 partial dots: 4 waves x LB lines x 4
    const int tid = threadIdx.x;

    stage_tables<T, Q>(P, G, C, TL, tid);

    int64_t line0 = 0, nl = 0;
    int z = 0, x0 = 0;
    if (DIR == 0) {
        line0 = (int64_t)mb.bx * LB;
        nl = (int64_t)P.ny * P.nz;
    } else {
        const int xc = mb.bx % P.nxc;
        z = mb.bx / P.nxc;
        x0 = xc * LB;
    }

    typedef typename MF<T>::v4 v4;
    v4 acc[2][NT][Q];
#pragma unroll
    for (int o = 0; o < 2; ++o)
#pragma unroll
        for (int u = 0; u < NT; ++u)
#pragma unroll
            for (int q = 0; q < Q; ++q) acc[o][u][q] = v4{0, 0, 0, 0};

    const int lane = tid & 63, ln = lane & 15, lk = lane >> 4;
    // wave index and its tile count are wave-uniform: keep them in scalar registers so the
    // per-tile guards below are scalar branches, not exec-masked ones
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    // Staging of one tile: a thread's PF elements are a fixed stride apart (x-lines: line tid/16,
    // points tid%16 + 16 r; y-lines: x offset tid%LB, points tid/LB + (256/LB) r), and all PF loads
    // are issued before the first LDS write, so a tile costs one memory round trip per thread.
    constexpr int PF = 4 * Q * NT;            // LB*Mp/256 >= LB*Np/256 elements per thread
    constexpr int JS = DIR == 0 ? 16 : LP_THREADS / LB;
    const int mrow = DIR == 0 ? (tid >> 4) : (tid % LB);
    const int jt = DIR == 0 ? (tid & 15) : (tid / LB);
    for (int t = 0; t < G.ntiles; ++t) {
        __syncthreads();
        const void* src = G.src[t];
        const int dt = G.src_dt[t];
        T pf[PF];
        int64_t base;
        int64_t step;
        bool okm;
        if (DIR == 0) {
            okm = true;
            base = 0;
            step = 1;
        } else {
            okm = x0 + mrow < P.nx;
            base = (int64_t)(x0 + mrow) + (int64_t)P.ps * z;
            step = (int64_t)P.rp;
        }
        int xbh[NT];
        if (DIR == 0) {
            const int l0 = (int)line0, z0 = l0 / P.ny;
#pragma unroll
            for (int u = 0; u < NT; ++u) xbh[u] = xline_base_at(P, l0, z0, mrow + 16 * u);
        }
        auto load = [&](auto tag) {
            using S_ = decltype(tag);
            const Rs rs = rs_make(src, fbytes<S_>(P));
#pragma unroll
            for (int r = 0; r < PF; ++r) {
                uint32_t e;
                if (DIR == 0) {
                    const int half = r / (PF / NT), rr = r % (PF / NT);
                    const int j = jt + JS * rr;
                    const int line = (int)line0 + mrow + 16 * half;
                    e = (line < (int)nl && j < N) ? (uint32_t)(xbh[half] + j) : Z_OOB;
                } else {
                    const int j = jt + JS * r;
                    e = (okm && j < N) ? (uint32_t)((int)base + (int)step * j) : Z_OOB;
                }
                pf[r] = (T)bldr<S_>(rs, e);
            }
        };
        if (dt == DT_F64) load(double{});
        else load(float{});
#pragma unroll
        for (int r = 0; r < PF; ++r) {
            if (DIR == 0) {
                const int half = r / (PF / NT), rr = r % (PF / NT);
                const int j = jt + JS * rr;
                if (j < Np) V[(mrow + 16 * half) * S + j] = pf[r];
            } else {
                const int j = jt + JS * r;
                if (j < Np) V[mrow * S + j] = pf[r];
            }
        }
        __syncthreads();
        if constexpr (KIND == 2) {
            if (G.has_lr) lr_terms<T, NT>(P, G, t, V, S, Np, DLp, DL, wave, lane);
        }
#pragma unroll
        for (int o = 0; o < 2; ++o) {
            if (o >= G.nout) break;
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                if (e >= G.out[o].nterms || G.out[o].term[e].tile != t || G.out[o].term[e].lr) continue;
                // table entry of T[i][j] is tab[i - j + Np + 4]; lane (ln, lk) reads i = i0 + ln, j = j0 + lk
                const T* Cs = C + G.out[o].term[e].slot * TL + Np + 4 + ln - lk + wave * 16;
                const T* Vr = V + ln * S + lk;
                // two k-steps per iteration with ping-pong operand registers: the loads of
                // one k-step are in flight while the other's MFMAs issue.  The last loads run
                // one k-step past the tile (values unused; the table is padded for them and
                // the line rows are followed by the table in LDS).
                T va0[NT], va1[NT];
                T cb0[Q], cb1[Q];
#pragma unroll
                for (int u = 0; u < NT; ++u) va0[u] = Vr[16 * u * S];
#pragma unroll
                for (int q = 0; q < Q; ++q) cb0[q] = Cs[64 * q];
                for (int j0 = 0; j0 < Np; j0 += 8) {
#pragma unroll
                    for (int u = 0; u < NT; ++u) va1[u] = Vr[16 * u * S + j0 + 4];
#pragma unroll
                    for (int q = 0; q < Q; ++q) cb1[q] = Cs[64 * q - j0 - 4];
                    LSB();
#pragma unroll
                    for (int u = 0; u < NT; ++u)
#pragma unroll
                        for (int q = 0; q < Q; ++q)
                            acc[o][u][q] = DIR == 0 ? MF<T>::mma(va0[u], cb0[q], acc[o][u][q])
                                                    : MF<T>::mma(cb0[q], va0[u], acc[o][u][q]);
                    LSB();
#pragma unroll
                    for (int u = 0; u < NT; ++u) va0[u] = Vr[16 * u * S + j0 + 8];
#pragma unroll
                    for (int q = 0; q < Q; ++q) cb0[q] = Cs[64 * q - j0 - 8];
                    LSB();
#pragma unroll
                    for (int u = 0; u < NT; ++u)
#pragma unroll
                        for (int q = 0; q < Q; ++q)
                            acc[o][u][q] = DIR == 0 ? MF<T>::mma(va1[u], cb1[q], acc[o][u][q])
                                                    : MF<T>::mma(cb1[q], va1[u], acc[o][u][q]);
                    LSB();
                }
            }
        }
    }

#pragma unroll
    for (int o = 0; o < 2; ++o) {
        if (o >= G.nout) break;
        const LOut& O = G.out[o];
        if constexpr (KIND == 2) {
            if (G.has_lr) lr_expand<T, DIR, Q, NT>(P, O, o, Np, DL, wave, lane, acc[o]);
        }
        const bool bad = tile_out<T, DIR, Q, NT>(P, O, acc[o], line0, x0, z, nl, N, wave, lane);
        if (O.nan_check) {
            if (__any(bad) && lane == 0) {
                const int step = __atomic_load_n(&P.flags[1], __ATOMIC_RELAXED);
                atomicCAS(&P.flags[0], 0, step > 0 ? step : 1);
            }
        }
    }
}

// Launch of one (precision, direction) family: Q = Mp/64 output tiles per wave, NT line tiles
// per work-group, KIND the profiling tag.  Dynamic LDS above 64 KB is enabled per instance on
// first use.
template <typename T, int DIR, int KIND, int NT, int Q>
hipError_t line_launch_one(const LPass* d, Members M, dim3 grid, size_t lds, hipStream_t s) {
    const void* fn;
    if constexpr (DIR == 0) fn = reinterpret_cast<const void*>(&k_line_pass<T, DIR, Q, KIND, NT>);
    else fn = reinterpret_cast<const void*>(&k_line_pass_one<T, DIR, Q, KIND, NT>);
    if (lds > 64 * 1024) {
        static bool done = false;
        if (!done) {
            const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            if (e != hipSuccess) return e;
            done = true;
        }
    }
    if constexpr (DIR == 0) hipLaunchKernelGGL((k_line_pass<T, DIR, Q, KIND, NT>), grid, dim3(LP_THREADS), lds, s, d, M);
    else hipLaunchKernelGGL((k_line_pass_one<T, DIR, Q, KIND, NT>), grid, dim3(LP_THREADS), lds, s, d, M);
    return hipGetLastError();
}

template <typename T, int DIR, int KIND, int NT>
hipError_t line_launch_q(const LPass* d, Members M, dim3 grid, int Q, size_t lds, hipStream_t s) {
    switch (Q) {
        case 1: return line_launch_one<T, DIR, KIND, NT, 1>(d, M, grid, lds, s);
        case 2: return line_launch_one<T, DIR, KIND, NT, 2>(d, M, grid, lds, s);
        case 3: return line_launch_one<T, DIR, KIND, NT, 3>(d, M, grid, lds, s);
        case 4: return line_launch_one<T, DIR, KIND, NT, 4>(d, M, grid, lds, s);
        case 5: return line_launch_one<T, DIR, KIND, NT, 5>(d, M, grid, lds, s);
        case 6: return line_launch_one<T, DIR, KIND, NT, 6>(d, M, grid, lds, s);
        case 7: return line_launch_one<T, DIR, KIND, NT, 7>(d, M, grid, lds, s);
        case 8: return line_launch_one<T, DIR, KIND, NT, 8>(d, M, grid, lds, s);
        default: return hipErrorInvalidValue;
    }
}

// Every instance of one (precision, direction) family; explicitly instantiated by the
// hqt_line_*.hip translation units so that the four families compile in parallel.
template <typename T, int DIR>
hipError_t line_launch_td(const LPass* d, Members M, dim3 grid, int Q, int nt, int kind, size_t lds, hipStream_t s) {
    if (nt == 2) {
        if (kind == 0) return line_launch_q<T, DIR, 0, 2>(d, M, grid, Q, lds, s);
        if (kind == 1) return line_launch_q<T, DIR, 1, 2>(d, M, grid, Q, lds, s);
        return line_launch_q<T, DIR, 2, 2>(d, M, grid, Q, lds, s);
    }
    if (kind == 0) return line_launch_q<T, DIR, 0, 1>(d, M, grid, Q, lds, s);
    if (kind == 1) return line_launch_q<T, DIR, 1, 1>(d, M, grid, Q, lds, s);
    return line_launch_q<T, DIR, 2, 1>(d, M, grid, Q, lds, s);
}

}  // namespace hqt
