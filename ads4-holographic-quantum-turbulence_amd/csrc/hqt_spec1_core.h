// Factored-DFT line passes of the RK stages, the constraint and the Fourier filter, single-region form (gfx950).
//
// Reference: tevol_driven_cluster.m:60-61 (func_Lx / func_Ly), :129 (the constraint's d_x A_x, d_y A_y) and
// :142 with src/numerics/ff.m:7-11 (the filter), third_party/dmsuite/fourdif.m.  The transform is the two-stage
// real DFT of
// hqt_spec_core.h (N = N1 N2, stage A a real DFT of length N2 of each of the N1 sub-sequences,
// stage B a complex DFT of length N1 per kept column); what differs is how a work-group uses the CU:
//
//   * one output per group: a pass's groups are split on the host so that a group has one output (a source
//     with two outputs becomes two groups, each transforming the source again), so the spectrum never has to
//     outlive its output and a work-group needs one LDS line region (the filter's x-pass groups carry a second,
//     non-transform output: the cross term's partials);
//   * stage B fused with the spectral multiply and its own inverse: a column task transforms its column
//     forward, multiplies it by the output term's eigenvalues and transforms it back in registers, in
//     place in the LDS line region.  No trig table in LDS: a work-group holds 16 lines (38.5 KB in FP64,
//     half that in FP32), so four FP64 (more FP32) work-groups share a CU;
//   * stage A accumulates in streaming form (acc[k] += v_j * cos(2 pi j k / N2) over the folded inputs
//     v_j), with the cos / sin constants of the compile-time sizes read by scalar loads into SGPRs, so the
//     phase holds one accumulator per output and no constant in vector registers; the inverse writes the
//     line back in natural order (even half to both mirror positions, then the odd half subtracted and
//     added), so the output phase reads it like the input phase wrote it;
//   * an output with two sources (d_y(d_y A_x - d_x A_y)) stores the first source's result to the output field
//     and reads it back as the second source's addend; an output's addend is loaded just before the store.
//     Nothing but the transform's own values is held in registers across the transforms, and every transform
//     phase is compiled once (rolled loops over sources and outputs): the FP64 stage passes fit 128 VGPRs without
//     spills, so four work-groups share a CU (4 x 38.5 KB of LDS; line_x_stage 68.5 -> 63.2 us, line_y_stage
//     91.8 -> 83.2 us at 301^2 x 20 against three).
//
// The stage passes apply d/dx (i p s), d2/dx2 (-(p s)^2) and -d/dx (-i p s), s the signed wavenumber (0 at
// the Nyquist mode of even N): exactly the eigenvalues of the fourdif matrix Lx = p Dx and of Lx Lx.  The
// filter's passes (FILT) apply the even part of its keep-mask as the multiplier and carry the mask's rank-4 odd
// part through the cross-term slice matrices G (hqt_spec_core.h describes the decomposition).
#pragma once
#include "hqt_spec_core.h"

namespace hqt {

// Byte offset of the filter cross-term area (80 doubles, then 16 x 4 values of the pass precision, spec_cross_bytes)
// in the dynamic LDS of a single-region pass: after the line region, 8-byte aligned.
__host__ __device__ inline int spec1_cross_offset(int S, int es) { return (16 * S * es + 7) & ~7; }

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)

// A uniform table pointer viewed in the constant address space: loads with compile-time offsets from it are scalar
// loads (the values land in SGPRs, shared by the wave).  The pointer passes through an empty asm statement at each
// use site, so the loads stay in the phase that uses them instead of being hoisted (and kept live) across phases.
template <typename T> using cptr = const __attribute__((address_space(4))) T*;
template <typename T> __device__ __forceinline__ cptr<T> phase_table(const void* p) {
    uint64_t v = reinterpret_cast<uint64_t>(p);
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    asm volatile("" : "+s"(lo), "+s"(hi));
    return (cptr<T>)(((uint64_t)hi << 32) | lo);
}

// acc[K] (+)= v cos(2 pi J K / N2), K = 0 .. H  (INIT: acc[K] = v cos(...)); ca[m] = cos(2 pi m / N2), m <= H
template <typename G, int J, bool INIT, typename T, int... K>
__device__ __forceinline__ void cos_axpy(T (&acc)[G::H + 1], T v, const T (&ca)[G::H + 1], std::integer_sequence<int, K...>) {
    if constexpr (INIT) ((acc[K] = v * cos_m<G>(ca, (J * K) % G::N2)), ...);
    else ((acc[K] = fmat(v, cos_m<G>(ca, (J * K) % G::N2), acc[K])), ...);
}
// acc[K + 1] (+)= v sin(2 pi J (K + 1) / N2), K + 1 = 1 .. HI (acc[0] unused); sa[m] = sin(2 pi m / N2)
template <typename G, int J, bool INIT, typename T, int... K>
__device__ __forceinline__ void sin_axpy(T (&acc)[G::H + 1], T v, const T (&sa)[G::H + 1], std::integer_sequence<int, K...>) {
    if constexpr (INIT) ((acc[K + 1] = v * sin_m<G>(sa, (J * (K + 1)) % G::N2)), ...);
    else ((acc[K + 1] = fmat(v, sin_m<G>(sa, (J * (K + 1)) % G::N2), acc[K + 1])), ...);
}

// Forward stage A, even half: acc[k] = x[0] + sum_{j=1..HI} (x[j] + x[N2 - j]) cos(2 pi j k / N2)
// (+ x[H] cos(pi k) for even N2); x[j] = row[N1 j].
template <typename G, typename T, int... J>
__device__ __forceinline__ void fwd_cos_stream(const T* row, T (&acc)[G::H + 1], const T (&ca)[G::H + 1],
                                               std::integer_sequence<int, J...>) {
    constexpr int N1 = G::N1, N2 = G::N2;
    cos_axpy<G, 0, true>(acc, row[0], ca, std::make_integer_sequence<int, G::H + 1>{});
    ((cos_axpy<G, J + 1, false>(acc, row[N1 * (J + 1)] + row[N1 * (N2 - J - 1)], ca,
                                std::make_integer_sequence<int, G::H + 1>{})), ...);
    if constexpr (G::EV2) cos_axpy<G, G::H, false>(acc, row[N1 * G::H], ca, std::make_integer_sequence<int, G::H + 1>{});
}
// Forward stage A, odd half: acc[k] = sum_{j=1..HI} (x[j] - x[N2 - j]) sin(2 pi j k / N2), k = 1 .. HI.
template <typename G, typename T, int J0, int... J>
__device__ __forceinline__ void fwd_sin_stream(const T* row, T (&acc)[G::H + 1], const T (&sa)[G::H + 1],
                                               std::integer_sequence<int, J0, J...>) {
    constexpr int N1 = G::N1, N2 = G::N2;
    sin_axpy<G, 1, true>(acc, row[N1] - row[N1 * (N2 - 1)], sa, std::make_integer_sequence<int, G::HI>{});
    ((sin_axpy<G, J + 1, false>(acc, row[N1 * (J + 1)] - row[N1 * (N2 - J - 1)], sa,
                                std::make_integer_sequence<int, G::HI>{})), ...);
}
// Inverse stage A, even half: acc[m] = c0 + sum_{k=1..HI} c_k cos(2 pi k m / N2) (+ c_H cos(pi m)); c_k = row[k]
// scaled (1 / N at k = 0 and H, 2 / N else).
template <typename G, typename T, int... K>
__device__ __forceinline__ void inv_cos_stream(const T* row, T rn, T rn2, T (&acc)[G::H + 1], const T (&ca)[G::H + 1],
                                               std::integer_sequence<int, K...>) {
    cos_axpy<G, 0, true>(acc, row[0] * rn, ca, std::make_integer_sequence<int, G::H + 1>{});
    ((cos_axpy<G, K + 1, false>(acc, row[K + 1] * rn2, ca, std::make_integer_sequence<int, G::H + 1>{})), ...);
    if constexpr (G::EV2) cos_axpy<G, G::H, false>(acc, row[G::H] * rn, ca, std::make_integer_sequence<int, G::H + 1>{});
}
// Inverse stage A, odd half: acc[m] = sum_{k=1..HI} s_k sin(2 pi k m / N2), s_k = row[H + k] 2 / N.
template <typename G, typename T, int K0, int... K>
__device__ __forceinline__ void inv_sin_stream(const T* row, T rn2, T (&acc)[G::H + 1], const T (&sa)[G::H + 1],
                                               std::integer_sequence<int, K0, K...>) {
    sin_axpy<G, 1, true>(acc, row[G::H + 1] * rn2, sa, std::make_integer_sequence<int, G::HI>{});
    ((sin_axpy<G, K + 1, false>(acc, row[G::H + K + 1] * rn2, sa, std::make_integer_sequence<int, G::HI>{})), ...);
}

// The pass body for work-group (bx0, by0) of an nbx x nby grid.  A group has one output whose terms are one per
// source (term e on source e); every term a d/dx, d2/dx2 or -d/dx multiplier.
// FILT: the Fourier filter's passes (ff.m:7-11).  A group has one source and one output whose term is the filter's
// real part (kind 4: the even part of the keep-mask, a spectral multiplier; plus the line's first value times the
// operator's row sum at the output); the x-pass group's second output (gmode 1) takes the spectrum's coefficients at
// the unpaired masked modes into per-block partials of the 4 x 4 slice matrix G, the y-pass output (gmode 2) adds
// U_x G of its line at those modes before the inverse transform (the rank-4 cross term of the mask's odd part).
template <typename T, int DIR, int N1, int N2, typename R, bool FILT = false>
__device__ __forceinline__ void spec1_body(const LPass* __restrict__ Pp, const Members M, const int xcd_order,
                                           const int bx0, const int nbx, const int by0, const int nby) {
    typedef SpecGeom<N1, N2> G;
    constexpr int LB = 16, N = G::N, H = G::H, HI = G::HI;
    constexpr int PF = (N + 15) / 16;                 // samples per thread and line
    static_assert(G::HI >= 1, "stage A needs a complex column");
    static_assert(LB * N1 <= 128, "one stage-A task per lane of two waves");
    const int lin0 = bx0 + nbx * by0;
    int bxl, gy;
    if (xcd_order) {
        const int lin = xcd_block(lin0, nbx * nby);
        bxl = lin / nby;
        gy = lin - bxl * nby;
    } else {
        bxl = bx0;
        gy = by0;
    }
    const MemberBlock mb = member_block_at(M, bxl, nbx);
    const LPass& P = Pp[mb.m];
    if (gy >= P.ngroups) return;
    if (__builtin_nontemporal_load(&P.flags[0]) != 0) return;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int S = P.S;
    T* Rg = reinterpret_cast<T*>(smem_raw);           // LB lines (row stride S), transformed in place
    const LGroup& Gr = P.g[gy];
    const LOut& O = Gr.out[0];
    // the filter's cross term: G and its per-wave partials (80 doubles) after the line region, then the lines'
    // coefficients at the unpaired modes (LB x 4)
    double* const GL = reinterpret_cast<double*>(smem_raw + spec1_cross_offset(S, (int)sizeof(T)));
    T* const DL = reinterpret_cast<T*>(GL + 80);
    const LOut* const Og = (FILT && DIR == 0 && Gr.nout == 2 && Gr.out[1].gmode == 1) ? &Gr.out[1] : nullptr;
    const bool gadd = FILT && DIR == 1 && O.gmode == 2;
    // compute roles by a rotated thread index vt: a work-group of odd index hands waves 2-3 the roles of waves
    // 0-1 and the reverse, so the heavier roles (the even half of stage A, stage B's second round) fall on
    // every SIMD of a CU in turn instead of always on the SIMDs of waves 0-1
    const int rot = ((bx0 + by0) & 1) * 128;
    const int wave = __builtin_amdgcn_readfirstlane((((int)threadIdx.x + rot) & (LP_THREADS - 1)) >> 6);
    const bool odd = wave >= 2;                       // stage A: waves 0-1 the even half, waves 2-3 the odd half
    const bool wact = 64 * (wave & 1) < LB * N1;      // the wave holds stage-A tasks
    const int nl = P.ny * P.nz;
    const int tile0 = mb.bx * P.tpb;
    const int ntile = min(P.tpb, P.ntiles - tile0);
    const int nsrc = Gr.ntiles;
    const T p = (T)P.sp;
    // twiddles: per-lane reads, through a global-address-space pointer (a generic one is read with flat loads,
    // which also count against the LDS wait counter, and is re-read after every LDS store it might alias)
    typedef const __attribute__((address_space(1))) T gT;
    gT* __restrict__ const twg = (gT*)(static_cast<const T*>(P.st) + G::O_TW);

    // Every phase forms its per-thread indices from a fresh copy of the thread index (an empty asm statement
    // the compiler cannot see through), so they are computed where used instead of being hoisted out of the
    // unit loop and kept live, or spilled, across the transforms.
    auto fresh = [](int v) {
        asm volatile("" : "+v"(v));
        return v;
    };
    // staging map: x-lines: line ut/16, samples ut%16 + 16 r; y-lines: x offset ut%16, samples ut/16 + 16 r
    auto row_of = [](int ut) { return DIR == 0 ? (ut >> 4) : (ut & 15); };
    auto sam_of = [](int ut) { return DIR == 0 ? (ut & 15) : (ut >> 4); };
    // element index of sample jt + 16 r of the thread's line: b = the line's first element + jt's offset
    const uint32_t estep = DIR == 0 ? 16u : 16u * (uint32_t)P.rp;
    auto eidx = [&](bool lok, int jt, uint32_t b, int r) -> uint32_t {
        return (lok && jt + 16 * r < N) ? b + estep * (uint32_t)r : Z_OOB;
    };

    // ---- stage A in place (lane task: line la, sub-sequence n1a): every task's reads precede (a barrier) every
    // task's writes
    auto fwd_a = [&]() {
        const int ut = fresh((int)threadIdx.x);
        const int ta = 64 * (wave & 1) + (ut & 63);     // (the lane is unchanged by the rotation)
        const bool act = ta < LB * N1;
        const int la = act ? ta % LB : 0, n1a = act ? ta / LB : 0;
        const T* row = Rg + la * S + n1a;
        T acc[H + 1];
        const cptr<T> tb = phase_table<T>(P.st);
        if (!odd) {
            T ca[H + 1];
#pragma unroll
            for (int m = 0; m <= H; ++m) ca[m] = tb[G::O_CA + m];
            if (wact) fwd_cos_stream<G>(row, acc, ca, std::make_integer_sequence<int, HI>{});
        } else {
            T sa[H + 1];
#pragma unroll
            for (int m = 0; m <= H; ++m) sa[m] = tb[G::O_SA + m];
            if (wact) fwd_sin_stream<G>(row, acc, sa, std::make_integer_sequence<int, HI>{});
        }
        __syncthreads();
        if (act) {
            T* dst = Rg + la * S + n1a * N2;
            if (!odd) {
#pragma unroll
                for (int k = 0; k <= H; ++k) dst[k] = acc[k];
            } else {
#pragma unroll
                for (int k = 1; k <= HI; ++k) dst[H + k] = -acc[k];
            }
        }
        __syncthreads();
    };

    auto inv_a = [&](T* Rb) {
        const int ut = fresh((int)threadIdx.x);
        const int ta = 64 * (wave & 1) + (ut & 63);     // (the lane is unchanged by the rotation)
        const bool act = ta < LB * N1;
        const int la = act ? ta % LB : 0, n1a = act ? ta / LB : 0;
        const T rn = (T)(1.0 / N), rn2 = (T)(2.0 / N);
        T* row = Rb + la * S + n1a * N2;
        T acc[H + 1];
        const cptr<T> tb = phase_table<T>(P.st);
        if (!odd) {
            T ca[H + 1];
#pragma unroll
            for (int m = 0; m <= H; ++m) ca[m] = tb[G::O_CA + m];
            if (wact) inv_cos_stream<G>(row, rn, rn2, acc, ca, std::make_integer_sequence<int, HI>{});
        } else {
            T sa[H + 1];
#pragma unroll
            for (int m = 0; m <= H; ++m) sa[m] = tb[G::O_SA + m];
            if (wact) inv_sin_stream<G>(row, rn2, acc, sa, std::make_integer_sequence<int, HI>{});
        }
        // the line back in natural order, y[n1 + N1 n2] = E[n2] - O[n2] (n2 <= HI) and E[m] + O[m] at
        // n2 = N2 - m: the even half is written to both positions, then the odd half is subtracted and
        // added (every address a compile-time offset from the task's first sample)
        __syncthreads();
        T* y = Rb + la * S + n1a;
        if (act && !odd) {
#pragma unroll
            for (int m = 0; m <= H; ++m) {
                y[N1 * m] = acc[m];
                if (m >= 1 && N2 - m != m) y[N1 * (N2 - m)] = acc[m];
            }
        }
        __syncthreads();
        if (act && odd) {
#pragma unroll
            for (int m = 1; m <= HI; ++m) {
                y[N1 * m] -= acc[m];
                y[N1 * (N2 - m)] += acc[m];
            }
        }
        __syncthreads();
    };

    // ---- stage B fused with the multiplier and its inverse: column task (line l, k2), generic columns
    // k2 = 1 .. HI then the real columns k2 = 0 (and H for even N2), in place
    constexpr int NTB = LB * (HI + G::NSP);
    // the filter's real part at DFT index k: 1 - (mask(k) + mask(-k)) / 2, or its masked part -(mask(k) +
    // mask(-k)) / 2 in the correction form (the output then adds the source itself)
    auto fmul = [&](int k) -> T {
        const int kn = k == 0 ? 0 : N - k;
        const T m = (T)0.5 * (T)((k >= P.mlo && k <= P.mhi) + (kn >= P.mlo && kn <= P.mhi));
        return P.fcorr ? -m : (T)1 - m;
    };
    auto mul_b = [&](int kind, T kap, int k, T xr, T xi, T& yr, T& yi) {
        if (FILT && kind == 4) {
            const T m = fmul(k);
            yr = m * xr; yi = m * xi;
        } else if (kind == 2) {
            const T m = -kap * kap;
            yr = m * xr; yi = m * xi;
        } else {
            const T kk = kind == 1 ? kap : -kap;
            yr = -kk * xi; yi = kk * xr;
        }
    };
    // kind: the multiplier of the output
    auto fused_b = [&](int kind0) {
        const cptr<T> tb = phase_table<T>(P.st);
        T cb[N1], sb[N1];
#pragma unroll
        for (int m = 0; m < N1; ++m) { cb[m] = tb[G::O_CB + m]; sb[m] = tb[G::O_SB + m]; }
        for (int t = (fresh((int)threadIdx.x) + rot) & (LP_THREADS - 1); t < NTB; t += LP_THREADS) {
            const int l = t % LB, c = t / LB;
            T ar[N1], ai[N1], fr[N1], fi[N1], zr[N1], zi[N1];
            if (c < HI) {
                const int k2 = 1 + c;
                const int off = l * S + k2;
                const T* col = Rg + off;
                gT* tw = twg + 2 * k2 * N1;
                T tc[N1], ts[N1];                         // exp(-2 pi i n1 k2 / N) = tc - i ts, read once per task
#pragma unroll
                for (int n1 = 0; n1 < N1; ++n1) { tc[n1] = tw[2 * n1]; ts[n1] = tw[2 * n1 + 1]; }
#pragma unroll
                for (int n1 = 0; n1 < N1; ++n1) {
                    const T xr = col[n1 * N2], xi = col[n1 * N2 + H], cs = tc[n1], sn = ts[n1];
                    ar[n1] = fmat(xr, cs, xi * sn);          // (xr + i xi)(c - i s)
                    ai[n1] = fmat(xi, cs, -(xr * sn));
                }
                dft_n1<G, -1>(ar, ai, cb, sb, fr, fi);
                if (FILT && Og) {
                    // the coefficients (Re X[kq], -Im X[kq]) of the unpaired modes, from kq or its conjugate N - kq
                    // (exactly one of the two lies on a complex column)
                    for (int q = 0; q < P.nkp; ++q) {
                        const int kq = P.kp[q];
#pragma unroll
                        for (int k1 = 0; k1 < N1; ++k1) {
                            const int k = k2 + N2 * k1;
                            if (k == kq) { DL[l * 4 + 2 * q] = fr[k1]; DL[l * 4 + 2 * q + 1] = -fi[k1]; }
                            else if (k == N - kq) { DL[l * 4 + 2 * q] = fr[k1]; DL[l * 4 + 2 * q + 1] = fi[k1]; }
                        }
                    }
                }
                // multiplier, inverse DFT, twiddle back, into region dst; the generic columns' wavenumbers:
                // k = k2 + N2 k1 wraps to k - N exactly when 2 k1 >= N1
                auto back = [&](int kind, T* dst) {
#pragma unroll
                    for (int k1 = 0; k1 < N1; ++k1)
                        mul_b(kind, p * (T)(k2 + N2 * k1 - (2 * k1 >= N1 ? N : 0)), k2 + N2 * k1, fr[k1], fi[k1], ar[k1],
                              ai[k1]);
                    if (FILT && gadd) {
                        // the cross term (ke / 2) (c_1 + i c_0) at mode kq and its conjugate at N - kq, c = the
                        // line's U_x G row
                        for (int q = 0; q < P.nkp; ++q) {
                            const int kq = P.kp[q];
                            const T hc = (T)(0.5 * P.ke[q]);
                            const T dr = hc * DL[l * 4 + 2 * q + 1], di = hc * DL[l * 4 + 2 * q];
#pragma unroll
                            for (int k1 = 0; k1 < N1; ++k1) {
                                const int k = k2 + N2 * k1;
                                if (k == kq) { ar[k1] += dr; ai[k1] += di; }
                                else if (k == N - kq) { ar[k1] += dr; ai[k1] -= di; }
                            }
                        }
                    }
                    dft_n1<G, 1>(ar, ai, cb, sb, zr, zi);
#pragma unroll
                    for (int n1 = 0; n1 < N1; ++n1) {
                        const T cs = tc[n1], sn = ts[n1];  // times exp(+2 pi i n1 k2 / N)
                        dst[n1 * N2] = fmat(zr[n1], cs, -(zi[n1] * sn));
                        dst[n1 * N2 + H] = fmat(zi[n1], cs, zr[n1] * sn);
                    }
                };
                back(kind0, Rg + off);

            } else {
                const int k2 = c > HI ? H : 0;
                const int off = l * S + k2;
                const T* col = Rg + off;
                gT* tw = twg + 2 * k2 * N1;
                T tc[N1], ts[N1];
#pragma unroll
                for (int n1 = 0; n1 < N1; ++n1) { tc[n1] = tw[2 * n1]; ts[n1] = tw[2 * n1 + 1]; }
#pragma unroll
                for (int n1 = 0; n1 < N1; ++n1) {
                    const T xr = col[n1 * N2];
                    ar[n1] = xr * tc[n1];
                    ai[n1] = -(xr * ts[n1]);
                }
                dft_n1<G, -1>(ar, ai, cb, sb, fr, fi);
                auto back = [&](int kind, T* dst) {
#pragma unroll
                    for (int k1 = 0; k1 < N1; ++k1) {
                        if (FILT && kind == 4) {
                            const T m = fmul(k2 + N2 * k1);
                            ar[k1] = m * fr[k1]; ai[k1] = m * fi[k1];
                        } else {
                            spec_mul(kind, p, k2 + N2 * k1, N, fr[k1], fi[k1], ar[k1], ai[k1]);
                        }
                    }
                    dft_n1<G, 1>(ar, ai, cb, sb, zr, zi);
#pragma unroll
                    for (int n1 = 0; n1 < N1; ++n1) dst[n1 * N2] = fmat(zr[n1], tc[n1], -(zi[n1] * ts[n1]));
                };
                back(kind0, Rg + off);

            }
        }
        __syncthreads();
    };

    // store of output Ou from region Rr: the transform (+ hold: the addend and, for a two-source output, the first
    // source's result); the NaN flag from the stored values when the output checks them
    struct Line { bool ok; uint32_t e0, b; int mrow, jt; };
    // cst: the filter's line constant (its first value times the row sum of the output's operator)
    auto store = [&](const LOut& Ou, const T* Rr, const Line& L, const T (&hold)[PF], T cst) {
        bool bad = false;
        auto st = [&](auto tag) {
            using S_ = decltype(tag);
            const Rs rs = rs_make(Ou.out, fbytes<S_>(P));
#pragma unroll
            for (int r = 0; r < PF; ++r) {
                const T y = Rr[L.mrow * S + min(L.jt + 16 * r, N - 1)];
                const T v = FILT ? (y + cst) + hold[r] : y + hold[r];
                const uint32_t ei = eidx(L.ok, L.jt, L.b, r);
                bstr<S_>(rs, ei, (S_)v);
                bad |= ei != Z_OOB && v != v;
            }
        };
        if (Ou.out_dt == DT_F64) st(double{});
        else st(float{});
        if (Ou.nan_check && __any(bad) && (threadIdx.x & 63) == 0) {
            const int step = __atomic_load_n(&P.flags[1], __ATOMIC_RELAXED);
            atomicCAS(&P.flags[0], 0, step > 0 ? step : 1);
        }
    };
    // an output's addend into hold (zeros without one); partial: the output field itself, which holds the
    // first source's result of a two-source output (stored by the same threads at the same elements)
    auto load_addend = [&](const LOut& Ou, const Line& L, T (&hold)[PF], bool partial) {
        auto ld = [&](auto tag) {
            using S_ = decltype(tag);
            const Rs rs = rs_make(partial ? Ou.out : Ou.addend, fbytes<S_>(P));
#pragma unroll
            for (int r = 0; r < PF; ++r) hold[r] = (T)bldr<S_>(rs, eidx(L.ok, L.jt, L.b, r));
        };
        const void* a = partial ? Ou.out : Ou.addend;
        const int dt = partial ? Ou.out_dt : Ou.add_dt;
        if (!a) {
#pragma unroll
            for (int r = 0; r < PF; ++r) hold[r] = (T)0;
        } else if (dt == DT_F64) {
            ld(double{});
        } else {
            ld(float{});
        }
    };

    // The tile loop.  A unit is one source of a tile: it is staged, transformed forward, multiplied, transformed
    // back and its output stored.  A two-source output stores its first source's result and reads it back as the
    // second source's addend, so both sources run the same code; the loop over sources stays rolled, so every
    // transform phase is compiled once and a kernel's code fits the instruction cache.  An output's addend is
    // loaded just before its store, so it occupies no registers during the transforms.
    for (int it = 0; it < ntile; ++it) {
        const int lt = tile0 + it;
        int line0, z, x0;
        if (DIR == 0) {
            line0 = lt * LB;
            z = line0 / P.ny;
            x0 = 0;
        } else {
            line0 = 0;
            x0 = (lt % P.nxc) * LB;
            z = lt / P.nxc;
        }
        // the thread's line: valid, its first element, and the offset of its sample jt
        auto line_at = [&](int ut) {
            Line L;
            L.mrow = row_of(ut);
            L.jt = sam_of(ut);
            L.ok = DIR == 0 ? line0 + L.mrow < nl : x0 + L.mrow < P.nx;
            L.e0 = DIR == 0 ? (uint32_t)xline_base_at(P, line0, z, L.mrow) : (uint32_t)(x0 + L.mrow + P.ps * z);
            L.b = L.e0 + (DIR == 0 ? (uint32_t)L.jt : (uint32_t)L.jt * (uint32_t)P.rp);
            return L;
        };
        // hold: the output's addend (for a two-source output's second source, the first source's stored result)
        T hold[PF];
#pragma unroll 1
        for (int t = 0; t < nsrc; ++t) {
            // the source's lines, relative to their first value
            T cst = (T)0;
            {
                const Line L = line_at(fresh((int)threadIdx.x));
                R pa[PF], pref;
                auto load = [&](auto tag) {
                    using S_ = decltype(tag);
                    const Rs rs = rs_make(Gr.src[t], fbytes<S_>(P));
                    pref = (R)bldr<S_>(rs, L.ok ? L.e0 : Z_OOB);
#pragma unroll
                    for (int r = 0; r < PF; ++r) pa[r] = (R)bldr<S_>(rs, eidx(L.ok, L.jt, L.b, r));
                };
                if (Gr.src_dt[t] == DT_F64) load(double{});
                else load(float{});
                if constexpr (FILT) {
                    int rsum = 0;
                    for (int e = 0; e < O.nterms; ++e) rsum += O.term[e].rsum;
                    cst = rsum ? (T)rsum * (T)pref : (T)0;
                }
                __syncthreads();                 // the previous unit is done with the line region
#pragma unroll
                for (int r = 0; r < PF; ++r) {
                    const int j = L.jt + 16 * r;
                    if (j < N) Rg[L.mrow * S + j] = (T)(pa[r] - pref);
                }
                __syncthreads();
            }
            // the y-pass filter's cross term: U_x G of the tile's lines into DL (ends with a barrier); the x-pass's
            // coefficients of absent unpaired modes are zero (stage B writes those of the present ones)
            if (FILT && gadd) fold_lr_from_g<T, 1>(P, O, DL, GL, x0, z);
            if (FILT && Og && P.nkp < 2)
                for (int idx = (int)threadIdx.x; idx < LB * 4; idx += LP_THREADS)
                    if ((idx & 3) >> 1 >= P.nkp) DL[idx] = (T)0;
            fwd_a();
            int kind = 1;
            for (int e = 0; e < O.nterms; ++e)            // (a filter output's rank-4 term is its cross term)
                if (O.term[e].tile == t && !O.term[e].lr) kind = O.term[e].sk;
            fused_b(kind);
            // the x-pass filter's cross term: this 16-line block's partials of G from the coefficients in DL
            if (FILT && Og) fold_g_partials<T, 1>(P, *Og, DL, line0, nl);
            inv_a(Rg);
            const Line L = line_at(fresh((int)threadIdx.x));
            // (the barriers since the first source's store order it before this read of the same elements)
            load_addend(O, L, hold, t > 0);
            store(O, Rg, L, hold, cst);
        }
    }
}

template <typename T, int DIR, int N1, int N2, typename R, bool FILT>
__global__ __launch_bounds__(LP_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_spec1_pass(
    const LPass* __restrict__ Pp, const Members M, const int xcd_order) {
    spec1_body<T, DIR, N1, N2, R, FILT>(Pp, M, xcd_order, (int)blockIdx.x, (int)gridDim.x, (int)blockIdx.y,
                                        (int)gridDim.y);
}

// The constraint's x- and y-pass (one source and one derivative output each, independent of each other) in one
// launch: the first nb0 work-groups run the x-pass, the y-pass's nb1 start at nb0 rounded up to a multiple of 8
// (the work-groups between return at once), so both passes keep the XCD assignment (work-group i on XCD i mod 8)
// their own launches would have.
template <typename T, int N1, int N2>
__global__ __launch_bounds__(LP_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_spec_pair(
    const LPass* __restrict__ X, const LPass* __restrict__ Y, const int nb0, const int nb1, const int order0,
    const int order1) {
    const Members one{nullptr, 1};
    const int b = (int)blockIdx.x, nb0p = (nb0 + 7) & ~7;
    if (b < nb0) spec1_body<T, 0, N1, N2, T>(X, one, order0, b, nb0, 0, 1);
    else if (b >= nb0p) spec1_body<T, 1, N1, N2, T>(Y, one, order1, b - nb0p, nb1, 0, 1);
}

template <typename T, int N1, int N2>
hipError_t spec_pair_launch_one(const LPass* x, const LPass* y, int nb0, int nb1, size_t lds, hipStream_t s, int o0,
                                int o1) {
    const void* fn = reinterpret_cast<const void*>(&k_spec_pair<T, N1, N2>);
    if (lds > 64 * 1024) {
        static bool done = false;
        if (!done) {
            const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            if (e != hipSuccess) return e;
            done = true;
        }
    }
    hipLaunchKernelGGL((k_spec_pair<T, N1, N2>), dim3((unsigned)(((nb0 + 7) & ~7) + nb1)), dim3(LP_THREADS), lds, s, x, y, nb0,
                       nb1, o0, o1);
    return hipGetLastError();
}

// The constraint pair of one precision in the factored-DFT form (both passes of factorization fid); explicitly
// instantiated by hqt_spec_pair.hip.
template <typename T>
hipError_t spec_pair_launch(const LPass* x, const LPass* y, int nb0, int nb1, int fid, size_t lds, hipStream_t s,
                            int o0, int o1) {
    switch (fid) {
        case 0: return spec_pair_launch_one<T, 7, 43>(x, y, nb0, nb1, lds, s, o0, o1);
        case 1: return spec_pair_launch_one<T, 4, 8>(x, y, nb0, nb1, lds, s, o0, o1);
        case 2: return spec_pair_launch_one<T, 8, 8>(x, y, nb0, nb1, lds, s, o0, o1);
        case 3: return spec_pair_launch_one<T, 8, 12>(x, y, nb0, nb1, lds, s, o0, o1);
        case 4: return spec_pair_launch_one<T, 8, 16>(x, y, nb0, nb1, lds, s, o0, o1);
        case 5: return spec_pair_launch_one<T, 8, 24>(x, y, nb0, nb1, lds, s, o0, o1);
        case 6: return spec_pair_launch_one<T, 8, 32>(x, y, nb0, nb1, lds, s, o0, o1);
        default: return hipErrorInvalidValue;
    }
}

template <typename T, int DIR, int N1, int N2, typename R, bool FILT>
hipError_t spec1_launch_one(const LPass* d, Members M, dim3 grid, size_t lds, hipStream_t s, int order) {
    const void* fn = reinterpret_cast<const void*>(&k_spec1_pass<T, DIR, N1, N2, R, FILT>);
    if (lds > 64 * 1024) {
        static bool done = false;
        if (!done) {
            const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            if (e != hipSuccess) return e;
            done = true;
        }
    }
    hipLaunchKernelGGL((k_spec1_pass<T, DIR, N1, N2, R, FILT>), grid, dim3(LP_THREADS), lds, s, d, M, order);
    return hipGetLastError();
}

template <typename T, int DIR, typename R, bool FILT>
hipError_t spec1_launch_r(const LPass* d, Members M, dim3 grid, int fid, size_t lds, hipStream_t s, int order) {
    switch (fid) {
        case 0: return spec1_launch_one<T, DIR, 7, 43, R, FILT>(d, M, grid, lds, s, order);
        case 1: return spec1_launch_one<T, DIR, 4, 8, R, FILT>(d, M, grid, lds, s, order);
        case 2: return spec1_launch_one<T, DIR, 8, 8, R, FILT>(d, M, grid, lds, s, order);
        case 3: return spec1_launch_one<T, DIR, 8, 12, R, FILT>(d, M, grid, lds, s, order);
        case 4: return spec1_launch_one<T, DIR, 8, 16, R, FILT>(d, M, grid, lds, s, order);
        case 5: return spec1_launch_one<T, DIR, 8, 24, R, FILT>(d, M, grid, lds, s, order);
        case 6: return spec1_launch_one<T, DIR, 8, 32, R, FILT>(d, M, grid, lds, s, order);
        default: return hipErrorInvalidValue;
    }
}

// Every single-region factored-DFT instance of one (precision, direction, RK-stage or filter) family; explicitly
// instantiated by the hqt_spec*.hip translation units.  wide: a double source in an FP32 pass (prefetch registers
// in double).
template <typename T, int DIR, bool FILT>
hipError_t spec1_launch_td(const LPass* d, Members M, dim3 grid, int fid, size_t lds, hipStream_t s, int order,
                           int wide) {
    if constexpr (sizeof(T) == 8) {
        (void)wide;
        return spec1_launch_r<T, DIR, double, FILT>(d, M, grid, fid, lds, s, order);
    } else {
        if (wide) return spec1_launch_r<T, DIR, double, FILT>(d, M, grid, fid, lds, s, order);
        return spec1_launch_r<T, DIR, float, FILT>(d, M, grid, fid, lds, s, order);
    }
}

#endif

}  // namespace hqt
