// Factored-DFT line passes for gfx950 (MI355X): the x/y spectral derivatives of the RK stages
// applied in Fourier space through a two-stage (Cooley-Tukey) real DFT of every line.
//
// Reference: tevol_driven_cluster.m:60-61 (func_Lx / func_Ly: the dense fourdif matrix times every
// line) and third_party/dmsuite/fourdif.m.  The fourdif matrix Lx = px Dx is the circulant whose
// eigenvalues are i px k (k the signed wavenumber, 0 at the Nyquist mode of even N), and Lx Lx (the
// reference's second derivative, func_Lx applied twice) has -(px k)^2.  So d_x x = IDFT(i px k DFT(x))
// exactly: this pass computes each source line's DFT once, multiplies it by the eigenvalues of each
// output's operator and transforms back, instead of a dense (or even/odd folded) matrix product.
//
// The DFT of length N = N1 N2 is split (decimation in time): with n = n1 + N1 n2, k = k2 + N2 k1,
//     X[k2 + N2 k1] = sum_n1 W_N1^{n1 k1} W_N^{n1 k2} A[n1, k2],    A[n1, k2] = sum_n2 x[n1 + N1 n2] W_N2^{n2 k2}
// (W_M = exp(-2 pi i / M)).  Stage A is a real DFT of length N2 of each of the N1 sub-sequences
// (only k2 = 0 .. N2/2 kept: the rest is conjugate), computed in its even/odd folded form; stage B is
// a complex DFT of length N1 per (line, k2) after the twiddle W_N^{n1 k2}.  The inverse runs the two
// stages backwards and ends in a real inverse DFT of length N2 (the stage-B output is Hermitian in k2
// for a real result).  At N = 301 = 7 x 43 a transform costs about 26 multiply-adds per point against
// 151 for one folded matrix application, so the passes are bound by their HBM traffic.
//
// Both stages run on the VALU (on MI355X FP64 vector FMA has the same peak as FP64 MFMA, and the
// stages' small odd sizes would waste much of a 16 x 16 x 4 tile): a lane owns one stage-A
// sub-sequence (line, n1) or one stage-B column (line, k2), with the trigonometric constants of the
// compile-time sizes in registers.  Work decomposition (as the folded passes): a work-group owns a
// tile of 16 lines and one output group; a unit is one (tile, source), staged into LDS once and
// transformed in place; every output term on it then runs the spectral multiply and inverse stage B
// into a second LDS region (summing the terms of a two-source output), and an output is transformed
// back and written as soon as its last source is in.
//
// Each line enters relative to its first value x_0 (the deviation x - x_0 is formed in the source's
// own precision at staging): the derivatives annihilate constants, and in FP32 this keeps the
// rounding of the transforms from acting on a line's mean.
//
// The Fourier filter (ff.m:7-11, w = real(ifft2(b)) with the masked modes of b = fft2(x) zeroed) runs
// through the same passes.  For a real field and the separable keep-mask a(kx) b(ky) (the complement of
// fx | fy, tevol_driven_cluster.m:27-31), real(ifft2(a b fft2(x))) keeps the even part of the mask:
// w = (A_e (x) B_e) x + (A_o (x) B_o) x, a_e(k) = (a(k) + a(-k)) / 2, a_o = (a(k) - a(-k)) / 2.  A_e is a
// spectral multiplier (term kind 4, the line's x_0 times its row sum a(0) added back at the output).
// a_o is nonzero only on the unpaired masked modes (two for the reference's masks), so the second
// product is rank 4 per direction: the x-pass reads the four real DFT coefficients of those modes off
// each line's spectrum and writes per-block partials of the 4 x 4 slice matrix G (as the folded
// passes do); the y-pass forms U_x G per line and adds it to the spectrum at the same modes before
// the inverse transform.
#pragma once
#include <utility>

#include "hqt_fold_core.h"
#include "hqt_line_core.h"

namespace hqt {

// Line lengths with a factored-DFT pass: (N1, N2), N = N1 N2.
constexpr int kSpecFact[][2] = {{7, 43}, {4, 8}, {8, 8}, {8, 12}, {8, 16}, {8, 24}, {8, 32}};
constexpr int kSpecNF = (int)(sizeof(kSpecFact) / sizeof(kSpecFact[0]));

// Index into kSpecFact of a line length, or -1.
inline int spec_fact_id(int N) {
    for (int i = 0; i < kSpecNF; ++i)
        if (kSpecFact[i][0] * kSpecFact[i][1] == N) return i;
    return -1;
}

// Sizes of a factorization.  Stage-A slots of a sub-sequence: Re A[k2] at k2 (0 .. H), Im A[k2] at
// H + k2 (1 .. HI), N2 slots in all; the same slots hold the spectrum after stage B and, in the second
// region, the inverse's stage-B output and then its folded halves E (0 .. H) and O (H + 1 ..).
template <int N1_, int N2_> struct SpecGeom {
    static constexpr int N1 = N1_, N2 = N2_, N = N1_ * N2_;
    static constexpr int H = N2_ / 2, HI = (N2_ - 1) / 2;        // k2 <= H kept; 1 .. HI complex
    static constexpr int H1 = N1_ / 2, HI1 = (N1_ - 1) / 2;
    static constexpr bool EV1 = (N1_ % 2) == 0, EV2 = (N2_ % 2) == 0;
    static constexpr int NSP = EV2 ? 2 : 1;                       // real columns: k2 = 0 (and k2 = H)
    // trig table: cos, sin (2 pi m / N2), m = 0 .. H | cos, sin (2 pi m / N1), m < N1 | twiddles
    // cos, sin (2 pi n1 k2 / N) for k2 = 0 .. H, n1 < N1 (k2-major, pairs)
    static constexpr int O_CA = 0, O_SA = H + 1, O_CB = 2 * (H + 1), O_SB = 2 * (H + 1) + N1_,
                         O_TW = 2 * (H + 1) + 2 * N1_, TAB = O_TW + 2 * (H + 1) * N1_;
};

// Entries of the trig table of a factorization (host: table size).
inline int spec_table_len(int fid) {
    const int N1 = kSpecFact[fid][0], N2 = kSpecFact[fid][1], H = N2 / 2;
    return 2 * (H + 1) + 2 * N1 + 2 * (H + 1) * N1;
}

// Byte offset of the filter cross-term area (80 doubles, then 16 x 4 values of the pass precision) in the
// dynamic LDS of a factored-DFT pass: after the two line regions and the trig table, 8-byte aligned.
__host__ __device__ inline int spec_cross_offset(int S, int tab, int es) {
    return ((2 * 16 * S + tab) * es + 7) & ~7;
}
// Bytes of that area.
__host__ __device__ inline int spec_cross_bytes(int es) { return 80 * 8 + 16 * 4 * es; }

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)

// Fused multiply-add in the pass precision (__builtin_fma alone would widen float operands to double).
template <typename T> __device__ __forceinline__ T fmat(T a, T b, T c) {
    if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
    else return __builtin_fma(a, b, c);
}

// cos / sin (2 pi m / N2) for a compile-time m in [0, N2) from the H + 1 stored values.
template <typename G, typename T> __device__ __forceinline__ T cos_m(const T (&ca)[G::H + 1], int m) {
    return m <= G::H ? ca[m] : ca[G::N2 - m];
}
template <typename G, typename T> __device__ __forceinline__ T sin_m(const T (&sa)[G::H + 1], int m) {
    return m <= G::H ? sa[m] : -sa[G::N2 - m];
}

// out[k] = sum_{j=0..H} in[j] cos(2 pi j k / N2), k = 0 .. H  (the even half of a real DFT of length N2),
// and out[k] = sum_{j=1..HI} in[j] sin(2 pi j k / N2), k = 1 .. HI (the odd half).  Expanded by fold
// expressions over index sequences: every trig index (j k mod N2) is a constant expression, so the
// constants are plain register operands and no loop is left for the compiler to keep rolled.
template <typename G, int K, typename T, int... J>
__device__ __forceinline__ T cos_dot(const T (&in)[G::H + 1], const T (&ca)[G::H + 1], std::integer_sequence<int, J...>) {
    T acc = in[0];
    ((acc = fmat(in[J + 1], cos_m<G>(ca, ((J + 1) * K) % G::N2), acc)), ...);
    return acc;
}
template <typename G, typename T, int... K>
__device__ __forceinline__ void cos_tf_k(const T (&in)[G::H + 1], const T (&ca)[G::H + 1], T (&out)[G::H + 1],
                                         std::integer_sequence<int, K...>) {
    ((out[K] = cos_dot<G, K>(in, ca, std::make_integer_sequence<int, G::H>{})), ...);
}
template <typename G, typename T>
__device__ __forceinline__ void cos_tf(const T (&in)[G::H + 1], const T (&ca)[G::H + 1], T (&out)[G::H + 1]) {
    cos_tf_k<G>(in, ca, out, std::make_integer_sequence<int, G::H + 1>{});
}
template <typename G, int K, typename T, int... J>
__device__ __forceinline__ T sin_dot(const T (&in)[G::HI + 1], const T (&sa)[G::H + 1], std::integer_sequence<int, J...>) {
    T acc = (T)0;
    ((acc = fmat(in[J + 1], sin_m<G>(sa, ((J + 1) * K) % G::N2), acc)), ...);
    return acc;
}
template <typename G, typename T, int... K>
__device__ __forceinline__ void sin_tf_k(const T (&in)[G::HI + 1], const T (&sa)[G::H + 1], T (&out)[G::HI + 1],
                                         std::integer_sequence<int, K...>) {
    ((out[K + 1] = sin_dot<G, K + 1>(in, sa, std::make_integer_sequence<int, G::HI>{})), ...);
}
template <typename G, typename T>
__device__ __forceinline__ void sin_tf(const T (&in)[G::HI + 1], const T (&sa)[G::H + 1], T (&out)[G::HI + 1]) {
    sin_tf_k<G>(in, sa, out, std::make_integer_sequence<int, G::HI>{});
}

// Complex DFT of length N1 in its folded form: z[k] = sum_n a[n] exp(sgn 2 pi i n k / N1) (sgn = -1
// forward, +1 inverse).  cb, sb: cos, sin (2 pi m / N1).
template <typename G, int SGN, typename T>
__device__ __forceinline__ void dft_n1(const T (&ar)[G::N1], const T (&ai)[G::N1], const T (&cb)[G::N1],
                                       const T (&sb)[G::N1], T (&zr)[G::N1], T (&zi)[G::N1]) {
    constexpr int N1 = G::N1, HI1 = G::HI1;
    T ur[HI1 + 1], ui[HI1 + 1], vr[HI1 + 1], vi[HI1 + 1];
#pragma unroll
    for (int j = 1; j <= HI1; ++j) {
        ur[j] = ar[j] + ar[N1 - j]; ui[j] = ai[j] + ai[N1 - j];
        vr[j] = ar[j] - ar[N1 - j]; vi[j] = ai[j] - ai[N1 - j];
    }
#pragma unroll
    for (int k = 0; k <= G::H1; ++k) {
        T pr = ar[0], pi = ai[0], qr = (T)0, qi = (T)0;
        if constexpr (G::EV1) {
            const T s = (k & 1) ? (T)-1 : (T)1;
            pr = fmat(s, ar[G::H1], pr);
            pi = fmat(s, ai[G::H1], pi);
        }
#pragma unroll
        for (int j = 1; j <= HI1; ++j) {
            const T c = cb[(j * k) % N1], s = sb[(j * k) % N1];
            pr = fmat(ur[j], c, pr); pi = fmat(ui[j], c, pi);
            qr = fmat(vr[j], s, qr); qi = fmat(vi[j], s, qi);
        }
        // z[k] = P + sgn i Q, z[N1 - k] = P - sgn i Q
        zr[k] = pr - SGN * qi; zi[k] = pi + SGN * qr;
        if (k >= 1 && 2 * k != N1) { zr[N1 - k] = pr + SGN * qi; zi[N1 - k] = pi - SGN * qr; }
    }
}

// Spectral multiplier of a term at wavenumber index k (0 <= k < N): kind 1 d/dx (i p s), 2 d2/dx2
// (-(p s)^2), 3 -d/dx (-i p s); s = the signed wavenumber, 0 at the Nyquist mode of even N.
template <typename T>
__device__ __forceinline__ void spec_mul(int kind, T p, int k, int N, T xr, T xi, T& yr, T& yi) {
    const int sk = 2 * k < N ? k : (2 * k > N ? k - N : 0);
    const T kap = p * (T)sk;
    if (kind == 2) {
        const T m = -kap * kap;
        yr = m * xr; yi = m * xi;
    } else if (kind == 1) {
        yr = -kap * xi; yi = kap * xr;
    } else {
        yr = kap * xi; yi = -kap * xr;
    }
}

// R: type of the prefetch registers (double when a source of the pass is double).  FILT: the filter's
// passes (term kind 4, the cross term's gmode outputs, the line constant at the output); the RK-stage
// passes are compiled without that code.
// The pass body for work-group (bx0, by0) of an nbx x nby grid (the kernels below pass their own block
// indices, or a sub-range of them).
template <typename T, int DIR, int N1, int N2, typename R, bool FILT>
__device__ __forceinline__ void spec_body(const LPass* __restrict__ Pp, const Members M, const int xcd_order,
                                          const int bx0, const int nbx, const int by0, const int nby) {
    typedef SpecGeom<N1, N2> G;
    constexpr int LB = 16, N = G::N, H = G::H, HI = G::HI;
    constexpr int PF = (N + 15) / 16;                 // samples per thread and line
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
    T* Rg = reinterpret_cast<T*>(smem_raw);           // lines / spectrum, LB x S
    T* Bg = Rg + LB * S;                              // inverse stage-B output / E, O halves, LB x S
    T* TW = Bg + LB * S;                              // trig table (SpecGeom layout)
    // filter cross term (passes with gmode outputs only): G and its per-wave partials (80 doubles), then the
    // lines' rank-4 dots or coefficients (LB x 4)
    double* GL = reinterpret_cast<double*>(smem_raw + spec_cross_offset(S, G::TAB, (int)sizeof(T)));
    T* DL = reinterpret_cast<T*>(GL + 80);
    const LGroup& Gr = P.g[gy];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nl = P.ny * P.nz;
    const int tile0 = mb.bx * P.tpb;
    const int ntile = min(P.tpb, P.ntiles - tile0);
    const int nunit = ntile * Gr.ntiles;
    const T p = (T)P.sp;

    // trig constants of the compile-time sizes and the twiddles, into LDS once; each phase loads the ones it
    // uses into registers
    const T* tab = static_cast<const T*>(P.st);
    for (int e = tid; e < G::TAB; e += LP_THREADS) TW[e] = tab[e];
    const T* CA = TW + G::O_CA;
    const T* SA = TW + G::O_SA;
    const T* CB = TW + G::O_CB;
    const T* SB = TW + G::O_SB;
    const T* TWD = TW + G::O_TW;

    // staging map: x-lines: line tid/16, samples tid%16 + 16 r; y-lines: x offset tid%16, samples tid/16 + 16 r
    const int mrow = DIR == 0 ? (tid >> 4) : (tid & 15);
    const int jt = DIR == 0 ? (tid & 15) : (tid >> 4);
    R pa[PF], pref = (R)0;
    auto geom = [&](int lt, int& line0, int& z, int& x0) {
        if (DIR == 0) {
            line0 = lt * LB;
            z = line0 / P.ny;
            x0 = 0;
        } else {
            line0 = 0;
            x0 = (lt % P.nxc) * LB;
            z = lt / P.nxc;
        }
    };
    // element index of sample j of this thread's line (Z_OOB past the grid)
    auto elem = [&](int line0, int z, int x0, int xb, int j) -> uint32_t {
        if (DIR == 0) return (line0 + mrow < nl && j < N) ? (uint32_t)(xb + j) : Z_OOB;
        const int x = x0 + mrow;
        return (x < P.nx && j < N) ? (uint32_t)(x + P.rp * j + P.ps * z) : Z_OOB;
    };
    auto fetch = [&](int unit) {
        const int lt = tile0 + unit / Gr.ntiles, t = unit % Gr.ntiles;
        int line0, z, x0;
        geom(lt, line0, z, x0);
        const int xb = DIR == 0 ? xline_base_at(P, line0, z, mrow) : 0;
        auto load = [&](auto tag) {
            using S_ = decltype(tag);
            const Rs rs = rs_make(Gr.src[t], fbytes<S_>(P));
            pref = (R)bldr<S_>(rs, elem(line0, z, x0, xb, 0));
#pragma unroll
            for (int r = 0; r < PF; ++r) pa[r] = (R)bldr<S_>(rs, elem(line0, z, x0, xb, jt + 16 * r));
        };
        if (Gr.src_dt[t] == DT_F64) load(double{});
        else load(float{});
    };
    // x - x_0 into the line region, formed in the prefetch precision
    auto deposit = [&]() {
#pragma unroll
        for (int r = 0; r < PF; ++r) {
            const int j = jt + 16 * r;
            if (j < N) Rg[mrow * S + j] = (T)(pa[r] - pref);
        }
    };

    // ---- stage A (forward): lane task (line l, sub-sequence n1), t = l + LB n1 < LB N1 <= 128; waves 0-1
    // the even half (cos), waves 2-3 the odd half (sin).  Every sub-sequence is read before any slot is
    // written (the stage-A layout overwrites the interleaved samples in place).
    static_assert(LB * N1 <= 128, "one stage-A task per lane of two waves");
    const bool odd = wave >= 2;
    const int ta = 64 * (wave & 1) + lane;
    const bool act = ta < LB * N1;
    const bool wact = 64 * (wave & 1) < LB * N1;     // wave-uniform: the wave holds a task
    const int la = act ? ta % LB : 0, n1a = act ? ta / LB : 0;
    auto fwd_a = [&]() {
        const T* row = Rg + la * S + n1a;
        T outv[H + 1];
        if (wact) {
            if (!odd) {
                T ca[H + 1];
#pragma unroll
                for (int m = 0; m <= H; ++m) ca[m] = CA[m];
                T f[H + 1];
                f[0] = row[0];
#pragma unroll
                for (int j = 1; j <= HI; ++j) f[j] = row[N1 * j] + row[N1 * (N2 - j)];
                if constexpr (G::EV2) f[H] = row[N1 * H];
                cos_tf<G>(f, ca, outv);
            } else {
                T sa[H + 1], g[HI + 1], o[HI + 1];
#pragma unroll
                for (int m = 0; m <= H; ++m) sa[m] = SA[m];
                g[0] = (T)0;
#pragma unroll
                for (int j = 1; j <= HI; ++j) g[j] = row[N1 * j] - row[N1 * (N2 - j)];
                sin_tf<G>(g, sa, o);
#pragma unroll
                for (int k = 1; k <= HI; ++k) outv[k] = -o[k];
            }
        }
        __syncthreads();
        if (act) {
            T* dst = Rg + la * S + n1a * N2;
            if (!odd) {
#pragma unroll
                for (int k = 0; k <= H; ++k) dst[k] = outv[k];
            } else {
#pragma unroll
                for (int k = 1; k <= HI; ++k) dst[H + k] = outv[k];
            }
        }
        __syncthreads();
    };

    // ---- stage B: column task (line l, k2); the generic columns k2 = 1 .. HI and then the real columns
    // k2 = 0 (and k2 = H for even N2) share one task range, so the lanes left over by the first take the
    // second.  Forward in place; the inverse writes (or, acc, adds to) the second region.
    constexpr int NTB = LB * (HI + G::NSP);
    // the generic columns' wavenumbers: k = k2 + N2 k1 wraps to k - N exactly when 2 k1 >= N1 (k2 < N2 / 2)
    T pc[N1];
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) pc[k1] = p * (T)(N2 * k1 - (2 * k1 >= N1 ? N : 0));
    auto fwd_b = [&]() {
        T cb[N1], sb[N1];
#pragma unroll
        for (int m = 0; m < N1; ++m) { cb[m] = CB[m]; sb[m] = SB[m]; }
        for (int t = tid; t < NTB; t += LP_THREADS) {
            const int l = t % LB, c = t / LB;
            T ar[N1], ai[N1], zr[N1], zi[N1];
            if (c < HI) {
                const int k2 = 1 + c;
                T* col = Rg + l * S + k2;
                const T* tw = TWD + 2 * k2 * N1;
#pragma unroll
                for (int n1 = 0; n1 < N1; ++n1) {
                    const T xr = col[n1 * N2], xi = col[n1 * N2 + H], cs = tw[2 * n1], sn = tw[2 * n1 + 1];
                    ar[n1] = fmat(xr, cs, xi * sn);          // (xr + i xi)(c - i s)
                    ai[n1] = fmat(xi, cs, -(xr * sn));
                }
                dft_n1<G, -1>(ar, ai, cb, sb, zr, zi);
#pragma unroll
                for (int k1 = 0; k1 < N1; ++k1) { col[k1 * N2] = zr[k1]; col[k1 * N2 + H] = zi[k1]; }
            } else {
                // real column: Hermitian in k1, half of the values are stored
                const int k2 = c > HI ? H : 0;
                T* col = Rg + l * S + k2;
                const T* tw = TWD + 2 * k2 * N1;
#pragma unroll
                for (int n1 = 0; n1 < N1; ++n1) {
                    const T xr = col[n1 * N2];
                    ar[n1] = xr * tw[2 * n1];
                    ai[n1] = -(xr * tw[2 * n1 + 1]);
                }
                dft_n1<G, -1>(ar, ai, cb, sb, zr, zi);
                const int q = k2 ? N1 - 1 : N1;                // conjugate partner of k1 is q - k1 (mod N1)
#pragma unroll
                for (int k1 = 0; k1 < N1; ++k1) {
                    const int pk = (q - k1) % N1;
                    if (k1 == pk) col[k1 * N2] = zr[k1];
                    else if (k1 < pk) { col[k1 * N2] = zr[k1]; col[pk * N2] = zi[k1]; }
                }
            }
        }
    };

    // the filter's real part at DFT index k: 1 - (mask(k) + mask(-k)) / 2, or its masked part -(mask(k) +
    // mask(-k)) / 2 in the correction form (the output then adds the source itself)
    auto fmul = [&](int k) -> T {
        const int kn = k == 0 ? 0 : N - k;
        const T m = (T)0.5 * (T)((k >= P.mlo && k <= P.mhi) + (kn >= P.mlo && kn <= P.mhi));
        return P.fcorr ? -m : (T)1 - m;
    };
    // the multiplier of a generic column: d/dx i kap, d2/dx2 -kap^2, -d/dx -i kap, the filter's real part
    auto gmul = [&](int kind, T kap, int k, T xr, T xi, T& yr, T& yi) {
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
    // gadd: add the filter's rank-4 cross term at the unpaired modes, (ke / 2) (c_1 + i c_0) for mode kp and
    // its conjugate at N - kp, from the line's coefficients c = DL[l][2p .. 2p + 1]
    auto inv_b = [&](int kind, bool acc, bool gadd) {
        T cb[N1], sb[N1];
#pragma unroll
        for (int m = 0; m < N1; ++m) { cb[m] = CB[m]; sb[m] = SB[m]; }
        for (int t = tid; t < NTB; t += LP_THREADS) {
            const int l = t % LB, c = t / LB;
            T ar[N1], ai[N1], zr[N1], zi[N1];
            if (c < HI) {
                const int k2 = 1 + c;
                const T* col = Rg + l * S + k2;
                T* dst = Bg + l * S + k2;
                const T* tw = TWD + 2 * k2 * N1;
                const T pk2 = p * (T)k2;
#pragma unroll
                for (int k1 = 0; k1 < N1; ++k1)
                    gmul(kind, pk2 + pc[k1], k2 + N2 * k1, col[k1 * N2], col[k1 * N2 + H], ar[k1], ai[k1]);
                if (FILT && gadd) {
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
                    const T cs = tw[2 * n1], sn = tw[2 * n1 + 1];  // times exp(+2 pi i n1 k2 / N)
                    T br = fmat(zr[n1], cs, -(zi[n1] * sn));
                    T bi = fmat(zi[n1], cs, zr[n1] * sn);
                    if (acc) { br += dst[n1 * N2]; bi += dst[n1 * N2 + H]; }
                    dst[n1 * N2] = br;
                    dst[n1 * N2 + H] = bi;
                }
            } else {
                const int k2 = c > HI ? H : 0;
                const T* col = Rg + l * S + k2;
                T* dst = Bg + l * S + k2;
                const T* tw = TWD + 2 * k2 * N1;
                const int q = k2 ? N1 - 1 : N1;
#pragma unroll
                for (int k1 = 0; k1 < N1; ++k1) {
                    const int pk = (q - k1) % N1;
                    T xr, xi;
                    if (k1 == pk) { xr = col[k1 * N2]; xi = (T)0; }
                    else if (k1 < pk) { xr = col[k1 * N2]; xi = col[pk * N2]; }
                    else { xr = col[pk * N2]; xi = -col[k1 * N2]; }
                    if (FILT && kind == 4) {
                        const T m = fmul(k2 + N2 * k1);
                        ar[k1] = m * xr; ai[k1] = m * xi;
                    } else {
                        spec_mul(kind, p, k2 + N2 * k1, N, xr, xi, ar[k1], ai[k1]);
                    }
                }
                dft_n1<G, 1>(ar, ai, cb, sb, zr, zi);
#pragma unroll
                for (int n1 = 0; n1 < N1; ++n1) {
                    T br = fmat(zr[n1], tw[2 * n1], -(zi[n1] * tw[2 * n1 + 1]));   // real part
                    if (acc) br += dst[n1 * N2];
                    dst[n1 * N2] = br;
                }
            }
        }
    };

    // ---- stage A (inverse), in place in the second region: E (slots 0 .. H) and O (H + 1 ..), scaled by 1 / N
    auto inv_a = [&]() {
        const T rn = (T)(1.0 / N), rn2 = (T)(2.0 / N);
        T* row = Bg + la * S + n1a * N2;
        T outv[H + 1];
        if (wact) {
            if (!odd) {
                T ca[H + 1];
#pragma unroll
                for (int m = 0; m <= H; ++m) ca[m] = CA[m];
                T c[H + 1];
                c[0] = row[0] * rn;
#pragma unroll
                for (int k = 1; k <= HI; ++k) c[k] = row[k] * rn2;
                if constexpr (G::EV2) c[H] = row[H] * rn;
                cos_tf<G>(c, ca, outv);
            } else {
                T sa[H + 1], g[HI + 1], o[HI + 1];
#pragma unroll
                for (int m = 0; m <= H; ++m) sa[m] = SA[m];
                g[0] = (T)0;
#pragma unroll
                for (int k = 1; k <= HI; ++k) g[k] = row[H + k] * rn2;
                sin_tf<G>(g, sa, o);
#pragma unroll
                for (int k = 1; k <= HI; ++k) outv[k] = o[k];
            }
        }
        __syncthreads();
        if (act) {
            if (!odd) {
#pragma unroll
                for (int k = 0; k <= H; ++k) row[k] = outv[k];
            } else {
#pragma unroll
                for (int k = 1; k <= HI; ++k) row[H + k] = outv[k];
            }
        }
        __syncthreads();
    };

    // ---- output: y[n1 + N1 n2] = E[n2] -+ O[n2] (+ addend), stored; returns whether a stored value is NaN.
    // (Loading the addend earlier, before an inverse stage, costs more in register spills than its latency.)
    auto store = [&](const LOut& O, int line0, int z, int x0) -> bool {
        const int xb = DIR == 0 ? xline_base_at(P, line0, z, mrow) : 0;
        // the line's constant: its first value times the row sum of the output's operator (the filter's
        // real part; one source per such output)
        int rs = 0;
        if (FILT)
            for (int e = 0; e < O.nterms; ++e) rs += O.term[e].rsum;
        const T cst = rs ? (T)rs * (T)pref : (T)0;
        const T* row = Bg + mrow * S;
        T add[PF];
        auto load_add = [&](auto tag) {
            using S_ = decltype(tag);
            const Rs rs = rs_make(O.addend, fbytes<S_>(P));
#pragma unroll
            for (int r = 0; r < PF; ++r) add[r] = (T)bldr<S_>(rs, elem(line0, z, x0, xb, jt + 16 * r));
        };
        if (!O.addend) {
#pragma unroll
            for (int r = 0; r < PF; ++r) add[r] = (T)0;
        } else if (O.add_dt == DT_F64) {
            load_add(double{});
        } else {
            load_add(float{});
        }
        bool bad = false;
        auto st = [&](auto tag) {
            using S_ = decltype(tag);
            const Rs rs = rs_make(O.out, fbytes<S_>(P));
            const int n1j = jt % N1, n2j = jt / N1;
#pragma unroll
            for (int r = 0; r < PF; ++r) {
                const int j = jt + 16 * r;
                // sample j = n1 + N1 n2 from jt's and the compile-time offset 16 r's digits
                int n1 = n1j + (16 * r) % N1, n2 = n2j + (16 * r) / N1;
                if (n1 >= N1) { n1 -= N1; ++n2; }
                if (j >= N) { n1 = 0; n2 = 0; }            // samples past the line: no LDS read out of the row
                const T* e = row + n1 * N2;
                const int m = n2 <= H ? n2 : N2 - n2;      // E[m] -+ O[m]: - for n2 <= HI, + past the half
                const T ov = (m >= 1 && m <= HI) ? e[H + m] : (T)0;
                const T y = n2 <= HI ? e[m] - ov : e[m] + ov;
                const T v = (y + cst) + add[r];
                const uint32_t ei = elem(line0, z, x0, xb, j);
                bstr<S_>(rs, ei, (S_)v);
                bad |= ei != Z_OOB && v != v;
            }
        };
        if (O.out_dt == DT_F64) st(double{});
        else st(float{});
        return bad;
    };

    // A unit's lines are loaded right before they are staged (most groups have one source and the tuned passes
    // take one tile per work-group, so a unit rarely has a successor whose lines could be fetched during it;
    // holding them would cost registers the transforms use).
    for (int unit = 0; unit < nunit; ++unit) {
        const int lt = tile0 + unit / Gr.ntiles, t = unit % Gr.ntiles;
        int line0, z, x0;
        geom(lt, line0, z, x0);
        fetch(unit);
        __syncthreads();                 // the previous unit is done with the line region
        deposit();
        __syncthreads();
        fwd_a();
        fwd_b();
        __syncthreads();
        for (int o = 0; o < Gr.nout; ++o) {
            const LOut& O = Gr.out[o];
            if (FILT && DIR == 0 && O.gmode == 1) {
                // filter x-pass: the unit's rank-4 dots (cos, sin sums of the unpaired modes: Re X[kp], -Im X[kp])
                // from the spectrum, then this 16-line block's partials of the cross-term matrix
                for (int idx = tid; idx < LB * 4; idx += LP_THREADS) {
                    const int l = idx >> 2, sl = idx & 3, q = sl >> 1;
                    T v = (T)0;
                    if (q < P.nkp) {
                        int k = P.kp[q], sg = 1;
                        int k2 = k % N2;
                        if (!(k2 >= 1 && k2 <= HI)) { k = N - k; k2 = k % N2; sg = -1; }   // stored conjugate
                        const T* e = Rg + l * S + (k / N2) * N2 + k2;
                        v = (sl & 1) ? (T)(-sg) * e[H] : e[0];
                    }
                    DL[idx] = v;
                }
                __syncthreads();
                fold_g_partials<T, 1>(P, O, DL, line0, nl);
                __syncthreads();
                continue;
            }
            const bool gadd = FILT && DIR == 1 && O.gmode == 2;
            if (gadd) fold_lr_from_g<T, 1>(P, O, DL, GL, x0, z);   // DL[l][s] = sum_r U_x[x0 + l][r] G[r][s]
            // the output is finished when its last source is this unit's
            bool last = false;
            for (int e = 0; e < O.nterms; ++e) last |= O.term[e].tile == t;
            for (int e = 0; e < O.nterms; ++e) last &= O.term[e].tile <= t;
            bool first = true;
            for (int e = 0; e < O.nterms; ++e) {
                const LTerm& tm = O.term[e];
                if (tm.tile < t) first = false;
                if (tm.tile != t || tm.lr) continue;          // a rank-4 term is the output's gmode
                inv_b(tm.sk, !first, gadd);
                __syncthreads();
                first = false;
            }
            if (!last) continue;
            inv_a();
            const bool bad = store(O, line0, z, x0);
            if (O.nan_check && __any(bad) && lane == 0) {
                const int step = __atomic_load_n(&P.flags[1], __ATOMIC_RELAXED);
                atomicCAS(&P.flags[0], 0, step > 0 ? step : 1);
            }
            __syncthreads();             // the second region is read before the next output writes it
        }
    }
}

template <typename T, int DIR, int N1, int N2, typename R, bool FILT>
__global__ __launch_bounds__(LP_THREADS) __attribute__((amdgpu_waves_per_eu(2))) void k_spec_pass(
    const LPass* __restrict__ Pp, const Members M, const int xcd_order) {
    spec_body<T, DIR, N1, N2, R, FILT>(Pp, M, xcd_order, (int)blockIdx.x, (int)gridDim.x, (int)blockIdx.y,
                                       (int)gridDim.y);
}

template <typename T, int DIR, int N1, int N2, typename R, bool FILT>
hipError_t spec_launch_one(const LPass* d, Members M, dim3 grid, size_t lds, hipStream_t s, int order) {
    const void* fn = reinterpret_cast<const void*>(&k_spec_pass<T, DIR, N1, N2, R, FILT>);
    if (lds > 64 * 1024) {
        static bool done = false;
        if (!done) {
            const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            if (e != hipSuccess) return e;
            done = true;
        }
    }
    hipLaunchKernelGGL((k_spec_pass<T, DIR, N1, N2, R, FILT>), grid, dim3(LP_THREADS), lds, s, d, M, order);
    return hipGetLastError();
}

template <typename T, int DIR, typename R, bool FILT>
hipError_t spec_launch_r(const LPass* d, Members M, dim3 grid, int fid, size_t lds, hipStream_t s, int order) {
    switch (fid) {
        case 0: return spec_launch_one<T, DIR, 7, 43, R, FILT>(d, M, grid, lds, s, order);
        case 1: return spec_launch_one<T, DIR, 4, 8, R, FILT>(d, M, grid, lds, s, order);
        case 2: return spec_launch_one<T, DIR, 8, 8, R, FILT>(d, M, grid, lds, s, order);
        case 3: return spec_launch_one<T, DIR, 8, 12, R, FILT>(d, M, grid, lds, s, order);
        case 4: return spec_launch_one<T, DIR, 8, 16, R, FILT>(d, M, grid, lds, s, order);
        case 5: return spec_launch_one<T, DIR, 8, 24, R, FILT>(d, M, grid, lds, s, order);
        case 6: return spec_launch_one<T, DIR, 8, 32, R, FILT>(d, M, grid, lds, s, order);
        default: return hipErrorInvalidValue;
    }
}

// Every factored-DFT instance of one (precision, direction, filter or RK-stage) family; explicitly
// instantiated by the hqt_spec_*.hip translation units.  fid: index into kSpecFact; wide: a double source
// in an FP32 pass (prefetch registers in double).
template <typename T, int DIR, bool FILT>
hipError_t spec_launch_td(const LPass* d, Members M, dim3 grid, int fid, size_t lds, hipStream_t s, int order,
                          int wide) {
    if constexpr (sizeof(T) == 8) {
        (void)wide;
        return spec_launch_r<T, DIR, double, FILT>(d, M, grid, fid, lds, s, order);
    } else {
        if (wide) return spec_launch_r<T, DIR, double, FILT>(d, M, grid, fid, lds, s, order);
        return spec_launch_r<T, DIR, float, FILT>(d, M, grid, fid, lds, s, order);
    }
}

#endif

}  // namespace hqt
