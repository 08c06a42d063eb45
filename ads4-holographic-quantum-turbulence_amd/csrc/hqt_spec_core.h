// Factored-DFT line transforms for gfx950 (MI355X): the geometry, trig-table layout and small building blocks
// shared by the single-region passes (hqt_spec1_core.h) of the RK stages, the constraint and the filter.
//
// Reference: tevol_driven_cluster.m:60-61 (func_Lx / func_Ly: the dense fourdif matrix times every
// line) and third_party/dmsuite/fourdif.m.  The fourdif matrix Lx = px Dx is the circulant whose
// eigenvalues are i px k (k the signed wavenumber, 0 at the Nyquist mode of even N), and Lx Lx (the
// reference's second derivative, func_Lx applied twice) has -(px k)^2.  So d_x x = IDFT(i px k DFT(x))
// exactly: a pass computes each source line's DFT, multiplies it by the eigenvalues of the output's
// operator and transforms back, instead of a dense (or even/odd folded) matrix product.
//
// The DFT of length N = N1 N2 is split (decimation in time): with n = n1 + N1 n2, k = k2 + N2 k1,
//     X[k2 + N2 k1] = sum_n1 W_N1^{n1 k1} W_N^{n1 k2} A[n1, k2],    A[n1, k2] = sum_n2 x[n1 + N1 n2] W_N2^{n2 k2}
// (W_M = exp(-2 pi i / M)).  Stage A is a real DFT of length N2 of each of the N1 sub-sequences
// (only k2 = 0 .. N2/2 kept: the rest is conjugate), computed in its even/odd folded form; stage B is
// a complex DFT of length N1 per (line, k2) after the twiddle W_N^{n1 k2}.  The inverse runs the two
// stages backwards and ends in a real inverse DFT of length N2 (the stage-B output is Hermitian in k2
// for a real result).  At N = 301 = 7 x 43 a transform costs about 26 multiply-adds per point against
// 151 for one folded matrix application.
//
// Both stages run on the VALU (on MI355X FP64 vector FMA has the same peak as FP64 MFMA, and the
// stages' small odd sizes would waste much of a 16 x 16 x 4 tile).
//
// The Fourier filter (ff.m:7-11, w = real(ifft2(b)) with the masked modes of b = fft2(x) zeroed) runs
// through the same transforms.  For a real field and the separable keep-mask a(kx) b(ky) (the complement of
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

#endif

}  // namespace hqt
