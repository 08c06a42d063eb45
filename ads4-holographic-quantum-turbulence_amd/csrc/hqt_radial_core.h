// Radial (Chebyshev z) kernels for gfx950: the per-column part of the AdS4 RK4 stage,
// the A_t constraint solve with the boundary reset, d_z A_t, and the winding removal.
//
// One thread owns one boundary-plane column (x, y) and holds it in registers; the
// nz x nz operators (Lz, Lz1) are wave-uniform, so every matrix entry is a scalar
// operand (s_load) of a v_fma_f64.  All radial arithmetic is FP64 in both precisions;
// the storage type TS of the x/y fields is double (plain form) or float (mixed form).
//
// Reference (scripts/simulation/tevol_driven_cluster.m):
//   F1, F3, F4 point-wise parts + func_Lz / func_Lzr     :62-65, :72-75
//   F2 (only the z=0 row is consumed, k2(:,:,nz) at :128) :76
//   RK4 stage states / combine / update                  :120-128
//   A_t constraint solve, boundary reset                 :129-133
//   winding removal                                      :136-139
#pragma once
#include "hqt_kernels.h"
#include <math.h>

namespace hqt {

template <typename TS> __device__ __forceinline__ double ld(const void* p, int64_t i) {
    return (double)static_cast<const TS*>(p)[i];
}
template <typename TS> __device__ __forceinline__ void st(void* p, int64_t i, double v) {
    static_cast<TS*>(p)[i] = (TS)v;
}

__device__ __forceinline__ bool stopped(const int32_t* flags) {
    return __builtin_nontemporal_load(&flags[0]) != 0;
}

// RK4 bookkeeping for one value of one evolved field (:120-128).
//   stage 1: acc = k;        next = z + k/2
//   stage 2: acc = acc + 2k; next = z + k/2
//   stage 3: acc = acc + 2k; next = z + k
//   stage 4: z = z + (1/6)*(acc + k)
template <typename TS>
__device__ __forceinline__ void rk_update(int stage, double kv, void* acc, void* z, void* nxt, int64_t i) {
    if (stage == 1) {
        st<TS>(acc, i, kv);
        st<TS>(nxt, i, ld<TS>(z, i) + kv * 0.5);
    } else if (stage == 2) {
        st<TS>(acc, i, ld<TS>(acc, i) + 2.0 * kv);
        st<TS>(nxt, i, ld<TS>(z, i) + kv * 0.5);
    } else if (stage == 3) {
        st<TS>(acc, i, ld<TS>(acc, i) + 2.0 * kv);
        st<TS>(nxt, i, ld<TS>(z, i) + kv);
    } else {
        const double fin = (1.0 / 6.0) * (ld<TS>(acc, i) + kv);
        st<TS>(z, i, ld<TS>(z, i) + fin);
    }
}

// ---- Psi equation (F1, :72-73) ------------------------------------------------------------
template <int NZ, typename TS>
__device__ void zstage_psi(const ZStage& A, int64_t c) {
    const int64_t n = A.nxy;
    const double* __restrict__ Lz = A.zo.Lz;
    double dR[NZ], dI[NZ];
    {
        double sR[NZ], sI[NZ];
#pragma unroll
        for (int k = 0; k < NZ; ++k) { sR[k] = ld<TS>(A.sR, c + n * k); sI[k] = ld<TS>(A.sI, c + n * k); }
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            double ar = 0.0, ai = 0.0;
#pragma unroll
            for (int m = 0; m < NZ; ++m) { ar = fma(Lz[k * NZ + m], sR[m], ar); ai = fma(Lz[k * NZ + m], sI[m], ai); }
            dR[k] = ar; dI[k] = ai;
        }
    }
    double eR[NZ], eI[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        double d2r = 0.0, d2i = 0.0;
#pragma unroll
        for (int m = 0; m < NZ; ++m) { d2r = fma(Lz[k * NZ + m], dR[m], d2r); d2i = fma(Lz[k * NZ + m], dI[m], d2i); }
        const int64_t p = c + n * k;
        const double sr = ld<TS>(A.sR, p), si = ld<TS>(A.sI, p);
        const double at = A.At[p], dat = A.dzAt[p];
        const double ax = ld<TS>(A.sAx, p), ay = ld<TS>(A.sAy, p);
        const double dxr = ld<TS>(A.dxR, p), dxi = ld<TS>(A.dxI, p);
        const double dyr = ld<TS>(A.dyR, p), dyi = ld<TS>(A.dyI, p);
        const double lpr = ld<TS>(A.lapR, p), lpi = ld<TS>(A.lapI, p);
        const double dv = ld<TS>(A.divA, p);
        // 1i*At*dPsi + 0.5i*Psi*dAt + 0.5*f*d2Psi - 1.5*z^2*dPsi - 0.5*z*Psi
        double er = -at * dI[k], ei = at * dR[k];
        er += (-0.5 * si) * dat; ei += (0.5 * sr) * dat;
        er += A.zo.hf[k] * d2r;  ei += A.zo.hf[k] * d2i;
        er += A.zo.cd[k] * dR[k]; ei += A.zo.cd[k] * dI[k];
        er += A.zo.cp[k] * sr;   ei += A.zo.cp[k] * si;
        // -0.5*Ax*Ax*Psi - 0.5*Ay*Ay*Psi
        const double a2 = (-0.5 * ax) * ax, b2 = (-0.5 * ay) * ay;
        er += a2 * sr; ei += a2 * si;
        er += b2 * sr; ei += b2 * si;
        // +0.5*(dxx+dyy)Psi - 0.5i*Psi*(dxAx+dyAy) - 1i*(Ax*dxPsi + Ay*dyPsi)
        er += 0.5 * lpr; ei += 0.5 * lpi;
        er += (0.5 * si) * dv; ei += (-0.5 * sr) * dv;
        const double qr = ax * dxr + ay * dyr, qi = ax * dxi + ay * dyi;
        er += qi; ei -= qr;
        eR[k] = er; eI[k] = ei;
    }
    const double* __restrict__ Lz1 = A.zo.Lz1;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        double fr = 0.0, fi = 0.0;
        if (k < NZ - 1) {
#pragma unroll
            for (int m = 0; m < NZ - 1; ++m) { fr = fma(Lz1[k * (NZ - 1) + m], eR[m], fr); fi = fma(Lz1[k * (NZ - 1) + m], eI[m], fi); }
        }
        const int64_t p = c + n * k;
        if (A.mode == 1) { A.F1R[p] = fr; A.F1I[p] = fi; continue; }
        rk_update<TS>(A.stage, A.h * fr, A.aR, A.zR, A.nR, p);
        rk_update<TS>(A.stage, A.h * fi, A.aI, A.zI, A.nI, p);
    }
}

// ---- A_x / A_y equations (F3, F4, :74-75) -------------------------------------------------
// dir 0: F3 = Re Lzr[0.5*Lz(dxAt + f*Lz Ax) + 0.5*curl3 + Im(conj(Psi)*dxPsi) - Ax*|Psi|^2]
// dir 1: F4 = Re Lzr[0.5*Lz(dyAt + f*Lz Ay) + 0.5*curl4 + Im(conj(Psi)*dyPsi) - Ay*|Psi|^2]
template <int NZ, typename TS, int DIR>
__device__ void zstage_a(const ZStage& A, int64_t c) {
    const int64_t n = A.nxy;
    const double* __restrict__ Lz = A.zo.Lz;
    const void* sA = DIR == 0 ? A.sAx : A.sAy;
    const void* dAt = DIR == 0 ? A.dxAt : A.dyAt;
    const void* curl = DIR == 0 ? A.curl3 : A.curl4;
    const void* dPR = DIR == 0 ? A.dxR : A.dyR;
    const void* dPI = DIR == 0 ? A.dxI : A.dyI;
    double u[NZ];
    {
        double s[NZ];
#pragma unroll
        for (int k = 0; k < NZ; ++k) s[k] = ld<TS>(sA, c + n * k);
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            double a = 0.0;
#pragma unroll
            for (int m = 0; m < NZ; ++m) a = fma(Lz[k * NZ + m], s[m], a);
            u[k] = ld<TS>(dAt, c + n * k) + A.zo.f1[k] * a;
        }
    }
    double e[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        double v = 0.0;
#pragma unroll
        for (int m = 0; m < NZ; ++m) v = fma(Lz[k * NZ + m], u[m], v);
        const int64_t p = c + n * k;
        const double sr = ld<TS>(A.sR, p), si = ld<TS>(A.sI, p);
        const double av = ld<TS>(sA, p);
        const double cur = sr * ld<TS>(dPI, p) - si * ld<TS>(dPR, p);
        e[k] = 0.5 * v + 0.5 * ld<TS>(curl, p) + cur - ((av * sr) * sr + (av * si) * si);
    }
    const double* __restrict__ Lz1 = A.zo.Lz1;
    void* acc = DIR == 0 ? A.aAx : A.aAy;
    void* zf = DIR == 0 ? A.zAx : A.zAy;
    void* nx = DIR == 0 ? A.nAx : A.nAy;
    double* Fo = DIR == 0 ? A.F3 : A.F4;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        double f = 0.0;
        if (k < NZ - 1) {
#pragma unroll
            for (int m = 0; m < NZ - 1; ++m) f = fma(Lz1[k * (NZ - 1) + m], e[m], f);
        }
        const int64_t p = c + n * k;
        if (A.mode == 1) { Fo[p] = f; continue; }
        rk_update<TS>(A.stage, A.h * f, acc, zf, nx, p);
    }
}

// ---- F2 = (1-z^3) * Lz(dxAx + dyAy) (:76); the RK step consumes only the z=0 row (:128) ----
template <int NZ, typename TS>
__device__ void zstage_f2(const ZStage& A, int64_t c) {
    const int64_t n = A.nxy;
    const double* __restrict__ Lz = A.zo.Lz;
    double dv[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) dv[k] = ld<TS>(A.divA, c + n * k);
    if (A.mode == 1) {
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            double a = 0.0;
#pragma unroll
            for (int m = 0; m < NZ; ++m) a = fma(Lz[k * NZ + m], dv[m], a);
            A.F2[c + n * k] = A.zo.f2[k] * a;
        }
        return;
    }
    double b = 0.0;
#pragma unroll
    for (int m = 0; m < NZ; ++m) b = fma(Lz[(NZ - 1) * NZ + m], dv[m], b);
    const double kv = A.h * (A.zo.f2[NZ - 1] * b);
    double* ar = A.arho;
    if (A.stage == 1) ar[c] = kv;
    else if (A.stage < 4) ar[c] = ar[c] + 2.0 * kv;
    else {
        const double fin = (1.0 / 6.0) * (ar[c] + kv);
#pragma unroll
        for (int k = 0; k < NZ; ++k) A.rho[c + n * k] = A.rho[c + n * k] - fin;
    }
}

template <int NZ, typename TS>
__global__ __launch_bounds__(256) void k_zstage(const ZStage A) {
    if (stopped(A.flags)) return;
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= A.nxy) return;
    switch (blockIdx.y) {
        case 0: zstage_psi<NZ, TS>(A, c); break;
        case 1: zstage_a<NZ, TS, 0>(A, c); break;
        case 2: zstage_a<NZ, TS, 1>(A, c); break;
        default: zstage_f2<NZ, TS>(A, c); break;
    }
}

// ---- post-step: A_t constraint solve and boundary reset (:129-133), O for the winding (:136)
template <int NZ, typename TS>
__global__ __launch_bounds__(256) void k_zpost(const ZPost A) {
    if (stopped(A.flags)) return;
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= A.nxy) return;
    const int64_t n = A.nxy;
    const double* __restrict__ Lz = A.zo.Lz;
    const double* __restrict__ Lz1 = A.zo.Lz1;
    double R[NZ], I[NZ], w[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) { R[k] = ld<TS>(A.R, c + n * k); I[k] = ld<TS>(A.I, c + n * k); }
    {
        double dv[NZ];
#pragma unroll
        for (int k = 0; k < NZ; ++k) dv[k] = ld<TS>(A.divA, c + n * k);
        // cons = real(Lz(divA) + 1i*(conj(Psi)*Lz(Psi) - Psi*Lz(conj(Psi)))) = Lz(divA) - 2*Im(conj(Psi)*Lz Psi)
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            double a = 0.0, dr = 0.0, di = 0.0;
#pragma unroll
            for (int m = 0; m < NZ; ++m) {
                a = fma(Lz[k * NZ + m], dv[m], a);
                dr = fma(Lz[k * NZ + m], R[m], dr);
                di = fma(Lz[k * NZ + m], I[m], di);
            }
            const double pim = R[k] * di - I[k] * dr;
            w[k] = a - 2.0 * pim;
        }
    }
    // Z2 = Lzr(Lzr(cons) - rhoi) + mu
    double w2[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        double a = 0.0;
        if (k < NZ - 1) {
#pragma unroll
            for (int m = 0; m < NZ - 1; ++m) a = fma(Lz1[k * (NZ - 1) + m], w[m], a);
        }
        w2[k] = a - A.rho[c + n * k];
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        double a = 0.0;
        if (k < NZ - 1) {
#pragma unroll
            for (int m = 0; m < NZ - 1; ++m) a = fma(Lz1[k * (NZ - 1) + m], w2[m], a);
        }
        A.At[c + n * k] = (k < NZ - 1) ? a + A.mu : A.mu;
    }
    // boundary reset: Z1(:,:,nz)=0; Z3(:,:,nz)=ax_b; Z4(:,:,nz)=ay_b  (Z2 row set above)
    const int64_t pb = c + n * (NZ - 1);
    st<TS>(A.R, pb, 0.0); st<TS>(A.I, pb, 0.0);
    st<TS>(A.Ax, pb, A.axb[c]); st<TS>(A.Ay, pb, A.ayb[c]);
    // O1 = func_Lz(Z1) at z=0, after the reset
    double orr = 0.0, oii = 0.0;
#pragma unroll
    for (int m = 0; m < NZ - 1; ++m) { orr = fma(Lz[(NZ - 1) * NZ + m], R[m], orr); oii = fma(Lz[(NZ - 1) * NZ + m], I[m], oii); }
    A.O[2 * c] = orr; A.O[2 * c + 1] = oii;
}

// Per-nz launchers; each hqt_radial_nz_*.hip translation unit instantiates a range of nz.
template <int NZ>
hipError_t zstage_launch_nz(const ZStage& A, int fp32, hipStream_t s) {
    dim3 grid((unsigned)((A.nxy + 255) / 256), 4);
    if (fp32) hipLaunchKernelGGL((k_zstage<NZ, float>), grid, dim3(256), 0, s, A);
    else hipLaunchKernelGGL((k_zstage<NZ, double>), grid, dim3(256), 0, s, A);
    return hipGetLastError();
}
template <int NZ>
hipError_t zpost_launch_nz(const ZPost& A, int fp32, hipStream_t s) {
    dim3 grid((unsigned)((A.nxy + 255) / 256));
    if (fp32) hipLaunchKernelGGL((k_zpost<NZ, float>), grid, dim3(256), 0, s, A);
    else hipLaunchKernelGGL((k_zpost<NZ, double>), grid, dim3(256), 0, s, A);
    return hipGetLastError();
}

}  // namespace hqt
