// Radial (Chebyshev z) kernels for gfx950: the per-column part of the AdS4 RK4 stage and
// the A_t constraint solve with the boundary reset.
//
// Every nz x nz radial operator (Lz = d_z, Lz1 = the integration of func_Lzr) is applied
// with FP64 MFMA: one wave owns 16 boundary-plane columns, a column vector is held in
// registers as "slots" (slot s, lane group g = lane/16 holds row 4s + g of column lane%16),
// which is both the B-operand layout of v_mfma_f64_16x16x4 for k-block s and the layout of
// its accumulator (gfx950: accumulator register r of row tile t holds row 16t + 4r + g).
// A matvec is therefore RT x KS MFMAs with the operator tiles kept in registers, and
// matvecs chain (Lz(Lz Psi), Lzr(E)) without leaving registers.  The point-wise parts of
// the right-hand sides run on the same slots, with 16 consecutive columns per lane group,
// so every global access is a 128-byte coalesced segment.
//
// All radial arithmetic is FP64 in both precisions; TS is the storage type of the x/y
// fields (double: plain form, float: mixed form).
//
// Reference (scripts/simulation/tevol_driven_cluster.m):
//   F1, F3, F4 point-wise parts + func_Lz / func_Lzr     :62-65, :72-75
//   F2 (only the z=0 row is consumed, k2(:,:,nz) at :128) :76
//   RK4 stage states / combine / update                  :120-128
//   A_t constraint solve, boundary reset, O              :129-136
#pragma once
#include "hqt_kernels.h"
#include <math.h>

namespace hqt {

typedef double v4d __attribute__((ext_vector_type(4)));

template <typename TS> __device__ __forceinline__ double ld(const void* p, int64_t i) {
    return (double)static_cast<const TS*>(p)[i];
}
template <typename TS> __device__ __forceinline__ void st(void* p, int64_t i, double v) {
    static_cast<TS*>(p)[i] = (TS)v;
}

__device__ __forceinline__ bool stopped(const int32_t* flags) {
    return __builtin_nontemporal_load(&flags[0]) != 0;
}

template <int NZ> struct ZG {
    static constexpr int KS = (NZ + 3) / 4;     // slots (k-blocks of 4 rows)
    static constexpr int RT = (NZ + 15) / 16;   // 16-row output tiles
};

// Operator tiles: A[t][s] = M[16t + m][4s + kk] (m = lane%16, kk = lane/16), zero outside n x n.
template <int NZ>
__device__ __forceinline__ void load_op(const double* __restrict__ M, int ldm, int n, double (&A)[ZG<NZ>::RT][ZG<NZ>::KS],
                                        int m, int kk) {
#pragma unroll
    for (int t = 0; t < ZG<NZ>::RT; ++t)
#pragma unroll
        for (int s = 0; s < ZG<NZ>::KS; ++s) {
            const int i = 16 * t + m, j = 4 * s + kk;
            A[t][s] = (i < n && j < n) ? M[i * ldm + j] : 0.0;
        }
}

// out = A * in on the slot layout (rows >= n of the operator are zero, so padded slots stay 0).
template <int NZ>
__device__ __forceinline__ void mv(const double (&A)[ZG<NZ>::RT][ZG<NZ>::KS], const double (&in)[ZG<NZ>::KS],
                                   double (&out)[ZG<NZ>::KS]) {
#pragma unroll
    for (int t = 0; t < ZG<NZ>::RT; ++t) {
        v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < ZG<NZ>::KS; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(A[t][s], in[s], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (4 * t + r < ZG<NZ>::KS) out[4 * t + r] = acc[r];
    }
}

// Load a [nxy, nz] field into slots (0 outside the grid).
template <int NZ, typename TS>
__device__ __forceinline__ void load_slots(const void* f, int64_t c, bool cok, int g, int64_t n, double (&v)[ZG<NZ>::KS]) {
#pragma unroll
    for (int s = 0; s < ZG<NZ>::KS; ++s) {
        const int k = 4 * s + g;
        v[s] = (cok && k < NZ) ? ld<TS>(f, c + n * k) : 0.0;
    }
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

// ---- one RK stage for 16 columns per wave (F1, F3, F4, F2 of :72-76 + RK bookkeeping) ----
//   E1 = 1i*At*dPsi + 0.5i*Psi*dAt + 0.5*f*d2Psi - 1.5*z^2*dPsi - 0.5*z*Psi
//        - 0.5*(Ax^2 + Ay^2)*Psi + 0.5*(dxx+dyy)Psi - 0.5i*Psi*(dxAx+dyAy) - 1i*(Ax*dxPsi + Ay*dyPsi)
//   E3 = 0.5*Lz(dxAt + f*Lz Ax) + 0.5*curl3 + Im(conj(Psi)*dxPsi) - Ax*|Psi|^2
//   E4 = 0.5*Lz(dyAt + f*Lz Ay) + 0.5*curl4 + Im(conj(Psi)*dyPsi) - Ay*|Psi|^2
//   F1 = Lzr(E1), F3 = Lzr(E3), F4 = Lzr(E4), F2 = (1-z^3)*Lz(dxAx + dyAy)
// (the reference's current term -0.5i*(conj(Psi)*dPsi - Psi*d conj(Psi)) equals Im(conj(Psi)*dPsi)).
// Every input field is read once; the stage state Psi and the x/y derivatives of Psi stay in
// registers across the three equations.
template <int NZ, typename TS, int DIR>
__device__ __forceinline__ void zstage_a(const ZStage& A, int64_t c, bool cok, int g, const double (&Lz)[ZG<NZ>::RT][ZG<NZ>::KS],
                                         const double (&L1)[ZG<NZ>::RT][ZG<NZ>::KS], const double (&sR)[ZG<NZ>::KS],
                                         const double (&sI)[ZG<NZ>::KS], const double (&dR)[ZG<NZ>::KS],
                                         const double (&dI)[ZG<NZ>::KS]) {
    constexpr int KS = ZG<NZ>::KS;
    const int64_t n = A.nxy;
    const void* sA = DIR == 0 ? A.sAx : A.sAy;
    const void* dAt = DIR == 0 ? A.dxAt : A.dyAt;
    const void* curl = DIR == 0 ? A.curl3 : A.curl4;
    double a[KS], u[KS], e[KS];
    load_slots<NZ, TS>(sA, c, cok, g, n, a);
    mv<NZ>(Lz, a, u);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + g;
        u[s] = (cok && k < NZ) ? ld<TS>(dAt, c + n * k) + A.zo.f1[k] * u[s] : 0.0;
    }
    mv<NZ>(Lz, u, e);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + g;
        double ev = 0.0;
        if (cok && k < NZ) {
            const int64_t p = c + n * k;
            const double sr = sR[s], si = sI[s], av = a[s];
            const double cur = sr * dI[s] - si * dR[s];
            ev = 0.5 * e[s] + 0.5 * ld<TS>(curl, p) + cur - ((av * sr) * sr + (av * si) * si);
        }
        u[s] = ev;
    }
    mv<NZ>(L1, u, e);
    void* acc = DIR == 0 ? A.aAx : A.aAy;
    void* zf = DIR == 0 ? A.zAx : A.zAy;
    void* nx = DIR == 0 ? A.nAx : A.nAy;
    double* Fo = DIR == 0 ? A.F3 : A.F4;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + g;
        if (!cok || k >= NZ) continue;
        const int64_t p = c + n * k;
        if (A.mode == 1) { Fo[p] = e[s]; continue; }
        rk_update<TS>(A.stage, A.h * e[s], acc, zf, nx, p);
    }
}

// grid: x = ceil(nxy/64) blocks of 4 waves, 16 columns per wave.  WPE: the occupancy the
// register allocation targets (2 waves per SIMD fit every nz without spills; 3 hide more memory
// latency at the cost of some spills from nz = 18 up); the host picks per (nz, precision) by timing.
template <int NZ, typename TS, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_zstage(const ZStage* __restrict__ Ap,
                                                                                            const Members M) {
    const MemberBlock mb = member_block(M);
    const ZStage& A = Ap[mb.m];
    if (stopped(A.flags)) return;
    constexpr int KS = ZG<NZ>::KS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int64_t c = (int64_t)mb.bx * 64 + wave * 16 + m;
    const bool cok = c < A.nxy;
    const int64_t n = A.nxy;
    double Lz[ZG<NZ>::RT][KS], L1[ZG<NZ>::RT][KS];
    load_op<NZ>(A.zo.Lz, NZ, NZ, Lz, m, g);
    load_op<NZ>(A.zo.Lz1, NZ - 1, NZ - 1, L1, m, g);
    double sR[KS], sI[KS], xR[KS], xI[KS], yR[KS], yI[KS], dv[KS];
    load_slots<NZ, TS>(A.sR, c, cok, g, n, sR);
    load_slots<NZ, TS>(A.sI, c, cok, g, n, sI);
    load_slots<NZ, TS>(A.dxR, c, cok, g, n, xR);
    load_slots<NZ, TS>(A.dxI, c, cok, g, n, xI);
    load_slots<NZ, TS>(A.dyR, c, cok, g, n, yR);
    load_slots<NZ, TS>(A.dyI, c, cok, g, n, yI);
    load_slots<NZ, TS>(A.divA, c, cok, g, n, dv);
    // ---- F1 ----
    {
        double dR[KS], dI[KS], eR[KS], eI[KS];
        mv<NZ>(Lz, sR, dR);
        mv<NZ>(Lz, sI, dI);
        {
            double d2R[KS], d2I[KS];
            mv<NZ>(Lz, dR, d2R);
            mv<NZ>(Lz, dI, d2I);
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const int k = 4 * s + g;
                double er = 0.0, ei = 0.0;
                if (cok && k < NZ) {
                    const int64_t p = c + n * k;
                    const double sr = sR[s], si = sI[s];
                    const double at = A.At[p], dat = A.dzAt[p];
                    const double ax = ld<TS>(A.sAx, p), ay = ld<TS>(A.sAy, p);
                    const double lpr = ld<TS>(A.lapR, p), lpi = ld<TS>(A.lapI, p);
                    const double hf = A.zo.hf[k], cd = A.zo.cd[k], cp = A.zo.cp[k];
                    er = -at * dI[s];            ei = at * dR[s];
                    er += (-0.5 * si) * dat;     ei += (0.5 * sr) * dat;
                    er += hf * d2R[s];           ei += hf * d2I[s];
                    er += cd * dR[s];            ei += cd * dI[s];
                    er += cp * sr;               ei += cp * si;
                    const double a2 = (-0.5 * ax) * ax, b2 = (-0.5 * ay) * ay;
                    er += a2 * sr;               ei += a2 * si;
                    er += b2 * sr;               ei += b2 * si;
                    er += 0.5 * lpr;             ei += 0.5 * lpi;
                    er += (0.5 * si) * dv[s];    ei += (-0.5 * sr) * dv[s];
                    const double qr = ax * xR[s] + ay * yR[s], qi = ax * xI[s] + ay * yI[s];
                    er += qi;                    ei -= qr;
                }
                eR[s] = er;
                eI[s] = ei;
            }
        }
        mv<NZ>(L1, eR, dR);     // F1 = Lzr(E1); the last row of Lzr is zero
        mv<NZ>(L1, eI, dI);
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = 4 * s + g;
            if (!cok || k >= NZ) continue;
            const int64_t p = c + n * k;
            if (A.mode == 1) { A.F1R[p] = dR[s]; A.F1I[p] = dI[s]; continue; }
            rk_update<TS>(A.stage, A.h * dR[s], A.aR, A.zR, A.nR, p);
            rk_update<TS>(A.stage, A.h * dI[s], A.aI, A.zI, A.nI, p);
        }
    }
    // ---- F3, F4 ----
    zstage_a<NZ, TS, 0>(A, c, cok, g, Lz, L1, sR, sI, xR, xI);
    zstage_a<NZ, TS, 1>(A, c, cok, g, Lz, L1, sR, sI, yR, yI);
    // ---- F2 (z=0 row feeds the charge update at :128) ----
    {
        double w[KS];
        mv<NZ>(Lz, dv, w);
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = 4 * s + g;
            if (!cok || k >= NZ) continue;
            if (A.mode == 1) { A.F2[c + n * k] = A.zo.f2[k] * w[s]; continue; }
            if (k != NZ - 1) continue;
            const double kv = A.h * (A.zo.f2[k] * w[s]);
            double* ar = A.arho;
            if (A.stage == 1) ar[c] = kv;
            else if (A.stage < 4) ar[c] = ar[c] + 2.0 * kv;
            else {
                const double fin = (1.0 / 6.0) * (ar[c] + kv);
                for (int q = 0; q < NZ; ++q) A.rho[c + n * q] = A.rho[c + n * q] - fin;
            }
        }
    }
}

// ---- post-step: A_t constraint solve and boundary reset (:129-133), O for the winding (:136)
template <int NZ, typename TS>
__global__ __launch_bounds__(256) void k_zpost(const ZPost* __restrict__ Ap, const Members M) {
    const MemberBlock mb = member_block(M);
    const ZPost& A = Ap[mb.m];
    if (stopped(A.flags)) return;
    constexpr int KS = ZG<NZ>::KS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int64_t c = (int64_t)mb.bx * 64 + wave * 16 + m;
    const bool cok = c < A.nxy;
    const int64_t n = A.nxy;
    double Lz[ZG<NZ>::RT][KS], L1[ZG<NZ>::RT][KS];
    load_op<NZ>(A.zo.Lz, NZ, NZ, Lz, m, g);
    load_op<NZ>(A.zo.Lz1, NZ - 1, NZ - 1, L1, m, g);
    double R[KS], I[KS], a[KS], b[KS], w[KS];
    load_slots<NZ, TS>(A.R, c, cok, g, n, R);
    load_slots<NZ, TS>(A.I, c, cok, g, n, I);
    load_slots<NZ, TS>(A.divA, c, cok, g, n, a);
    mv<NZ>(Lz, a, w);                // Lz(dxAx + dyAy)
    mv<NZ>(Lz, R, a);                // Lz Re Psi
    mv<NZ>(Lz, I, b);                // Lz Im Psi
    // cons = real(Lz(divA) + 1i*(conj(Psi)*Lz(Psi) - Psi*Lz(conj(Psi)))) = Lz(divA) - 2*Im(conj(Psi)*Lz Psi)
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const double pim = R[s] * b[s] - I[s] * a[s];
        w[s] = w[s] - 2.0 * pim;
    }
    mv<NZ>(L1, w, a);                // Lzr(cons)
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + g;
        w[s] = (cok && k < NZ) ? a[s] - A.rho[c + n * k] : 0.0;
    }
    mv<NZ>(L1, w, a);                // Lzr(Lzr(cons) - rhoi)
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + g;
        if (!cok || k >= NZ) continue;
        A.At[c + n * k] = (k < NZ - 1) ? a[s] + A.mu : A.mu;
        if (k == NZ - 1) {
            // boundary reset: Z1(:,:,nz)=0; Z3(:,:,nz)=ax_b; Z4(:,:,nz)=ay_b  (Z2 row set above)
            const int64_t pb = c + n * k;
            st<TS>(A.R, pb, 0.0); st<TS>(A.I, pb, 0.0);
            st<TS>(A.Ax, pb, A.axb[c]); st<TS>(A.Ay, pb, A.ayb[c]);
            R[s] = 0.0;
            I[s] = 0.0;
        }
    }
    // O1 = func_Lz(Z1) at z=0 after the reset
    mv<NZ>(Lz, R, a);
    mv<NZ>(Lz, I, b);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + g;
        if (cok && k == NZ - 1) { A.O[2 * c] = a[s]; A.O[2 * c + 1] = b[s]; }
    }
}

// ---- d_z A_t of the frozen A_t; also advances the device step counter ----------------------
template <int NZ>
__global__ __launch_bounds__(256) void k_zprep(const ZPrep* __restrict__ Ap, const Members M) {
    const MemberBlock mb = member_block(M);
    const ZPrep& A = Ap[mb.m];
    if (stopped(A.flags)) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int64_t c = (int64_t)mb.bx * 64 + wave * 16 + m;
    if (A.bump && mb.bx == 0 && threadIdx.x == 0) A.flags[1] = A.flags[1] + 1;
    const bool cok = c < A.nxy;
    double Lz[ZG<NZ>::RT][ZG<NZ>::KS], a[ZG<NZ>::KS], d[ZG<NZ>::KS];
    load_op<NZ>(A.Lz, NZ, NZ, Lz, m, g);
    load_slots<NZ, double>(A.At, c, cok, g, A.nxy, a);
    mv<NZ>(Lz, a, d);
#pragma unroll
    for (int s = 0; s < ZG<NZ>::KS; ++s) {
        const int k = 4 * s + g;
        if (cok && k < NZ) A.dzAt[c + A.nxy * k] = d[s];
    }
}

// Per-nz launchers; each hqt_radial_nz_*.hip translation unit instantiates a range of nz.
template <int NZ>
hipError_t zstage_launch_nz(const ZStage* d, Members M, int blocks, int fp32, int wpe, hipStream_t s) {
    if (wpe == 3) {
        if (fp32) hipLaunchKernelGGL((k_zstage<NZ, float, 3>), dim3(blocks), dim3(256), 0, s, d, M);
        else hipLaunchKernelGGL((k_zstage<NZ, double, 3>), dim3(blocks), dim3(256), 0, s, d, M);
    } else {
        if (fp32) hipLaunchKernelGGL((k_zstage<NZ, float, 2>), dim3(blocks), dim3(256), 0, s, d, M);
        else hipLaunchKernelGGL((k_zstage<NZ, double, 2>), dim3(blocks), dim3(256), 0, s, d, M);
    }
    return hipGetLastError();
}
template <int NZ>
hipError_t zpost_launch_nz(const ZPost* d, Members M, int blocks, int fp32, hipStream_t s) {
    if (fp32) hipLaunchKernelGGL((k_zpost<NZ, float>), dim3(blocks), dim3(256), 0, s, d, M);
    else hipLaunchKernelGGL((k_zpost<NZ, double>), dim3(blocks), dim3(256), 0, s, d, M);
    return hipGetLastError();
}
template <int NZ>
hipError_t zprep_launch_nz(const ZPrep* d, Members M, int blocks, hipStream_t s) {
    hipLaunchKernelGGL((k_zprep<NZ>), dim3(blocks), dim3(256), 0, s, d, M);
    return hipGetLastError();
}

}  // namespace hqt
