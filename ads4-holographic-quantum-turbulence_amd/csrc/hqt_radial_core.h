// Radial (Chebyshev z) kernels for gfx950: the per-column part of the AdS4 RK4 stage and
// the A_t constraint solve with the boundary reset.
//
// Every nz x nz radial operator (Lz = d_z, Lz1 = the integration of func_Lzr) is applied
// with FP64 MFMA: one wave owns 16 boundary-plane columns, a column vector is held in
// registers as "slots" (slot s, lane group g = lane/16 holds row 4s + g of column lane%16),
// which is both the B-operand layout of v_mfma_f64_16x16x4 for k-block s and the layout of
// its accumulator (gfx950: accumulator register r of row tile t holds row 16t + 4r + g).
// A matvec is therefore RT x KS MFMAs, and matvecs chain (Lz(Lz Psi), Lzr(E)) without
// leaving registers.  The stage kernel reads its operator tiles from LDS; the small post-step
// kernels keep them in registers.  The point-wise parts of
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
#include "hqt_buffer.h"
#include "hqt_kernels.h"
#include <math.h>

namespace hqt {

typedef double v4d __attribute__((ext_vector_type(4)));

// Field and operator accesses through global-address-space pointers: through generic pointers the compiler
// issues flat loads, which count against the scalar-load wait counter too, so every wait for a kernel argument
// also waited for the field loads in flight.
template <typename TS> using gptr = __attribute__((address_space(1))) TS*;
template <typename TS> __device__ __forceinline__ double ld(const void* p, int64_t i) {
    return (double)((const gptr<const TS>)p)[i];
}
template <typename TS> __device__ __forceinline__ void st(void* p, int64_t i, double v) {
    ((gptr<TS>)p)[i] = (TS)v;
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
            A[t][s] = (i < n && j < n) ? ((gptr<const double>)M)[i * ldm + j] : 0.0;
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

// Operator tiles in LDS, one entry per lane: T[(t*KS + s)*64 + lane] = A[t][s] of that lane.
template <int NZ>
__device__ __forceinline__ void stage_op(const double* __restrict__ M, int ldm, int n, double* T) {
    constexpr int KS = ZG<NZ>::KS, RT = ZG<NZ>::RT;
    for (int idx = threadIdx.x; idx < RT * KS * 64; idx += blockDim.x) {
        const int lane = idx & 63, ts = idx >> 6, t = ts / KS, s = ts % KS;
        const int i = 16 * t + (lane & 15), j = 4 * s + (lane >> 4);
        T[idx] = (i < n && j < n) ? M[i * ldm + j] : 0.0;
    }
}
// out = A * in with the operator tiles read from LDS at every use (volatile: no reuse across
// calls, so the tiles occupy no registers between matvecs; the explicit LDS address space keeps
// them ds_reads, since address-space inference leaves volatile accesses generic).
template <int NZ>
__device__ __forceinline__ void mvl(const double* T, int lane, const double (&in)[ZG<NZ>::KS], double (&out)[ZG<NZ>::KS]) {
    constexpr int KS = ZG<NZ>::KS;
#pragma unroll
    for (int t = 0; t < ZG<NZ>::RT; ++t) {
        v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const double a = *(const volatile __attribute__((address_space(3))) double*)(
                (const __attribute__((address_space(3))) double*)(T) + (t * KS + s) * 64 + lane);
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, in[s], acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (4 * t + r < KS) out[4 * t + r] = acc[r];
    }
}

// Load a field of plane stride n into slots (0 outside the grid).
template <int NZ, typename TS>
__device__ __forceinline__ void load_slots(const void* f, int64_t c, bool cok, int g, int64_t n, double (&v)[ZG<NZ>::KS]) {
#pragma unroll
    for (int s = 0; s < ZG<NZ>::KS; ++s) {
        const int k = 4 * s + g;
        v[s] = (cok && k < NZ) ? ld<TS>(f, c + n * k) : 0.0;
    }
}

// ---- one RK stage for 16 columns per wave (F1, F3, F4, F2 of :72-76 + RK bookkeeping) ----
//   E1 = 1i*At*dPsi + 0.5i*Psi*dAt + 0.5*f*d2Psi - 1.5*z^2*dPsi - 0.5*z*Psi
//        - 0.5*(Ax^2 + Ay^2)*Psi + 0.5*(dxx+dyy)Psi - 0.5i*Psi*(dxAx+dyAy) - 1i*(Ax*dxPsi + Ay*dyPsi)
//   E3 = 0.5*Lz(dxAt + f*Lz Ax) + 0.5*curl3 + Im(conj(Psi)*dxPsi) - Ax*|Psi|^2
//   E4 = 0.5*Lz(dyAt + f*Lz Ay) + 0.5*curl4 + Im(conj(Psi)*dyPsi) - Ay*|Psi|^2
//   F1 = Lzr(E1), F3 = Lzr(E3), F4 = Lzr(E4), F2 = (1-z^3)*Lz(dxAx + dyAy)
// (the reference's current term -0.5i*(conj(Psi)*dPsi - Psi*d conj(Psi)) equals Im(conj(Psi)*dPsi)).
//
// Each wave's run time is a chain of memory round trips (a 301 x 301 grid is two rounds of
// resident blocks, so that chain, not bandwidth or occupancy, sets the kernel time).  The
// kernel therefore has few serial memory phases, and each phase's loads are issued before the
// previous phase's arithmetic: every field access is a buffer load/store whose out-of-range
// offset reads 0 or drops the write (columns past nxy, rows past nz: no branches), the RK
// bookkeeping selects buffer resources and values by stage instead of branching, the operator
// tiles and row coefficients sit in LDS, and scheduling barriers pin each prefetch to its phase.
// Order: F2 (and the charge update); one point-wise pass over every term of E1, E3, E4 without
// a z derivative; F3; F4; F1.
// Resources of one evolved field's RK bookkeeping at this stage (:120-128):
//   stage 1: acc = k;        next = z + k/2
//   stage 2: acc = acc + 2k; next = z + k/2
//   stage 3: acc = acc + 2k; next = z + k
//   stage 4: z = z + (1/6)*(acc + k)
struct RkRs { Rs acc_ld, z_ld, st1, st2; };
template <typename TS>
__device__ __forceinline__ RkRs rk_rs(int stage, void* acc, void* z, void* nxt, int64_t bytes) {
    const Rs ra = rs_make(acc, bytes), rz = rs_make(z, bytes), rn = rs_make(nxt, bytes), r0 = rs_make(acc, 0);
    return RkRs{stage == 1 ? r0 : ra, rz, stage == 4 ? rz : ra, stage == 4 ? r0 : rn};
}
template <int NZ> struct ZLds {
    static constexpr int KS = ZG<NZ>::KS, RT = ZG<NZ>::RT, OPN = RT * KS * 64, NZP = 4 * KS;
    double lz[OPN], l1[OPN];
    double hf[NZP], cd[NZP], cp[NZP], f1[NZP], f2[NZP];
};

// no instruction crosses this point: keeps each prefetch within its phase
#define ZSB() __builtin_amdgcn_sched_barrier(0)

template <int NZ, typename TS, int MODE>
__device__ __forceinline__ void zstage_body(const ZStage& A, const ZLds<NZ>& Z, int lane, int g, int64_t c, bool cok,
                                            const uint32_t (&e)[ZG<NZ>::KS], double (&sR)[ZG<NZ>::KS],
                                            double (&sI)[ZG<NZ>::KS], double (&dv)[ZG<NZ>::KS],
                                            const double (&at0)[ZG<NZ>::KS]) {
    constexpr int KS = ZG<NZ>::KS;
    const int64_t bt = A.ps * NZ * (int64_t)sizeof(TS), b8 = A.ps * NZ * 8;
    const int stage = A.stage;
    const double h = A.h;
    double w[KS], u[KS], p1r[KS], p1i[KS], p3[KS], p4[KS];
    // point-wise operands in flight during F2
    TS ax[KS], ay[KS], lr[KS], li[KS], xr[KS], xi[KS], yr[KS], yi[KS], c3[KS], c4[KS];
    double dat[KS];
    {
        const Rs rax = rs_make(A.sAx, bt), ray = rs_make(A.sAy, bt);
        const Rs rlr = rs_make(A.lapR, bt), rli = rs_make(A.lapI, bt);
        const Rs rxr = rs_make(A.dxR, bt), rxi = rs_make(A.dxI, bt), ryr = rs_make(A.dyR, bt), ryi = rs_make(A.dyI, bt);
        const Rs rc3 = rs_make(A.curl3, bt), rc4 = rs_make(A.curl4, bt);
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            ax[s] = bldr<TS>(rax, e[s]); ay[s] = bldr<TS>(ray, e[s]);
            lr[s] = bldr<TS>(rlr, e[s]); li[s] = bldr<TS>(rli, e[s]);
            xr[s] = bldr<TS>(rxr, e[s]); xi[s] = bldr<TS>(rxi, e[s]);
            yr[s] = bldr<TS>(ryr, e[s]); yi[s] = bldr<TS>(ryi, e[s]);
            c3[s] = bldr<TS>(rc3, e[s]); c4[s] = bldr<TS>(rc4, e[s]);
        }
    }
    ZSB();
    // ---- F2 (z=0 row feeds the charge update at :128) ----
    mvl<NZ>(Z.lz, lane, dv, w);
    if (MODE == 1) {
        const Rs r = rs_make(A.F2, b8);
#pragma unroll
        for (int s = 0; s < KS; ++s) bst<double>(r, e[s], Z.f2[4 * s + g] * w[s]);
    } else {
        constexpr int sl = (NZ - 1) / 4, gl = (NZ - 1) % 4;
        const double kvl = h * (Z.f2[NZ - 1] * w[sl]);
        const double kv = __shfl(kvl, (lane & 15) + 16 * gl);
        const Rs arl = rs_make(A.arho, stage == 1 ? 0 : A.nxy * 8);
        const Rs ars = rs_make(A.arho, stage == 4 ? 0 : A.nxy * 8);
        const uint32_t ec = cok ? (uint32_t)c : Z_OOB;
        const double ar0 = bld<double>(arl, ec);
        bst<double>(ars, g == 0 ? ec : Z_OOB, stage == 1 ? kv : ar0 + 2.0 * kv);
        const double fin = (1.0 / 6.0) * (ar0 + kv);
        const Rs rr = rs_make(A.rho, stage == 4 ? b8 : 0);
#pragma unroll
        for (int s = 0; s < KS; ++s) bst<double>(rr, e[s], bld<double>(rr, e[s]) - fin);
    }
    ZSB();
    mvl<NZ>(Z.lz, lane, at0, dat);   // d_z A_t of the frozen A_t
    // F3 operands in flight during the point-wise pass
    TS fa[KS], fd[KS];
    {
        const Rs ra = rs_make(A.sAx, bt), rd = rs_make(A.dxAt, bt);
#pragma unroll
        for (int s = 0; s < KS; ++s) { fa[s] = bldr<TS>(ra, e[s]); fd[s] = bldr<TS>(rd, e[s]); }
    }
    ZSB();
    // ---- point-wise parts of E1, E3, E4 ----
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const double sr = sR[s], si = sI[s];
        const double axv = ax[s], ayv = ay[s], dvs = dv[s];
        const double xrv = xr[s], xiv = xi[s], yrv = yr[s], yiv = yi[s];
        const double cp = Z.cp[4 * s + g];
        double er = (-0.5 * si) * dat[s], ei = (0.5 * sr) * dat[s];
        er += cp * sr;               ei += cp * si;
        const double a2 = (-0.5 * axv) * axv, b2 = (-0.5 * ayv) * ayv;
        er += a2 * sr;               ei += a2 * si;
        er += b2 * sr;               ei += b2 * si;
        er += 0.5 * (double)lr[s];   ei += 0.5 * (double)li[s];
        er += (0.5 * si) * dvs;      ei += (-0.5 * sr) * dvs;
        const double qr = axv * xrv + ayv * yrv, qi = axv * xiv + ayv * yiv;
        er += qi;                    ei -= qr;
        p1r[s] = er;
        p1i[s] = ei;
        p3[s] = 0.5 * (double)c3[s] + (sr * xiv - si * xrv) - ((axv * sr) * sr + (axv * si) * si);
        p4[s] = 0.5 * (double)c4[s] + (sr * yiv - si * yrv) - ((ayv * sr) * sr + (ayv * si) * si);
    }
    ZSB();
    // ---- F3, F4 ----
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {
        const double (&pa)[KS] = dir == 0 ? p3 : p4;
        const RkRs R = dir == 0 ? rk_rs<TS>(stage, A.aAx, A.zAx, A.nAx, bt) : rk_rs<TS>(stage, A.aAy, A.zAy, A.nAy, bt);
        TS ra0[KS], rz0[KS];
        if (MODE == 0) {
#pragma unroll
            for (int s = 0; s < KS; ++s) { ra0[s] = bldr<TS>(R.acc_ld, e[s]); rz0[s] = bldr<TS>(R.z_ld, e[s]); }
        }
#pragma unroll
        for (int s = 0; s < KS; ++s) w[s] = fa[s];
        mvl<NZ>(Z.lz, lane, w, u);
#pragma unroll
        for (int s = 0; s < KS; ++s) w[s] = (double)fd[s] + Z.f1[4 * s + g] * u[s];
        if (dir == 0) {
            // F4 operands in flight during the F3 matvecs
            const Rs ra = rs_make(A.sAy, bt), rd = rs_make(A.dyAt, bt);
#pragma unroll
            for (int s = 0; s < KS; ++s) { fa[s] = bldr<TS>(ra, e[s]); fd[s] = bldr<TS>(rd, e[s]); }
        }
        mvl<NZ>(Z.lz, lane, w, u);
#pragma unroll
        for (int s = 0; s < KS; ++s) w[s] = 0.5 * u[s] + pa[s];
        mvl<NZ>(Z.l1, lane, w, u);
        if (MODE == 1) {
            const Rs r = rs_make(dir == 0 ? A.F3 : A.F4, b8);
#pragma unroll
            for (int s = 0; s < KS; ++s) bst<double>(r, e[s], u[s]);
        } else {
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const double kv = h * u[s], a0 = ra0[s], z0 = rz0[s];
                const double v1 = stage == 4 ? z0 + (1.0 / 6.0) * (a0 + kv) : (stage == 1 ? kv : a0 + 2.0 * kv);
                const double v2 = z0 + (stage == 3 ? kv : kv * 0.5);
                bst<TS>(R.st1, e[s], v1);
                bst<TS>(R.st2, e[s], v2);
            }
        }
        ZSB();
    }
    // ---- F1 ----
    {
        const RkRs RR = rk_rs<TS>(stage, A.aR, A.zR, A.nR, bt), RI = rk_rs<TS>(stage, A.aI, A.zI, A.nI, bt);
        double atv[KS];
        TS aR0[KS], zR0[KS], aI0[KS], zI0[KS];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            atv[s] = at0[s];
            if (MODE == 0) {
                aR0[s] = bldr<TS>(RR.acc_ld, e[s]); zR0[s] = bldr<TS>(RR.z_ld, e[s]);
                aI0[s] = bldr<TS>(RI.acc_ld, e[s]); zI0[s] = bldr<TS>(RI.z_ld, e[s]);
            }
        }
        double dR[KS], dI[KS], d2R[KS], d2I[KS];
        mvl<NZ>(Z.lz, lane, sR, dR);
        mvl<NZ>(Z.lz, lane, sI, dI);
        mvl<NZ>(Z.lz, lane, dR, d2R);
        mvl<NZ>(Z.lz, lane, dI, d2I);
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = 4 * s + g;
            const double at = atv[s];
            const double hf = Z.hf[k], cd = Z.cd[k];
            double er = p1r[s] + hf * d2R[s], ei = p1i[s] + hf * d2I[s];
            er += cd * dR[s];            ei += cd * dI[s];
            er -= at * dI[s];            ei += at * dR[s];
            // rows past nz carry no data: keep them exactly zero for the Lzr matvec
            p1r[s] = k < NZ ? er : 0.0;
            p1i[s] = k < NZ ? ei : 0.0;
        }
        mvl<NZ>(Z.l1, lane, p1r, dR);     // F1 = Lzr(E1); the last row of Lzr is zero
        mvl<NZ>(Z.l1, lane, p1i, dI);
        if (MODE == 1) {
            const Rs rr = rs_make(A.F1R, b8), ri = rs_make(A.F1I, b8);
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                bst<double>(rr, e[s], dR[s]);
                bst<double>(ri, e[s], dI[s]);
            }
        } else {
#pragma unroll
            for (int s = 0; s < KS; ++s) {
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const RkRs& R = q == 0 ? RR : RI;
                    const double kv = h * (q == 0 ? dR[s] : dI[s]);
                    const double a0 = q == 0 ? aR0[s] : aI0[s], z0 = q == 0 ? zR0[s] : zI0[s];
                    const double v1 = stage == 4 ? z0 + (1.0 / 6.0) * (a0 + kv) : (stage == 1 ? kv : a0 + 2.0 * kv);
                    const double v2 = z0 + (stage == 3 ? kv : kv * 0.5);
                    bst<TS>(R.st1, e[s], v1);
                    bst<TS>(R.st2, e[s], v2);
                }
            }
        }
    }
}

// grid: x = ceil(nxy/64) blocks of 4 waves, 16 columns per wave.  WPE: the occupancy the
// register allocation targets (2: no spills up to nz = 20; 3 trades spills for occupancy).
template <int NZ, typename TS, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_zstage(const ZStage* __restrict__ Ap,
                                                                                            const Members M) {
    constexpr int KS = ZG<NZ>::KS;
    // XCD-aware block order: neighbouring column blocks share cache lines at their edges
    const MemberBlock mb = member_block_at(M, xcd_block((int)blockIdx.x, (int)gridDim.x), (int)gridDim.x);
    const ZStage& A = Ap[mb.m];
    const int stop = __builtin_amdgcn_readfirstlane(__builtin_nontemporal_load(&A.flags[0]));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int64_t c = (int64_t)mb.bx * 64 + wave * 16 + m;
    const bool cok = c < A.nxy;
    uint32_t e[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + g;
        e[s] = (cok && k < NZ) ? (uint32_t)(c + A.ps * k) : Z_OOB;
    }
    // first operands in flight before the operator staging
    const int64_t bt = A.ps * NZ * (int64_t)sizeof(TS);
    const Rs rsr = rs_make(A.sR, bt), rsi = rs_make(A.sI, bt), rdv = rs_make(A.divA, bt);
    double sR[KS], sI[KS], dv[KS], at0[KS];
    const Rs rat0 = rs_make(A.At, A.ps * NZ * 8);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        at0[s] = bld<double>(rat0, e[s]);
        dv[s] = bld<TS>(rdv, e[s]);
        sR[s] = bld<TS>(rsr, e[s]);
        sI[s] = bld<TS>(rsi, e[s]);
    }
    __shared__ ZLds<NZ> Z;
    stage_op<NZ>(A.zo.Lz, NZ, NZ, Z.lz);
    stage_op<NZ>(A.zo.Lz1, NZ - 1, NZ - 1, Z.l1);
    for (int i = threadIdx.x; i < 5 * ZLds<NZ>::NZP; i += 256) {
        const int q = i / ZLds<NZ>::NZP, k = i % ZLds<NZ>::NZP;
        const double* src = q == 0 ? A.zo.hf : q == 1 ? A.zo.cd : q == 2 ? A.zo.cp : q == 3 ? A.zo.f1 : A.zo.f2;
        double* dst = q == 0 ? Z.hf : q == 1 ? Z.cd : q == 2 ? Z.cp : q == 3 ? Z.f1 : Z.f2;
        dst[k] = k < NZ ? src[k] : 0.0;
    }
    __syncthreads();
    if (stop != 0) return;           // same value for the whole grid: the block returns together
    // the step counter advances at the step's first radial stage (no kernel before it in the step
    // reports a NaN step: the previous step's guard has already checked the state it reads)
    if (A.mode == 0 && A.stage == 1 && mb.bx == 0 && threadIdx.x == 0) A.flags[1] = A.flags[1] + 1;
    if (A.mode == 1) zstage_body<NZ, TS, 1>(A, Z, lane, g, c, cok, e, sR, sI, dv, at0);
    else zstage_body<NZ, TS, 0>(A, Z, lane, g, c, cok, e, sR, sI, dv, at0);
}

// ---- post-step: A_t constraint solve and boundary reset (:129-133), O for the winding (:136)
template <int NZ, typename TS>
__global__ __launch_bounds__(256) void k_zpost(const ZPost* __restrict__ Ap, const Members M) {
    const MemberBlock mb = member_block_at(M, xcd_block((int)blockIdx.x, (int)gridDim.x), (int)gridDim.x);
    const ZPost& A = Ap[mb.m];
    if (stopped(A.flags)) return;
    constexpr int KS = ZG<NZ>::KS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int64_t c = (int64_t)mb.bx * 64 + wave * 16 + m;
    const bool cok = c < A.nxy;
    const int64_t n = A.ps;
    double Lz[ZG<NZ>::RT][KS], L1[ZG<NZ>::RT][KS];
    load_op<NZ>(A.zo.Lz, NZ, NZ, Lz, m, g);
    load_op<NZ>(A.zo.Lz1, NZ - 1, NZ - 1, L1, m, g);
    double R[KS], I[KS], a[KS], b[KS], w[KS];
    load_slots<NZ, TS>(A.R, c, cok, g, n, R);
    load_slots<NZ, TS>(A.I, c, cok, g, n, I);
    load_slots<NZ, TS>(A.divA, c, cok, g, n, a);
    load_slots<NZ, TS>(A.divB, c, cok, g, n, b);
    // rho and the boundary values are loaded with the fields: one memory round trip for the kernel
    double rh[KS];
    load_slots<NZ, double>(A.rho, c, cok, g, n, rh);
    const double axb = cok ? ld<double>(A.axb, c) : 0.0, ayb = cok ? ld<double>(A.ayb, c) : 0.0;
#pragma unroll
    for (int s = 0; s < KS; ++s) a[s] += b[s];
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
        w[s] = (cok && k < NZ) ? a[s] - rh[s] : 0.0;
    }
    mv<NZ>(L1, w, a);                // Lzr(Lzr(cons) - rhoi)
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + g;
        if (!cok || k >= NZ) continue;
        st<double>(A.At, c + n * k, (k < NZ - 1) ? a[s] + A.mu : A.mu);
        if (k == NZ - 1) {
            // boundary reset: Z1(:,:,nz)=0; Z3(:,:,nz)=ax_b; Z4(:,:,nz)=ay_b  (Z2 row set above)
            const int64_t pb = c + n * k;
            st<TS>(A.R, pb, 0.0); st<TS>(A.I, pb, 0.0);
            st<TS>(A.Ax, pb, axb); st<TS>(A.Ay, pb, ayb);
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
        if (cok && k == NZ - 1) { st<double>(A.O, 2 * c, a[s]); st<double>(A.O, 2 * c + 1, b[s]); }
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

}  // namespace hqt
