// Radial kernels: nz dispatch and the small kernels that need no compile-time nz (d_z A_t,
// winding removal, NaN guard).  The per-nz templates live in hqt_radial_core.h and are
// instantiated by the hqt_radial_nz_*.hip translation units so the build runs in parallel.
#include "hqt_radial_core.h"
#include "hqt_angle.h"

namespace hqt {

// ---- winding: round(mean(sum(angle(circshift(O,-1,2)./O),1))/(2*pi)) (:137) --------------
__device__ __forceinline__ double block_sum_256(double v, double* sh) {
    sh[threadIdx.x] = v;
    for (int off = 128; off > 0; off >>= 1) {
        __syncthreads();
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    }
    __syncthreads();
    const double r = sh[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void k_wind_partial(const Wind* __restrict__ Ap, const Members M) {
    __shared__ double sh[256];
    const MemberBlock mb = member_block(M);
    const Wind& A = Ap[mb.m];
    if (stopped(A.flags)) return;
    const int j = mb.bx, jn = (j + 1) % A.ny;
    double s = 0.0;
    for (int x = threadIdx.x; x < A.nx; x += 256) {
        const int64_t a = (int64_t)x + A.rp * jn, b = (int64_t)x + A.rp * j;
        s += angle_ratio(A.O[2 * a], A.O[2 * a + 1], A.O[2 * b], A.O[2 * b + 1]);
    }
    const double t = block_sum_256(s, sh);
    if (threadIdx.x == 0) A.part[j] = t;
}

template <typename TS>
__global__ __launch_bounds__(256) void k_wind_apply(const Wind* __restrict__ Ap, const Members M) {
    __shared__ double sh[256];
    const MemberBlock mb = member_block(M);
    const Wind& A = Ap[mb.m];
    if (stopped(A.flags)) return;
    double s = 0.0;
    for (int j = threadIdx.x; j < A.ny; j += 256) s += A.part[j];
    const double tot = block_sum_256(s, sh);
    const double W = round((tot / A.ny) / (2.0 * M_PI));
    if (mb.bx == 0 && threadIdx.x == 0) A.W[0] = W;
    if (W == 0.0) return;
    const int64_t total = A.ps * A.nz;
    for (int64_t e = (int64_t)mb.bx * 256 + threadIdx.x; e < total; e += (int64_t)mb.nb * 256) {
        const int64_t c = e % A.ps;
        if (c >= A.nxy) continue;        // plane padding
        const int y = (int)(c / A.rp);
        double sn, cs;
        sincos(-A.yph[y] * W, &sn, &cs);
        const double r = ld<TS>(A.R, e), i = ld<TS>(A.I, e);
        st<TS>(A.R, e, r * cs - i * sn);
        st<TS>(A.I, e, r * sn + i * cs);
    }
}

// ---------------------------------------------------------------------------------------
#define HQT_NZ_LIST(X) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) \
    X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31) X(32)

bool zkernels_support_nz(int nz) { return nz >= 4 && nz <= 32; }

#define HQT_DECL(n) extern template hipError_t zstage_launch_nz<n>(const ZStage*, Members, int, int, int, hipStream_t); \
                    extern template hipError_t zpost_launch_nz<n>(const ZPost*, Members, int, int, hipStream_t);
HQT_NZ_LIST(HQT_DECL)
#undef HQT_DECL

hipError_t launch_zstage(const ZStage* d, Members M, int blocks, int nz, int fp32, int wpe, hipStream_t s) {
    switch (nz) {
#define HQT_ZS(n) case n: return zstage_launch_nz<n>(d, M, blocks, fp32, wpe, s);
        HQT_NZ_LIST(HQT_ZS)
#undef HQT_ZS
        default: return hipErrorInvalidValue;
    }
}
hipError_t launch_zpost(const ZPost* d, Members M, int blocks, int nz, int fp32, hipStream_t s) {
    switch (nz) {
#define HQT_ZP(n) case n: return zpost_launch_nz<n>(d, M, blocks, fp32, s);
        HQT_NZ_LIST(HQT_ZP)
#undef HQT_ZP
        default: return hipErrorInvalidValue;
    }
}


int wind_apply_blocks(int64_t nxy, int nz) {
    const int64_t b = (nxy * nz + 255) / 256;
    return (int)(b > 2048 ? 2048 : b);
}

hipError_t launch_wind_partial(const Wind* d, Members M, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_wind_partial, dim3(blocks), dim3(256), 0, s, d, M);
    return hipGetLastError();
}

hipError_t launch_wind_apply(const Wind* d, Members M, int blocks, int fp32, hipStream_t s) {
    if (fp32) hipLaunchKernelGGL(k_wind_apply<float>, dim3(blocks), dim3(256), 0, s, d, M);
    else hipLaunchKernelGGL(k_wind_apply<double>, dim3(blocks), dim3(256), 0, s, d, M);
    return hipGetLastError();
}

}  // namespace hqt

namespace hqt {
template <typename TS>
__global__ __launch_bounds__(256) void k_nancheck(const NanCheck* __restrict__ Ap, const Members M) {
    const MemberBlock mb = member_block(M);
    const NanCheck& A = Ap[mb.m];
    if (stopped(A.flags)) return;
    bool bad = false;
    for (int64_t e = (int64_t)mb.bx * 256 + threadIdx.x; e < A.n; e += (int64_t)mb.nb * 256) {
        const double a = ld<TS>(A.R, e), b = ld<TS>(A.I, e), c = ld<TS>(A.Ax, e), d = ld<TS>(A.Ay, e), f = A.At[e];
        if (a != a || b != b || c != c || d != d || f != f) bad = true;
    }
    if (__any(bad) && (threadIdx.x & 63) == 0) {
        const int step = __atomic_load_n(&A.flags[1], __ATOMIC_RELAXED);
        atomicCAS(&A.flags[0], 0, step > 0 ? step : 1);
    }
}

int nancheck_blocks(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (int)(b > 4096 ? 4096 : b);
}

hipError_t launch_nancheck(const NanCheck* d, Members M, int blocks, int fp32, hipStream_t s) {
    if (fp32) hipLaunchKernelGGL(k_nancheck<float>, dim3(blocks), dim3(256), 0, s, d, M);
    else hipLaunchKernelGGL(k_nancheck<double>, dim3(blocks), dim3(256), 0, s, d, M);
    return hipGetLastError();
}
}  // namespace hqt

namespace hqt {
__global__ __launch_bounds__(256) void k_copy_calib(const double4* __restrict__ src, double4* __restrict__ dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}
hipError_t launch_copy_calib(const void* src, void* dst, int64_t bytes, hipStream_t s) {
    const int64_t n = bytes / (int64_t)sizeof(double4);
    hipLaunchKernelGGL(k_copy_calib, dim3(16384), dim3(256), 0, s, (const double4*)src, (double4*)dst, n);
    return hipGetLastError();
}
}  // namespace hqt
