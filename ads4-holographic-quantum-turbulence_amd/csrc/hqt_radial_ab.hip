// The radial contraction in two forms, for a measured comparison: the FP64 MFMA matvec on the register slot
// layout that the radial stage kernel uses (mvl, hqt_radial_core.h) against a vector-FMA matvec fed through LDS
// (the LDS/wavefront path).  Each kernel applies the Chebyshev derivative Lz twice to every column of one field
// (u = Lz x, v = Lz u: func_Lz and func_Lz(func_Lz(.)) of tevol_driven_cluster.m:62, the chain F1 uses), with
// the columns laid out as the volume fields are (plane stride ps, 16 consecutive columns per lane group).
// Measurement helper with no counterpart in the reference (hqt_calibration_radial).
#include "hqt_radial_core.h"
#include "hqt_internal.h"
#include "hqt_operators.h"

#include <cmath>
#include <string>
#include <vector>

namespace hqt {
namespace {

// out = A * in with vector FMAs: the wave's 16 columns go to its LDS column buffer cb (NZP x 16), and each lane
// forms its KS rows from the row-major operator Ar (NZP x NZP, zero outside NZ x NZ) and the buffer.
template <int NZ>
__device__ __forceinline__ void mv_lds(const double* Ar, double* cb, int lane, const double (&in)[ZG<NZ>::KS],
                                       double (&out)[ZG<NZ>::KS]) {
    constexpr int KS = ZG<NZ>::KS, NZP = 4 * KS;
    const int m = lane & 15, g = lane >> 4;
#pragma unroll
    for (int s = 0; s < KS; ++s) cb[(4 * s + g) * 16 + m] = in[s];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const double* row = Ar + (4 * s + g) * NZP;
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < NZ; ++j) acc = __builtin_fma(row[j], cb[j * 16 + m], acc);
        out[s] = acc;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NZ, bool VALU>
__global__ __launch_bounds__(256) void k_radial_ab(const double* __restrict__ x, double* __restrict__ u,
                                                   double* __restrict__ v, const double* __restrict__ Lz, int64_t ncol,
                                                   int64_t ps) {
    constexpr int KS = ZG<NZ>::KS, RT = ZG<NZ>::RT, NZP = 4 * KS;
    __shared__ double tiles[VALU ? 1 : RT * KS * 64];
    __shared__ double rowm[VALU ? NZP * NZP : 1];
    __shared__ double cbuf[VALU ? 4 * NZP * 16 : 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, m = lane & 15, g = lane >> 4;
    if constexpr (VALU) {
        for (int idx = threadIdx.x; idx < NZP * NZP; idx += 256) {
            const int i = idx / NZP, j = idx % NZP;
            rowm[idx] = (i < NZ && j < NZ) ? Lz[i * NZ + j] : 0.0;
        }
    } else {
        stage_op<NZ>(Lz, NZ, NZ, tiles);
    }
    __syncthreads();
    const int64_t c = (int64_t)blockIdx.x * 64 + wave * 16 + m;
    const bool cok = c < ncol;
    double xs[KS], us[KS], vs[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + g;
        xs[s] = (cok && k < NZ) ? x[c + ps * k] : 0.0;
    }
    if constexpr (VALU) {
        mv_lds<NZ>(rowm, cbuf + wave * NZP * 16, lane, xs, us);
        mv_lds<NZ>(rowm, cbuf + wave * NZP * 16, lane, us, vs);
    } else {
        mvl<NZ>(tiles, lane, xs, us);
        mvl<NZ>(tiles, lane, us, vs);
    }
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + g;
        if (cok && k < NZ) {
            u[c + ps * k] = us[s];
            v[c + ps * k] = vs[s];
        }
    }
}

__global__ void k_radial_ab_fill(double* x, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = sin(0.001 * (double)i) + 0.5 * cos(0.013 * (double)i);
}

template <int NZ>
hipError_t radial_ab_launch(bool valu, const double* x, double* u, double* v, const double* Lz, int64_t ncol, int64_t ps,
                            hipStream_t s) {
    const int blocks = (int)((ncol + 63) / 64);
    if (valu) hipLaunchKernelGGL((k_radial_ab<NZ, true>), dim3(blocks), dim3(256), 0, s, x, u, v, Lz, ncol, ps);
    else hipLaunchKernelGGL((k_radial_ab<NZ, false>), dim3(blocks), dim3(256), 0, s, x, u, v, Lz, ncol, ps);
    return hipGetLastError();
}

}  // namespace
}  // namespace hqt

using namespace hqt;

// ncol columns of nz points (nz = 12 or 20), `form` 0: MFMA, 1: vector FMAs through LDS; `reps` timed launches
// after one warm-up launch; ms = average duration per launch; sumsq = sum of squares of v (the second derivative)
// over the field, to compare the two forms.
extern "C" int hqt_calibration_radial(int64_t ncol, int32_t nz, int32_t form, int32_t reps, double* ms, double* sumsq) {
    if (!ms || !sumsq || ncol <= 0 || ncol > (1 << 22) || (nz != 12 && nz != 20) || (form != 0 && form != 1) ||
        reps < 1 || reps > 10000)
        return hqt_fail(HQT_ERR_ARG, "hqt_calibration_radial: ncol in [1, 2^22], nz 12 or 20, form 0 or 1, reps in [1, 10000]");
    const int64_t ps = (ncol + 31) / 32 * 32;
    std::vector<double> zc, DM;
    chebdif(nz, 1, zc, DM);
    std::vector<double> lz((size_t)nz * nz);
    for (int i = 0; i < nz; ++i)
        for (int j = 0; j < nz; ++j) lz[(size_t)i * nz + j] = 2.0 * DM[(size_t)j * nz + i];   // Lz = 2 Dz (z in [0, 1])
    double *x = nullptr, *u = nullptr, *v = nullptr, *L = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s = nullptr;
    std::string err;
    hipError_t e = hipMalloc(&x, (size_t)ps * nz * 8);
    if (e == hipSuccess) e = hipMalloc(&u, (size_t)ps * nz * 8);
    if (e == hipSuccess) e = hipMalloc(&v, (size_t)ps * nz * 8);
    if (e == hipSuccess) e = hipMalloc(&L, lz.size() * 8);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipMemcpyAsync(L, lz.data(), lz.size() * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_radial_ab_fill, dim3(1024), dim3(256), 0, s, x, ps * nz);
        e = hipGetLastError();
    }
    auto launch = [&]() {
        return nz == 12 ? radial_ab_launch<12>(form == 1, x, u, v, L, ncol, ps, s)
                        : radial_ab_launch<20>(form == 1, x, u, v, L, ncol, ps, s);
    };
    if (e == hipSuccess) e = launch();
    if (e == hipSuccess) e = hipEventRecord(a, s);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch();
    if (e == hipSuccess) e = hipEventRecord(b, s);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    std::vector<double> hv;
    if (e == hipSuccess) {
        hv.resize((size_t)ps * nz);
        e = hipMemcpy(hv.data(), v, hv.size() * 8, hipMemcpyDeviceToHost);
    }
    int rc = HQT_OK;
    if (e != hipSuccess) {
        rc = hqt_fail(HQT_ERR_HIP, std::string("hqt_calibration_radial: ") + hipGetErrorString(e));
    } else {
        *ms = (double)t / reps;
        double acc = 0.0;
        for (int k = 0; k < nz; ++k)
            for (int64_t c = 0; c < ncol; ++c) acc += hv[(size_t)(c + ps * k)] * hv[(size_t)(c + ps * k)];
        *sumsq = acc;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(x); (void)hipFree(u); (void)hipFree(v); (void)hipFree(L);
    return rc;
}
