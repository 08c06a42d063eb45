// Boundary diagnostics on the device: the condensate O = (func_Lz(Z1))(:,:,end)
// (tevol_driven_cluster.m:136,148) and the plaquette vortex census of
// src/diagnostics/countN.m:7-21 (charges as find_vortices.m:103-131), so that error metrics
// need only O or a charge map downloaded, not the fields.
#include "hqt_angle.h"
#include "hqt_kernels.h"

namespace hqt {

constexpr int DG_T = 256;

// O(i,j) = sum_m Lz(nz,m) Psi(i,j,m): one thread per column, coalesced over the plane.
template <typename TS>
__global__ __launch_bounds__(DG_T) void k_order_param(const Census A) {
    const int64_t p = (int64_t)blockIdx.x * DG_T + threadIdx.x;     // point (i, j) of the plane, i + nx j
    if (p >= A.nxy) return;
    const int64_t col = p % A.nx + A.rp * (p / A.nx);                // its column in the padded layout
    const TS* R = (const TS*)A.R;
    const TS* I = (const TS*)A.I;
    double a = 0.0, b = 0.0;
    for (int m = 0; m < A.nz; ++m) {
        const double l = A.Lzb[m];
        a += l * (double)R[col + A.ps * m];
        b += l * (double)I[col + A.ps * m];
    }
    A.O[2 * p] = a;
    A.O[2 * p + 1] = b;
}

// Plaquette (i,j) with corners x1=O(i,j), x2=O(i+1,j), x3=O(i+1,j+1), x4=O(i,j+1) (periodic):
// w = (angle(x1/x2)+angle(x2/x3)+angle(x3/x4)+angle(x4/x1))/(2 pi); +1 when |w-1| <= 1e-4,
// -1 when |w+1| <= 1e-4 (countN.m:9-13).  Per block: counts of +1/-1 (integer atomics) and
// the partial max / sum of |O| over the block's points (fixed-order reduction, summed on the
// host in block order).
__global__ __launch_bounds__(DG_T) void k_census(const Census A) {
    __shared__ double smax[DG_T], ssum[DG_T];
    __shared__ int sp[DG_T], sm[DG_T];
    const int64_t p = (int64_t)blockIdx.x * DG_T + threadIdx.x;
    int cp = 0, cm = 0;
    double amax = 0.0, asum = 0.0;
    if (p < A.nxy) {
        const int i = (int)(p % A.nx), j = (int)(p / A.nx);
        const int i1 = (i + 1 == A.nx) ? 0 : i + 1, j1 = (j + 1 == A.ny) ? 0 : j + 1;
        const double* O = A.O;
        const int64_t k1 = p, k2 = i1 + (int64_t)A.nx * j, k3 = i1 + (int64_t)A.nx * j1, k4 = i + (int64_t)A.nx * j1;
        const double x1r = O[2 * k1], x1i = O[2 * k1 + 1], x2r = O[2 * k2], x2i = O[2 * k2 + 1];
        const double x3r = O[2 * k3], x3i = O[2 * k3 + 1], x4r = O[2 * k4], x4i = O[2 * k4 + 1];
        const double w = (angle_ratio(x1r, x1i, x2r, x2i) + angle_ratio(x2r, x2i, x3r, x3i) +
                          angle_ratio(x3r, x3i, x4r, x4i) + angle_ratio(x4r, x4i, x1r, x1i)) *
                         (1.0 / (2.0 * 3.14159265358979323846));
        cp = (w >= 1 - 1e-4 && w <= 1 + 1e-4) ? 1 : 0;
        cm = (w <= -1 + 1e-4 && w >= -1 - 1e-4) ? 1 : 0;
        if (A.charge) A.charge[p] = (int8_t)(cp - cm);
        amax = sqrt(x1r * x1r + x1i * x1i);
        asum = amax;
    }
    smax[threadIdx.x] = amax;
    ssum[threadIdx.x] = asum;
    sp[threadIdx.x] = cp;
    sm[threadIdx.x] = cm;
    for (int off = DG_T / 2; off > 0; off >>= 1) {
        __syncthreads();
        if ((int)threadIdx.x < off) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + off]);
            ssum[threadIdx.x] += ssum[threadIdx.x + off];
            sp[threadIdx.x] += sp[threadIdx.x + off];
            sm[threadIdx.x] += sm[threadIdx.x + off];
        }
    }
    if (threadIdx.x == 0) {
        A.part[2 * blockIdx.x] = smax[0];
        A.part[2 * blockIdx.x + 1] = ssum[0];
        if (sp[0]) atomicAdd(&A.counts[0], sp[0]);
        if (sm[0]) atomicAdd(&A.counts[1], sm[0]);
    }
}

int census_blocks(int64_t nxy) { return (int)((nxy + DG_T - 1) / DG_T); }

hipError_t launch_order_param(const Census& A, int fp32, hipStream_t s) {
    const dim3 g(census_blocks(A.nxy)), b(DG_T);
    if (fp32) hipLaunchKernelGGL(k_order_param<float>, g, b, 0, s, A);
    else hipLaunchKernelGGL(k_order_param<double>, g, b, 0, s, A);
    return hipGetLastError();
}

hipError_t launch_census(const Census& A, hipStream_t s) {
    hipLaunchKernelGGL(k_census, dim3(census_blocks(A.nxy)), dim3(DG_T), 0, s, A);
    return hipGetLastError();
}

}  // namespace hqt
