// The Fourier filter for any logical mask, for the stand-alone ff entry point (src/numerics/ff.m:7-11):
// w = real(ifft2(b)), b = fft2(x) with the masked modes zeroed, per slice.  Direct DFTs in FP64, one thread per
// output value: along x (real or complex input), along y (then the mask), back along y, back along x (only the
// real part, scaled by 1 / (nx ny)).  O(nx ny (nx + ny)) per slice: this is the general-mask utility, not the
// stepper's filter (the stepper's mask fx | fy, tevol_driven_cluster.m:27-31, is separable and runs through the
// factored line passes).
#include "hqt_kernels.h"

namespace hqt {

namespace {

constexpr int kFfmThreads = 256;

// tw[2 m], tw[2 m + 1] = cos, sin (2 pi m / N), m < N.  Every DFT index (j k mod N) is advanced by adding k and
// wrapping, so no product j k is formed.

// out[kx + nx l] = sum_x (re + i im)[x + nx l] exp(-2 pi i x kx / nx), l = y + ny s; im NULL for a real input
__global__ __launch_bounds__(kFfmThreads) void k_ffm_fwd_x(const double* __restrict__ re, const double* __restrict__ im,
                                                            double2* __restrict__ out, const double* __restrict__ tw,
                                                            int nx, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kFfmThreads + threadIdx.x;
    if (i >= n) return;
    const int kx = (int)(i % nx);
    const int64_t b = i - kx;
    double ar = 0.0, ai = 0.0;
    int m = 0;
    for (int x = 0; x < nx; ++x) {
        const double c = tw[2 * m], s = tw[2 * m + 1];
        const double xr = re[b + x], xi = im ? im[b + x] : 0.0;
        ar = fma(xr, c, fma(xi, s, ar));          // (xr + i xi)(c - i s)
        ai = fma(xi, c, fma(-xr, s, ai));
        m += kx;
        if (m >= nx) m -= nx;
    }
    out[i] = make_double2(ar, ai);
}

// out[kx + nx (ky + ny s)] = sum_y in[kx + nx (y + ny s)] exp(sg 2 pi i y ky / ny), sg = -1 forward, +1 inverse;
// forward: zero where mask[kx + nx ky + mstride s] (mstride 0: one mask for every slice)
template <bool FWD>
__global__ __launch_bounds__(kFfmThreads) void k_ffm_y(const double2* __restrict__ in, double2* __restrict__ out,
                                                        const double* __restrict__ tw, const uint8_t* __restrict__ mask,
                                                        int nx, int ny, int64_t mstride, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kFfmThreads + threadIdx.x;
    if (i >= n) return;
    const int kx = (int)(i % nx);
    const int64_t r = i / nx;
    const int ky = (int)(r % ny);
    const int64_t s = r / ny;
    const int64_t b = kx + (int64_t)nx * ny * s;
    if (FWD && mask[kx + (int64_t)nx * ky + mstride * s]) {
        out[i] = make_double2(0.0, 0.0);
        return;
    }
    const double sg = FWD ? -1.0 : 1.0;
    double ar = 0.0, ai = 0.0;
    int m = 0;
    for (int y = 0; y < ny; ++y) {
        const double c = tw[2 * m], sn = sg * tw[2 * m + 1];
        const double2 v = in[b + (int64_t)nx * y];
        ar = fma(v.x, c, fma(-v.y, sn, ar));      // (vr + i vi)(c + i sn)
        ai = fma(v.y, c, fma(v.x, sn, ai));
        m += ky;
        if (m >= ny) m -= ny;
    }
    out[i] = make_double2(ar, ai);
}

// w[x + nx l] = Re(sum_kx in[kx + nx l] exp(2 pi i x kx / nx)) * scale
__global__ __launch_bounds__(kFfmThreads) void k_ffm_inv_x(const double2* __restrict__ in, double* __restrict__ w,
                                                            const double* __restrict__ tw, int nx, double scale,
                                                            int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kFfmThreads + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % nx);
    const int64_t b = i - x;
    double ar = 0.0;
    int m = 0;
    for (int kx = 0; kx < nx; ++kx) {
        const double2 v = in[b + kx];
        ar = fma(v.x, tw[2 * m], fma(-v.y, tw[2 * m + 1], ar));
        m += x;
        if (m >= nx) m -= nx;
    }
    w[i] = ar * scale;
}

}  // namespace

hipError_t launch_ff_mask(const double* re, const double* im, const uint8_t* mask, int64_t mstride, const double* twx,
                          const double* twy, double2* a, double2* b, double* w, int nx, int ny, int nslices,
                          hipStream_t s) {
    const int64_t n = (int64_t)nx * ny * nslices;
    if (nx < 1 || ny < 1 || nslices < 1) return hipErrorInvalidValue;
    const unsigned blocks = (unsigned)((n + kFfmThreads - 1) / kFfmThreads);
    hipLaunchKernelGGL(k_ffm_fwd_x, dim3(blocks), dim3(kFfmThreads), 0, s, re, im, a, twx, nx, n);
    hipLaunchKernelGGL(k_ffm_y<true>, dim3(blocks), dim3(kFfmThreads), 0, s, a, b, twy, mask, nx, ny, mstride, n);
    hipLaunchKernelGGL(k_ffm_y<false>, dim3(blocks), dim3(kFfmThreads), 0, s, b, a, twy, mask, nx, ny, mstride, n);
    hipLaunchKernelGGL(k_ffm_inv_x, dim3(blocks), dim3(kFfmThreads), 0, s, a, w, twx, nx, 1.0 / ((double)nx * ny), n);
    return hipGetLastError();
}

}  // namespace hqt
