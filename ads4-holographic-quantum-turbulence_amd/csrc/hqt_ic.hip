// Initial state built on the device (tevol_driven_cluster.m:81-102; tevol_lattice.m:101-135).
//
// Host part (O(nz) work): the stationary profile on the chebdif(nz0) grid is splined to the
// run's radial nodes (interp1 'spline', not-a-knot, :84), the gauge phase
// theta = -Lzr(A_t/f) with the horizon limit -d_z A_t/3 is formed (:90-93), and the boundary
// charge rho(z=0) is taken (:98).  Device part (O(nx ny nz)): every column gets the profile
// times a boundary-plane pattern
//     amp(x,y)   = (1 + amp_noise g1) (1 - sum_v exp(-|r - r_v|^2 / sigma^2))      (:101; lattice :113-121)
//     phase(x,y) = phase_noise g2 + sum_v q_v atan2(y - Y_v, x - X_v)              (lattice :125-133)
//     Psi = Z1(z) e^{i theta(z)} amp e^{i phase},  A_t = Z2(z),  A_x = 0,  A_y = -Vy1(z),  rho = rho(z=0)
// with (g1, g2) standard normal deviates from Philox4x32-10 (Salmon et al., SC'11) keyed by
// the seed and counted by the column index, then Box-Muller, so the noise does not depend on
// the launch shape or the device.
#include "hqt_internal.h"
#include "hqt_operators.h"

#include <algorithm>
#include <cmath>

namespace hqt {

struct ICDev {
    int32_t nx, ny, nz, nv;
    int64_t nxy, ps, rp;  // points of a plane (nx ny); plane and row stride of the volume fields
    double Lx, Ly, amp_noise, phase_noise, inv_sigma2;
    uint32_t key0, key1;
    const double* prof;   // [4*nz]: Re(Z1 e^{i theta}), Im(Z1 e^{i theta}), Z2, -Vy1
    const double* vort;   // [3*nv]: X, Y, q
    double rho0;
    void *R, *I, *Ax, *Ay;
    double *At, *rho;
};

// Philox4x32-10: ten rounds of two 32x32->64 multiplies with the Weyl key schedule.
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

template <typename TS>
__global__ __launch_bounds__(256) void k_init_state(const ICDev A) {
    const int64_t col = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (col >= A.nxy) return;
    const int i = (int)(col % A.nx), j = (int)(col / A.nx);
    const double x = i * A.Lx / A.nx, y = j * A.Ly / A.ny;      // xx, yy (tevol_driven_cluster.m:42-43)
    double amp = 1.0, phase = 0.0;
    if (A.amp_noise != 0.0 || A.phase_noise != 0.0) {
        uint32_t c[4] = {(uint32_t)col, (uint32_t)(col >> 32), 0u, 0u};
        philox4x32_10(c, A.key0, A.key1);
        const double r1 = ((double)((((uint64_t)(c[0] >> 6)) << 27) | (c[1] >> 5)) + 0.5) * 0x1p-53;
        const double r2 = ((double)((((uint64_t)(c[2] >> 6)) << 27) | (c[3] >> 5)) + 0.5) * 0x1p-53;
        const double rad = sqrt(-2.0 * log(r1));
        double sn, cs;
        sincos(2.0 * 3.14159265358979323846 * r2, &sn, &cs);
        amp += A.amp_noise * rad * cs;
        phase += A.phase_noise * rad * sn;
    }
    double dip = 0.0;
    for (int v = 0; v < A.nv; ++v) {
        const double dx = x - A.vort[3 * v], dy = y - A.vort[3 * v + 1];
        phase += A.vort[3 * v + 2] * atan2(dy, dx);
        if (A.inv_sigma2 > 0.0) dip += exp(-(dx * dx + dy * dy) * A.inv_sigma2);
    }
    amp *= 1.0 - dip;
    double ps, pc;
    sincos(phase, &ps, &pc);
    const double ar = amp * pc, ai = amp * ps;
    TS* R = (TS*)A.R;
    TS* I = (TS*)A.I;
    TS* X = (TS*)A.Ax;
    TS* Y = (TS*)A.Ay;
    for (int k = 0; k < A.nz; ++k) {
        const int64_t p = i + A.rp * j + A.ps * k;
        const double zr = A.prof[k], zi = A.prof[A.nz + k];
        R[p] = (TS)(zr * ar - zi * ai);
        I[p] = (TS)(zr * ai + zi * ar);
        A.At[p] = A.prof[2 * A.nz + k];
        X[p] = (TS)0;
        Y[p] = (TS)A.prof[3 * A.nz + k];
        A.rho[p] = A.rho0;
    }
}

// interp1(X, Y, Xi, 'spline'): not-a-knot cubic spline through (X, Y), X strictly monotone.
// Slopes from the tridiagonal system of the C2 conditions with not-a-knot end rows, then
// piecewise cubic Hermite evaluation.
int spline_notaknot(int n, const double* Xin, const double* Yin, int m, const double* Xi, double* out) {
    if (n < 4) return -1;
    std::vector<double> x(Xin, Xin + n), y(Yin, Yin + n);
    if (x[0] > x[n - 1]) { std::reverse(x.begin(), x.end()); std::reverse(y.begin(), y.end()); }
    std::vector<double> dx(n - 1), sl(n - 1), lo(n, 0.0), di(n, 0.0), up(n, 0.0), b(n, 0.0);
    for (int i = 0; i < n - 1; ++i) {
        dx[i] = x[i + 1] - x[i];
        if (!(dx[i] > 0)) return -1;
        sl[i] = (y[i + 1] - y[i]) / dx[i];
    }
    for (int i = 1; i < n - 1; ++i) {
        lo[i] = dx[i];
        di[i] = 2 * (dx[i - 1] + dx[i]);
        up[i] = dx[i - 1];
        b[i] = 3 * (dx[i] * sl[i - 1] + dx[i - 1] * sl[i]);
    }
    const double d0 = x[2] - x[0], d1 = x[n - 1] - x[n - 3];
    di[0] = dx[1];
    up[0] = d0;
    b[0] = ((dx[0] + 2 * d0) * dx[1] * sl[0] + dx[0] * dx[0] * sl[1]) / d0;
    di[n - 1] = dx[n - 3];
    lo[n - 1] = d1;
    b[n - 1] = (dx[n - 2] * dx[n - 2] * sl[n - 3] + (2 * d1 + dx[n - 2]) * dx[n - 3] * sl[n - 2]) / d1;
    // Thomas elimination (the not-a-knot rows keep the system tridiagonal)
    for (int i = 1; i < n; ++i) {
        const double w = lo[i] / di[i - 1];
        di[i] -= w * up[i - 1];
        b[i] -= w * b[i - 1];
    }
    std::vector<double> s(n);
    s[n - 1] = b[n - 1] / di[n - 1];
    for (int i = n - 2; i >= 0; --i) s[i] = (b[i] - up[i] * s[i + 1]) / di[i];
    for (int q = 0; q < m; ++q) {
        const double t = Xi[q];
        int k = (int)(std::upper_bound(x.begin(), x.end(), t) - x.begin()) - 1;
        k = std::max(0, std::min(n - 2, k));
        const double h = dx[k], u = t - x[k];
        const double c3 = (s[k] + s[k + 1] - 2 * sl[k]) / (h * h), c2 = (3 * sl[k] - 2 * s[k] - s[k + 1]) / h;
        out[q] = ((c3 * u + c2) * u + s[k]) * u + y[k];
    }
    return 0;
}

}  // namespace hqt

extern "C" {

int hqt_spline(int32_t n, const double* X, const double* Y, int32_t m, const double* Xi, double* out) {
    if (!X || !Y || !Xi || !out || m < 0) return hqt_fail(HQT_ERR_ARG, "hqt_spline: NULL argument");
    if (hqt::spline_notaknot(n, X, Y, m, Xi, out)) return hqt_fail(HQT_ERR_ARG, "hqt_spline: need n >= 4 distinct monotone nodes");
    return HQT_OK;
}

int hqt_init_state(hqt_ctx* c, const hqt_init* in) {
    using namespace hqt;
    if (!c || !in || !in->Z1 || !in->Z2 || !in->Vy1 || !in->rhoi) return hqt_fail(HQT_ERR_ARG, "hqt_init_state: NULL argument");
    if (in->nz0 < 4) return hqt_fail(HQT_ERR_ARG, "hqt_init_state: profile needs nz0 >= 4");
    if (in->nv < 0 || (in->nv > 0 && (!in->vx || !in->vy || !in->q))) return hqt_fail(HQT_ERR_ARG, "hqt_init_state: bad vortex list");
    if (int rc = hqt_sync_streams(c)) return rc;
    const int nz = c->nz;
    // radial grids: profile nodes zz0 = (x0-1)/2+1 on chebdif(nz0), run nodes zz likewise (:44-45)
    std::vector<double> x0, D0, xr, Dr;
    chebdif(in->nz0, 1, x0, D0);
    chebdif(nz, 1, xr, Dr);
    std::vector<double> Zg(in->nz0), zz(nz);
    for (int k = 0; k < in->nz0; ++k) Zg[k] = (x0[k] - 1) * 1.0 / 2 + 1.0;
    for (int k = 0; k < nz; ++k) zz[k] = (xr[k] - 1) * 1.0 / 2 + 1.0;
    std::vector<double> Z1(nz), Z2(nz), V(nz), rh(nz);
    if (spline_notaknot(in->nz0, Zg.data(), in->Z1, nz, zz.data(), Z1.data()) ||
        spline_notaknot(in->nz0, Zg.data(), in->Z2, nz, zz.data(), Z2.data()) ||
        spline_notaknot(in->nz0, Zg.data(), in->Vy1, nz, zz.data(), V.data()) ||
        spline_notaknot(in->nz0, Zg.data(), in->rhoi, nz, zz.data(), rh.data()))
        return hqt_fail(HQT_ERR_ARG, "hqt_init_state: spline failed");
    // Z2_f = Z2 ./ f below the horizon, -Lz(1,:) Z2 / 3 at it; theta = -Lzr(Z2_f) (:90-93)
    std::vector<double> Lz((size_t)nz * nz);
    for (int i = 0; i < nz; ++i)
        for (int k = 0; k < nz; ++k) Lz[(size_t)i * nz + k] = 2.0 * Dr[i + (size_t)nz * k];
    std::vector<double> Zf(nz);
    double p0 = 0;
    for (int k = 0; k < nz; ++k) p0 += Lz[k] * Z2[k];
    Zf[0] = -p0 / 3;
    for (int k = 1; k < nz; ++k) Zf[k] = Z2[k] / (1 - (zz[k] / 1.0) * (zz[k] / 1.0) * (zz[k] / 1.0));
    const int m1 = nz - 1;
    std::vector<double> A((size_t)m1 * m1), th(nz, 0.0);
    for (int i = 0; i < m1; ++i)
        for (int k = 0; k < m1; ++k) A[(size_t)i * m1 + k] = Lz[(size_t)i * nz + k];
    std::vector<double> w(Zf.begin(), Zf.begin() + m1);
    if (!lu_solve(m1, A, w.data(), 1)) return hqt_fail(HQT_ERR_ARG, "hqt_init_state: Lz(1:nz-1,1:nz-1) singular");
    for (int k = 0; k < m1; ++k) th[k] = -w[k];
    std::vector<double> prof(4 * (size_t)nz + 3 * (size_t)std::max(in->nv, 1));
    for (int k = 0; k < nz; ++k) {
        prof[k] = Z1[k] * std::cos(th[k]);
        prof[nz + k] = Z1[k] * std::sin(th[k]);
        prof[2 * nz + k] = Z2[k];
        prof[3 * nz + k] = -V[k];
    }
    for (int v = 0; v < in->nv; ++v) {
        prof[4 * nz + 3 * v] = in->vx[v];
        prof[4 * nz + 3 * v + 1] = in->vy[v];
        prof[4 * nz + 3 * v + 2] = in->q[v];
    }
    double* dbuf = nullptr;
    HIPC(hipMalloc(&dbuf, prof.size() * 8));
    hipError_t e = hipMemcpyAsync(dbuf, prof.data(), prof.size() * 8, hipMemcpyHostToDevice, c->stream);
    ICDev D{};
    D.nx = c->nx; D.ny = c->ny; D.nz = nz; D.nv = in->nv; D.nxy = c->nxy; D.ps = c->ps; D.rp = c->rp;
    D.Lx = c->p.Lx; D.Ly = c->p.Ly;
    D.amp_noise = in->amp_noise; D.phase_noise = in->phase_noise;
    D.inv_sigma2 = in->sigma > 0 ? 1.0 / (in->sigma * in->sigma) : 0.0;
    D.key0 = (uint32_t)in->seed; D.key1 = (uint32_t)(in->seed >> 32);
    D.prof = dbuf; D.vort = dbuf + 4 * nz; D.rho0 = rh[nz - 1];
    D.R = c->ZR; D.I = c->ZI; D.Ax = c->ZAx; D.Ay = c->ZAy; D.At = c->At; D.rho = c->rho;
    const dim3 grid((unsigned)((c->nxy + 255) / 256));
    if (e == hipSuccess) {
        if (c->fp32) hipLaunchKernelGGL(k_init_state<float>, grid, dim3(256), 0, c->stream, D);
        else hipLaunchKernelGGL(k_init_state<double>, grid, dim3(256), 0, c->stream, D);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemsetAsync(c->flags, 0, 4 * sizeof(int32_t), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dbuf);
    if (e != hipSuccess) return hqt_fail(HQT_ERR_HIP, std::string("hqt_init_state: ") + hipGetErrorString(e));
    c->have_state = true;
    return HQT_OK;
}

}  // extern "C"
