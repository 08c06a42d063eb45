// Batched stationary Newton solver on gfx950: NewtonRaphson_vy (src/solvers/NewtonRaphson_vy.m)
// for many independent (vy, initial guess) problems of one radial size nz, one work-group each.
//
// A work-group follows the reference iteration step for step:
//   - residual `fun` (:32-34) with the Dirichlet rows zeroed (:62,82);
//   - analytic Jacobian `dfun` (:37-39) with the rows replaced by BCz (:50-52,63,66);
//   - y = -dFun\Fun (:69), dense LU with partial pivoting and column-oriented triangular solves
//     (the LAPACK getrf/getrs order);
//   - the update and the new residual (:73-83);
//   - the stop tests: NaN update norm (:88-90), max|Fun_next| < tol (:93).
// A problem with a continuation seed walks mu from the seed to its target inside the same
// work-group, one solve per mu step, as the reference's stationary-state notebook does by hand
// (scripts/legacy/Statinary_state_with_v.m:11-17,50-52).
// The 3nz x 3nz Jacobian lives in LDS for nz <= 41 (the reference's nz = 30 takes 65 KB) and
// in a per-problem global workspace above that.  The radial operators Lz, Lz^2 and all
// vectors are always in LDS.
#include "hqt_internal.h"
#include <math.h>

#include <mutex>

namespace hqt {

constexpr int NW_T = 512;          // threads per problem
constexpr int NW_W = NW_T / 64;    // waves per problem
constexpr int NW_MAXNZ = 64;
constexpr size_t NW_LDS_LIMIT = 160 * 1024;
constexpr int NW_CONT_IT = 100;          // iteration limit of one continuation solve (it_max if smaller)
constexpr int NW_CONT_MAX_STEPS = 1024;  // solves per continuation, the seed's included

struct NwVec {
    double Z1[NW_MAXNZ], Z2[NW_MAXNZ], V[NW_MAXNZ], Z2f[NW_MAXNZ];
    double T1[NW_MAXNZ], T2[NW_MAXNZ], TV[NW_MAXNZ], T2f[NW_MAXNZ];
    double zz[NW_MAXNZ], f[NW_MAXNZ], df[NW_MAXNZ];
    double C1[NW_MAXNZ], C2[NW_MAXNZ], CV[NW_MAXNZ];   // last converged continuation point
    double F[3 * NW_MAXNZ], Fn[3 * NW_MAXNZ], y[3 * NW_MAXNZ], x[3 * NW_MAXNZ];
    double pv[NW_W];
    int pi[NW_W];
    double ra[NW_W], rb[NW_W];
};

static size_t newton_dyn_bytes(int nz, bool mlds) {
    return sizeof(double) * (size_t)nz * nz * (mlds ? 11 : 2);
}

bool newton_lds_resident(int nz) {
    return newton_dyn_bytes(nz, true) + sizeof(NwVec) <= NW_LDS_LIMIT;
}

// max|v| over the block, skipping NaN like MATLAB's max (NaN only when every entry is NaN)
__device__ __forceinline__ double block_max_skipnan(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
    for (int t = 1; t < NW_W; ++t) r = fmax(r, red[t]);
    return r;
}

// max|v| over the block with NaN propagation, like MATLAB's norm(y, inf)
__device__ __forceinline__ double nanmax2(double a, double b) {
    return (a != a || b != b) ? __builtin_nan("") : fmax(a, b);
}
__device__ __forceinline__ double block_max_nanprop(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v = nanmax2(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
    for (int t = 1; t < NW_W; ++t) r = nanmax2(r, red[t]);
    return r;
}

// fun(Z1,Z2,Vy1,Z2_f) (:32-34) into out[3nz], one thread per (equation, node), Dirichlet rows
// zeroed.  Caller syncs before and after.
__device__ void nw_residual(int nz, const double* Lz, const double* L2, const NwVec& s, const double* a,
                            const double* b, const double* v, const double* bf, double* out) {
    const int r = threadIdx.x;
    if (r >= 3 * nz) return;
    const int eq = (r >= nz) + (r >= 2 * nz), i = r - eq * nz;
    const double* l2 = L2 + i * nz;
    const double* l1 = Lz + i * nz;
    const double ai = a[i];
    double o;
    if (eq == 0) {
        double a2 = 0, a1 = 0;
        for (int j = 0; j < nz; ++j) { a2 += l2[j] * a[j]; a1 += l1[j] * a[j]; }
        o = s.f[i] * a2 + s.df[i] * a1 - v[i] * v[i] * ai + b[i] * bf[i] * ai - s.zz[i] * ai;
    } else if (eq == 1) {
        double b2 = 0;
        for (int j = 0; j < nz; ++j) b2 += l2[j] * b[j];
        o = s.f[i] * b2 - 2 * ai * ai * b[i];
    } else {
        double v2 = 0, v1 = 0;
        for (int j = 0; j < nz; ++j) { v2 += l2[j] * v[j]; v1 += l1[j] * v[j]; }
        o = -s.f[i] * v2 - s.df[i] * v1 + 2 * v[i] * ai * ai;
    }
    const bool bc = (r == nz - 1) || (r == nz) || (r == 2 * nz - 1) || (r == 3 * nz - 1);
    out[r] = bc ? 0.0 : o;
}

// Z2_f = Z2./f below the horizon, -Lz(1,:)*Z2/3 at the horizon node (:26-27, :76-78)
__device__ void nw_z2f(int nz, const double* Lz, const NwVec& s, const double* b, double* bf) {
    for (int i = threadIdx.x; i < nz; i += NW_T) {
        if (i == 0) {
            double p = 0;
            for (int j = 0; j < nz; ++j) p += Lz[j] * b[j];
            bf[0] = -p / 3;
        } else {
            bf[i] = b[i] / s.f[i];
        }
    }
}

// One NewtonRaphson_vy solve from the iterate in s.Z1, s.Z2, s.V (V = Vy1, already shifted as :28-29):
// Z2_f and the first residual (:26-27, :62), then the iteration (:60-96).  On return s.Z1, s.Z2, s.V
// hold the last iterate and K, err the reference's `it` and `error`; the result is 1 when
// max|Fun_next| < tol.  Every branch is taken on block-uniform values, so all threads return together.
__device__ int nw_solve(int nz, const double* Lz, const double* L2, double* M, NwVec& s, int it_max, double tol,
                        int& K_out, double& err_out) {
    const int n = 3 * nz;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    nw_z2f(nz, Lz, s, s.Z2, s.Z2f);
    __syncthreads();
    nw_residual(nz, Lz, L2, s, s.Z1, s.Z2, s.V, s.Z2f, s.F);
    __syncthreads();

    int K = 0, converged = 0;
    double err = __builtin_nan("");
    while (K < it_max) {
        // ---- dFun (:37-39) with BCz rows; y <- Fun ----
        for (int r = w; r < n; r += NW_W) {
            const int br = (r >= nz) + (r >= 2 * nz), i = r - br * nz;
            const bool bc = (r == nz - 1) || (r == nz) || (r == 2 * nz - 1) || (r == 3 * nz - 1);
            for (int q = lane; q < n; q += 64) {
                const int bq = (q >= nz) + (q >= 2 * nz), j = q - bq * nz;
                double v = 0.0;
                if (bc) {
                    v = (q == r) ? 1.0 : 0.0;
                } else if (br == bq) {
                    const double l2 = L2[i * nz + j], l1 = Lz[i * nz + j];
                    if (br == 0) v = s.f[i] * l2 + s.df[i] * l1;
                    else if (br == 1) v = s.f[i] * l2;
                    else v = -s.f[i] * l2 - s.df[i] * l1;
                    if (i == j) {
                        if (br == 0) v += s.Z2[i] * s.Z2f[i] - s.V[i] * s.V[i] - s.zz[i];
                        else if (br == 1) v -= 2 * s.Z1[i] * s.Z1[i];
                        else v += 2 * s.Z1[i] * s.Z1[i];
                    }
                } else if (i == j) {
                    if (br == 0) v = (bq == 1) ? 2 * s.Z2f[i] * s.Z1[i] : -2 * s.V[i] * s.Z1[i];
                    else if (br == 1) v = (bq == 0) ? -4 * s.Z1[i] * s.Z2[i] : 0.0;
                    else v = (bq == 0) ? 4 * s.V[i] * s.Z1[i] : 0.0;
                }
                M[r * n + q] = v;
            }
        }
        for (int i = tid; i < n; i += NW_T) s.y[i] = s.F[i];
        __syncthreads();

        // ---- LU with partial pivoting; forward elimination applied to y on the fly ----
        bool singular = false;
        for (int k = 0; k < n; ++k) {
            double bv = -1.0;
            int bi = n;
            for (int i = k + tid; i < n; i += NW_T) {
                const double a = fabs(M[i * n + k]);
                if (a > bv) { bv = a; bi = i; }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(bv, off);
                const int oi = __shfl_xor(bi, off);
                if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
            }
            if (lane == 0) { s.pv[w] = bv; s.pi[w] = bi; }
            __syncthreads();
            bv = s.pv[0];
            bi = s.pi[0];
            for (int t = 1; t < NW_W; ++t) {
                const double ov = s.pv[t];
                const int oi = s.pi[t];
                if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
            }
            if (!(bv > 0.0)) { singular = true; break; }   // same decision in every thread
            if (bi != k) {
                for (int q = k + tid; q < n; q += NW_T) {
                    const double t = M[k * n + q];
                    M[k * n + q] = M[bi * n + q];
                    M[bi * n + q] = t;
                }
                if (tid == 0) { const double t = s.y[k]; s.y[k] = s.y[bi]; s.y[bi] = t; }
                __syncthreads();
            }
            // rank-1 update of the trailing block.  Each wave takes rows i = k+1+w (mod NW_W), four
            // rows per batch so that their LDS loads overlap; row k (the pivot row) is held in
            // registers.  Multipliers use the reciprocal of the pivot, as LAPACK's getf2 does.
            const double rdk = 1.0 / M[k * n + k];
            const double yk = s.y[k];
            const int q0 = k + 1 + lane;
            double rk[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) rk[c] = (q0 + 64 * c < n) ? M[k * n + q0 + 64 * c] : 0.0;
            int i = k + 1 + w;
            for (; i + 3 * NW_W < n; i += 4 * NW_W) {
                double l[4], m[4][3];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const double* row = M + (i + u * NW_W) * n;
                    l[u] = row[k];
#pragma unroll
                    for (int c = 0; c < 3; ++c) m[u][c] = (q0 + 64 * c < n) ? row[q0 + 64 * c] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    double* row = M + (i + u * NW_W) * n;
                    const double lu = l[u] * rdk;
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        if (q0 + 64 * c < n) row[q0 + 64 * c] = m[u][c] - lu * rk[c];
                    if (lane == 0) s.y[i + u * NW_W] -= lu * yk;
                }
            }
            for (; i < n; i += NW_W) {
                double* row = M + i * n;
                const double lu = row[k] * rdk;
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    if (q0 + 64 * c < n) row[q0 + 64 * c] -= lu * rk[c];
                if (lane == 0) s.y[i] -= lu * yk;
            }
            __syncthreads();
        }
        if (singular) { err = __builtin_nan(""); break; }
        // back substitution, column-oriented; x = -U \ y is the Newton update
        for (int k = n - 1; k >= 0; --k) {
            const double xk = s.y[k] / M[k * n + k];
            for (int i = tid; i < k; i += NW_T) s.y[i] -= M[i * n + k] * xk;
            if (tid == 0) s.x[k] = -xk;
            __syncthreads();
        }

        // ---- update, new residual, stop tests (:73-96) ----
        for (int i = tid; i < nz; i += NW_T) {
            s.T1[i] = s.Z1[i] + s.x[i];
            s.T2[i] = s.Z2[i] + s.x[nz + i];
            s.TV[i] = s.V[i] + s.x[2 * nz + i];
        }
        __syncthreads();
        nw_z2f(nz, Lz, s, s.T2, s.T2f);
        __syncthreads();
        nw_residual(nz, Lz, L2, s, s.T1, s.T2, s.TV, s.T2f, s.Fn);
        __syncthreads();
        double la = __builtin_nan(""), lb = 0.0;
        for (int i = tid; i < n; i += NW_T) {
            la = fmax(la, fabs(s.Fn[i]));
            lb = nanmax2(lb, fabs(s.x[i]));
        }
        const double err2 = block_max_skipnan(la, s.ra);
        err = block_max_nanprop(lb, s.rb);
        for (int i = tid; i < nz; i += NW_T) {
            s.Z1[i] = s.T1[i];
            s.Z2[i] = s.T2[i];
            s.V[i] = s.TV[i];
            s.Z2f[i] = s.T2f[i];
        }
        for (int i = tid; i < n; i += NW_T) s.F[i] = s.Fn[i];
        __syncthreads();
        if (err != err) break;
        if (err2 < tol) { converged = 1; break; }
        ++K;
    }
    K_out = K;
    err_out = err;
    return converged;
}

// Restart point of a continuation solve: Z1, Z2 rescaled by r with A_t(z=0) set to mu
// (Statinary_state_with_v.m:50), Vy1 = vy + vy1 - vy1(end) (NewtonRaphson_vy.m:28-29).
__device__ void nw_restart(int nz, NwVec& s, const double* a, const double* b, const double* v, double vy, double r,
                           double mu) {
    for (int i = threadIdx.x; i < nz; i += NW_T) {
        s.Z1[i] = a[i];
        s.Z2[i] = (i == nz - 1) ? mu : b[i] * r;
        s.V[i] = vy + (v[i] - v[nz - 1]);
    }
    __syncthreads();
}

template <bool MLDS>
__global__ __launch_bounds__(NW_T) void k_newton(const NewtonJob* __restrict__ jobs) {
    extern __shared__ double dyn[];
    __shared__ NwVec s;
    const NewtonJob J = jobs[blockIdx.x];
    const int nz = J.nz;
    const int tid = threadIdx.x;
    double* Lz = dyn;
    double* L2 = dyn + nz * nz;
    double* M = MLDS ? dyn + 2 * nz * nz : J.work;   // row-major 3nz x 3nz

    for (int e = tid; e < nz * nz; e += NW_T) { Lz[e] = J.Lz[e]; L2[e] = J.Lz2[e]; }
    const double zh3 = J.zh * J.zh * J.zh;
    for (int i = tid; i < nz; i += NW_T) {
        const double zi = J.zz[i], q = zi / J.zh;
        s.zz[i] = zi;
        s.f[i] = 1 - q * q * q;
        s.df[i] = -3 * zi * zi / zh3;
    }
    int K = 0, conv = 0, steps = 0;
    double err = __builtin_nan("");
    if (J.mu_seed != J.mu_seed) {
        // one solve from the caller's guess (NewtonRaphson_vy.m:1-105)
        for (int i = tid; i < nz; i += NW_T) {
            s.Z1[i] = J.Z1[i];
            s.Z2[i] = J.Z2[i];
            s.V[i] = J.vy + (J.vy1[i] - J.vy1[nz - 1]);   // :28-29
        }
        __syncthreads();
        conv = nw_solve(nz, Lz, L2, M, s, J.it_max, J.tol, K, err);
    } else {
        // mu-continuation: the guess rescaled to the seed's A_t(z=0) = mu_seed and solved there, then mu
        // stepped to the guess's own A_t(z=0) in steps of at most dmu, each from the last solution
        // rescaled (Statinary_state_with_v.m:11-17,50-52).  A step that does not converge is retried at
        // half the length, down to dmu/64.
        const double mt = J.Z2[nz - 1];
        const int itc = J.it_max < NW_CONT_IT ? J.it_max : NW_CONT_IT;
        double prev = J.mu_seed;
        nw_restart(nz, s, J.Z1, J.Z2, J.vy1, J.vy, J.mu_seed / mt, J.mu_seed);
        conv = nw_solve(nz, Lz, L2, M, s, itc, J.tol, K, err);
        steps = 1;
        if (conv) {
            double h = copysign(J.dmu, mt - prev);
            const double hmin = J.dmu * (1.0 / 64);
            while (prev != mt && steps < NW_CONT_MAX_STEPS) {
                for (int i = tid; i < nz; i += NW_T) { s.C1[i] = s.Z1[i]; s.C2[i] = s.Z2[i]; s.CV[i] = s.V[i]; }
                __syncthreads();
                bool ok = false;
                while (!ok && steps < NW_CONT_MAX_STEPS) {
                    double m = prev + h;
                    if (h > 0 ? m >= mt : m <= mt) m = mt;
                    nw_restart(nz, s, s.C1, s.C2, s.CV, J.vy, m / prev, m);
                    ok = nw_solve(nz, Lz, L2, M, s, itc, J.tol, K, err) != 0;
                    ++steps;
                    if (ok) {
                        prev = m;
                    } else {
                        h *= 0.5;
                        if (fabs(h) < hmin) break;
                    }
                }
                if (!ok) break;
            }
            conv = prev == mt;
        }
    }
    for (int i = tid; i < nz; i += NW_T) {
        J.z1[i] = s.Z1[i];
        J.z2[i] = s.Z2[i];
        J.VY1[i] = s.V[i];
    }
    if (tid == 0) {
        J.out[0] = (double)K;
        J.out[1] = err;
        J.out[2] = conv ? 1.0 : 0.0;
        J.out[3] = (double)steps;
    }
}

hipError_t launch_newton(const NewtonJob* jobs, int njobs, int maxnz, bool lds, hipStream_t st) {
    if (maxnz < 2 || maxnz > NW_MAXNZ || njobs <= 0) return hipErrorInvalidValue;
    if (lds && !newton_lds_resident(maxnz)) return hipErrorInvalidValue;
    const size_t dyn = newton_dyn_bytes(maxnz, lds);
    hipError_t e;
    if (lds) {
        e = hipFuncSetAttribute((const void*)k_newton<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_newton<true>, dim3(njobs), dim3(NW_T), dyn, st, jobs);
    } else {
        e = hipFuncSetAttribute((const void*)k_newton<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_newton<false>, dim3(njobs), dim3(NW_T), dyn, st, jobs);
    }
    return hipGetLastError();
}

}  // namespace hqt

// ---- allocation-free batched Newton: hqt_newton_vy_batch_noalloc --------------------------------------
// The device workspace of hqt_newton_vy_batch (operators, inputs, results, job descriptors and, for
// nz > 41, the per-problem Jacobians) placed in a fixed array instead of a hipMalloc'd buffer.  Each GPU
// holds its own copy of the array for the life of the process.  Its host-side name is not a device
// address: the address is looked up once per device with hipGetSymbolAddress and kept in g_scratch.
static __device__ __align__(256) unsigned char scratch[HQT_NEWTON_NOALLOC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu[kMaxDevices];     // one call at a time per device: the calls share its array
void* g_scratch[kMaxDevices] = {};        // device address of `scratch` on each device (0 until set up)
}  // namespace

extern "C" int hqt_newton_vy_batch_noalloc(int32_t nz, double zh, double small, double D, int32_t nb,
                                           const double* Z1, const double* Z2, const double* vy1, const double* vy,
                                           int32_t it_max, double tol, double* z1, double* z2, double* VY1,
                                           int32_t* it, double* err, uint8_t* conv) {
    static const char* who = "hqt_newton_vy_batch_noalloc";
    if (int rc = newton_batch_args(who, nz, nb, Z1, Z2, vy1, vy, z1, z2, VY1, it, err, conv, it_max, tol)) return rc;
    const size_t need = newton_batch_bytes(nz, nb);
    if (need > sizeof(scratch))
        return hqt_fail(HQT_ERR_ARG, std::string(who) + ": " + std::to_string(nb) + " problems of nz = " +
                                         std::to_string(nz) + " need " + std::to_string(need) +
                                         " bytes of device workspace, more than the " + std::to_string(sizeof(scratch)) +
                                         "-byte allocation-free workspace (use hqt_newton_vy_batch or smaller batches)");
    int dev = -1;
    HIPC(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return hqt_fail(HQT_ERR_ARG, std::string(who) + ": device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu[dev]);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return hqt_fail(HQT_ERR_HIP, std::string(who) + ": hipGetSymbolAddress(scratch): " + hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    // the null stream: no stream is created; the library's own streams are non-blocking and do not wait on it
    const int rc = newton_batch_run_in(nullptr, static_cast<unsigned char*>(g_scratch[dev]), nz, zh, small, D, nb, Z1,
                                       Z2, vy1, vy, it_max, tol, z1, z2, VY1, it, err, conv);
    if (rc) return rc;
    return newton_batch_status(who, nb, conv);
}

