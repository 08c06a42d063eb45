// Host side of the batched Newton solver (src/solvers/NewtonRaphson_vy.m on the GPU): radial
// operators, one job descriptor per problem, a reusable device workspace, and the C ABI entry
// hqt_newton_vy_batch.
#include "hqt_internal.h"
#include "hqt_operators.h"

#include <cstring>

NewtonWork::~NewtonWork() {
    if (buf) (void)hipFree(buf);
}

namespace {
size_t align256(size_t n) { return (n + 255) & ~size_t(255); }
}  // namespace

int newton_batch_run(hipStream_t s, NewtonWork& w, int nz, double zh, double small, double D, int nb,
                     const double* Z1, const double* Z2, const double* vy1, const double* vy, int it_max, double tol,
                     double* z1, double* z2, double* VY1, int32_t* it, double* err, uint8_t* conv) {
    // radial grid and operators (NewtonRaphson_vy.m:12-21), as in the host solver
    std::vector<double> zc, DM;
    chebdif(nz, 1, zc, DM);
    const double zf = 1 - small, dlz = zf - D, pz = 2 / dlz;
    const size_t nn = (size_t)nz * nz, n3 = 3 * (size_t)nz, vec = (size_t)nz * nb;
    // host image: [Lz | Lz^2 | zz | Z1 Z2 vy1 (inputs)]
    std::vector<double> host(2 * nn + nz + 3 * vec);
    double* Lz = host.data();
    double* Lz2 = Lz + nn;
    double* zz = Lz2 + nn;
    for (int i = 0; i < nz; ++i) zz[i] = (zc[i] - 1) * dlz / 2 + zf;
    for (int i = 0; i < nz; ++i)
        for (int j = 0; j < nz; ++j) Lz[(size_t)i * nz + j] = pz * DM[i + (size_t)nz * j];
    for (int i = 0; i < nz; ++i)
        for (int j = 0; j < nz; ++j) {
            double acc = 0;
            for (int k = 0; k < nz; ++k) acc += Lz[(size_t)i * nz + k] * Lz[(size_t)k * nz + j];
            Lz2[(size_t)i * nz + j] = acc;
        }
    double* hin = zz + nz;
    std::memcpy(hin, Z1, vec * 8);
    std::memcpy(hin + vec, Z2, vec * 8);
    std::memcpy(hin + 2 * vec, vy1, vec * 8);
    // device layout: host image | results 3*vec | out 3*nb | jobs | per-problem Jacobians (when not in LDS)
    const bool lds = newton_lds_resident(nz);
    const size_t o_img = 0, o_res = align256(host.size() * 8), o_out = o_res + align256(3 * vec * 8),
                 o_jobs = o_out + align256(3 * (size_t)nb * 8), o_work = o_jobs + align256((size_t)nb * sizeof(NewtonJob)),
                 total = o_work + (lds ? 0 : (size_t)nb * n3 * n3 * 8);
    if (total > w.cap) {
        if (w.buf) {
            HIPC(hipStreamSynchronize(s));
            (void)hipFree(w.buf);
            w.buf = nullptr;
            w.cap = 0;
        }
        const hipError_t e = hipMalloc(&w.buf, total);
        if (e != hipSuccess) { w.buf = nullptr; return hqt_fail(HQT_ERR_OOM, "hipMalloc(Newton workspace)"); }
        w.cap = total;
    }
    double* dimg = reinterpret_cast<double*>(w.buf + o_img);
    double* dres = reinterpret_cast<double*>(w.buf + o_res);
    double* dout = reinterpret_cast<double*>(w.buf + o_out);
    NewtonJob* djobs = reinterpret_cast<NewtonJob*>(w.buf + o_jobs);
    double* dwork = reinterpret_cast<double*>(w.buf + o_work);
    const double* din = dimg + 2 * nn + nz;
    std::vector<NewtonJob> jobs(nb);
    for (int b = 0; b < nb; ++b) {
        NewtonJob& J = jobs[b];
        J.nz = nz;
        J.it_max = it_max;
        J.zh = zh;
        J.tol = tol;
        J.vy = vy[b];
        J.Lz = dimg;
        J.Lz2 = dimg + nn;
        J.zz = dimg + 2 * nn;
        J.Z1 = din + (size_t)b * nz;
        J.Z2 = din + vec + (size_t)b * nz;
        J.vy1 = din + 2 * vec + (size_t)b * nz;
        J.z1 = dres + (size_t)b * nz;
        J.z2 = dres + vec + (size_t)b * nz;
        J.VY1 = dres + 2 * vec + (size_t)b * nz;
        J.work = lds ? nullptr : dwork + (size_t)b * n3 * n3;
        J.out = dout + 3 * (size_t)b;
    }
    std::vector<double> res(3 * vec + 3 * (size_t)nb);
    HIPC(hipMemcpyAsync(dimg, host.data(), host.size() * 8, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(djobs, jobs.data(), jobs.size() * sizeof(NewtonJob), hipMemcpyHostToDevice, s));
    HIPC(launch_newton(djobs, nb, nz, s));
    HIPC(hipMemcpyAsync(res.data(), dres, 3 * vec * 8, hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(res.data() + 3 * vec, dout, 3 * (size_t)nb * 8, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    std::memcpy(z1, res.data(), vec * 8);
    std::memcpy(z2, res.data() + vec, vec * 8);
    std::memcpy(VY1, res.data() + 2 * vec, vec * 8);
    const double* o = res.data() + 3 * vec;
    for (int b = 0; b < nb; ++b) {
        it[b] = (int32_t)o[3 * b];
        err[b] = o[3 * b + 1];
        conv[b] = o[3 * b + 2] != 0 ? 1 : 0;
    }
    return HQT_OK;
}

extern "C" int hqt_newton_vy_batch(int32_t nz, double zh, double small, double D, int32_t nb, const double* Z1,
                                   const double* Z2, const double* vy1, const double* vy, int32_t it_max, double tol,
                                   double* z1, double* z2, double* VY1, int32_t* it, double* err, uint8_t* conv) {
    if (nz < 3 || nz > 64) return hqt_fail(HQT_ERR_ARG, "hqt_newton_vy_batch: nz must lie in [3, 64]");
    if (nb < 1) return hqt_fail(HQT_ERR_ARG, "hqt_newton_vy_batch: nb must be positive");
    if (!Z1 || !Z2 || !vy1 || !vy || !z1 || !z2 || !VY1 || !it || !err || !conv)
        return hqt_fail(HQT_ERR_ARG, "hqt_newton_vy_batch: NULL argument");
    if (it_max <= 0) it_max = 1000;
    if (!(tol > 0)) tol = 1e-6;
    hipStream_t s = nullptr;
    HIPC(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int rc;
    {
        NewtonWork w;
        rc = newton_batch_run(s, w, nz, zh, small, D, nb, Z1, Z2, vy1, vy, it_max, tol, z1, z2, VY1, it, err, conv);
    }
    (void)hipStreamDestroy(s);
    if (rc) return rc;
    int nconv = 0;
    for (int b = 0; b < nb; ++b) nconv += conv[b];
    if (nconv != nb)
        return hqt_fail(HQT_ERR_NOCONV, "hqt_newton_vy_batch: " + std::to_string(nb - nconv) + " of " + std::to_string(nb) +
                                            " problems did not converge");
    return HQT_OK;
}
