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

// Byte offsets of the device layout: host image (operators, inputs) | results 3*vec | out 3*nb | jobs |
// per-problem Jacobians (when they do not fit in LDS).
struct NewtonLayout { size_t img, res, out, jobs, work, total; };

NewtonLayout newton_layout(int nz, int nb) {
    const size_t nn = (size_t)nz * nz, n3 = 3 * (size_t)nz, vec = (size_t)nz * nb;
    NewtonLayout L;
    L.img = 0;
    L.res = align256((2 * nn + nz + 3 * vec) * 8);
    L.out = L.res + align256(3 * vec * 8);
    L.jobs = L.out + align256(3 * (size_t)nb * 8);
    L.work = L.jobs + align256((size_t)nb * sizeof(NewtonJob));
    L.total = L.work + (newton_lds_resident(nz) ? 0 : (size_t)nb * n3 * n3 * 8);
    return L;
}
}  // namespace

size_t newton_batch_bytes(int nz, int nb) { return newton_layout(nz, nb).total; }

int newton_batch_run(hipStream_t s, NewtonWork& w, int nz, double zh, double small, double D, int nb,
                     const double* Z1, const double* Z2, const double* vy1, const double* vy, int it_max, double tol,
                     double* z1, double* z2, double* VY1, int32_t* it, double* err, uint8_t* conv) {
    const size_t total = newton_batch_bytes(nz, nb);
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
    return newton_batch_run_in(s, w.buf, nz, zh, small, D, nb, Z1, Z2, vy1, vy, it_max, tol, z1, z2, VY1, it, err,
                               conv);
}

int newton_batch_run_in(hipStream_t s, unsigned char* buf, int nz, double zh, double small, double D, int nb,
                        const double* Z1, const double* Z2, const double* vy1, const double* vy, int it_max,
                        double tol, double* z1, double* z2, double* VY1, int32_t* it, double* err, uint8_t* conv) {
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
    const bool lds = newton_lds_resident(nz);
    const NewtonLayout lay = newton_layout(nz, nb);
    const size_t o_img = lay.img, o_res = lay.res, o_out = lay.out, o_jobs = lay.jobs, o_work = lay.work;
    double* dimg = reinterpret_cast<double*>(buf + o_img);
    double* dres = reinterpret_cast<double*>(buf + o_res);
    double* dout = reinterpret_cast<double*>(buf + o_out);
    NewtonJob* djobs = reinterpret_cast<NewtonJob*>(buf + o_jobs);
    double* dwork = reinterpret_cast<double*>(buf + o_work);
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

int newton_batch_args(const char* who, int nz, int nb, const void* Z1, const void* Z2, const void* vy1, const void* vy,
                      const void* z1, const void* z2, const void* VY1, const void* it, const void* err,
                      const void* conv, int32_t& it_max, double& tol) {
    if (nz < 3 || nz > 64) return hqt_fail(HQT_ERR_ARG, std::string(who) + ": nz must lie in [3, 64]");
    if (nb < 1) return hqt_fail(HQT_ERR_ARG, std::string(who) + ": nb must be positive");
    if (!Z1 || !Z2 || !vy1 || !vy || !z1 || !z2 || !VY1 || !it || !err || !conv)
        return hqt_fail(HQT_ERR_ARG, std::string(who) + ": NULL argument");
    if (it_max <= 0) it_max = 1000;
    if (!(tol > 0)) tol = 1e-6;
    return HQT_OK;
}

int newton_batch_status(const char* who, int nb, const uint8_t* conv) {
    int nconv = 0;
    for (int b = 0; b < nb; ++b) nconv += conv[b];
    if (nconv != nb)
        return hqt_fail(HQT_ERR_NOCONV, std::string(who) + ": " + std::to_string(nb - nconv) + " of " +
                                            std::to_string(nb) + " problems did not converge");
    return HQT_OK;
}

extern "C" int hqt_newton_vy_batch(int32_t nz, double zh, double small, double D, int32_t nb, const double* Z1,
                                   const double* Z2, const double* vy1, const double* vy, int32_t it_max, double tol,
                                   double* z1, double* z2, double* VY1, int32_t* it, double* err, uint8_t* conv) {
    static const char* who = "hqt_newton_vy_batch";
    if (int rc = newton_batch_args(who, nz, nb, Z1, Z2, vy1, vy, z1, z2, VY1, it, err, conv, it_max, tol)) return rc;
    hipStream_t s = nullptr;
    HIPC(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int rc;
    {
        NewtonWork w;
        rc = newton_batch_run(s, w, nz, zh, small, D, nb, Z1, Z2, vy1, vy, it_max, tol, z1, z2, VY1, it, err, conv);
    }
    (void)hipStreamDestroy(s);
    if (rc) return rc;
    return newton_batch_status(who, nb, conv);
}
