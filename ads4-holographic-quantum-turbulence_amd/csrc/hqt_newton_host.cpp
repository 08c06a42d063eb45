// Host side of the batched Newton solver (src/solvers/NewtonRaphson_vy.m on the GPU): radial
// operators, one job descriptor per problem, a reusable device workspace, and the C ABI entry
// hqt_newton_vy_batch.
#include "hqt_internal.h"
#include "hqt_operators.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <tuple>

NewtonWork::~NewtonWork() {
    if (buf) (void)hipFree(buf);
}

namespace {
size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// Radial operators of one (nz, small, D) (NewtonRaphson_vy.m:12-21), as in the host solver:
// [Lz | Lz^2 | zz], Lz row-major.
void newton_operators(int nz, double small, double D, double* out) {
    std::vector<double> zc, DM;
    chebdif(nz, 1, zc, DM);
    const double zf = 1 - small, dlz = zf - D, pz = 2 / dlz;
    const size_t nn = (size_t)nz * nz;
    double* Lz = out;
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
}

using OpKey = std::tuple<int, double, double>;   // (nz, small, D)

// Device layout of one launch: host image [operator blocks | inputs] | results | out 4*nb | jobs |
// per-problem Jacobians of the problems whose Jacobian does not fit in LDS.
struct NewtonLayout {
    std::map<OpKey, size_t> ops;              // offset (doubles) of each operator block in the image
    std::vector<size_t> vin, vres, vwork;     // per problem: offsets (doubles) of inputs, results, Jacobian
    size_t img_d = 0, res_d = 0, work_d = 0;  // sizes (doubles)
    size_t res, out, jobs, work, total;       // byte offsets
};

NewtonLayout newton_layout(const NewtonProblems& P) {
    NewtonLayout L;
    for (int b = 0; b < P.nb; ++b) {
        const OpKey k{P.nz[b], P.small[b], P.D[b]};
        if (!L.ops.count(k)) {
            L.ops[k] = L.img_d;
            L.img_d += 2 * (size_t)P.nz[b] * P.nz[b] + P.nz[b];
        }
    }
    L.vin.resize(P.nb);
    L.vres.resize(P.nb);
    L.vwork.resize(P.nb);
    for (int b = 0; b < P.nb; ++b) {
        const size_t nz = P.nz[b], n3 = 3 * nz;
        L.vin[b] = L.img_d;
        L.img_d += n3;
        L.vres[b] = L.res_d;
        L.res_d += n3;
        L.vwork[b] = L.work_d;
        if (!newton_lds_resident((int)nz)) L.work_d += n3 * n3;
    }
    L.res = align256(L.img_d * 8);
    L.out = L.res + align256(L.res_d * 8);
    L.jobs = L.out + align256(4 * (size_t)P.nb * 8);
    L.work = L.jobs + align256((size_t)P.nb * sizeof(NewtonJob));
    L.total = L.work + L.work_d * 8;
    return L;
}

// The uniform batch of the C entry points as a problem list ([nz, nb] column-major arrays).
NewtonProblems uniform_problems(int nz, double zh, double small, double D, int nb, const double* Z1, const double* Z2,
                                const double* vy1, const double* vy, double* z1, double* z2, double* VY1, int32_t* it,
                                double* err, uint8_t* conv, const NewtonCont* cont) {
    NewtonProblems P;
    P.resize(nb);
    for (int b = 0; b < nb; ++b) {
        const size_t o = (size_t)b * nz;
        P.nz[b] = nz;
        P.zh[b] = zh;
        P.small[b] = small;
        P.D[b] = D;
        P.vy[b] = vy[b];
        P.mu_seed[b] = cont ? cont->mu_seed[b] : std::numeric_limits<double>::quiet_NaN();
        P.dmu[b] = cont ? cont->dmu[b] : 0.0;
        P.Z1[b] = Z1 + o;
        P.Z2[b] = Z2 + o;
        P.vy1[b] = vy1 + o;
        P.z1[b] = z1 + o;
        P.z2[b] = z2 + o;
        P.VY1[b] = VY1 + o;
    }
    P.it = it;
    P.err = err;
    P.conv = conv;
    P.steps = cont ? cont->steps : nullptr;
    return P;
}
}  // namespace

void NewtonProblems::resize(int n) {
    nb = n;
    nz.assign(n, 0);
    zh.assign(n, 1.0);
    small.assign(n, 0.0);
    D.assign(n, 0.0);
    vy.assign(n, 0.0);
    mu_seed.assign(n, std::numeric_limits<double>::quiet_NaN());
    dmu.assign(n, 0.0);
    Z1.assign(n, nullptr);
    Z2.assign(n, nullptr);
    vy1.assign(n, nullptr);
    z1.assign(n, nullptr);
    z2.assign(n, nullptr);
    VY1.assign(n, nullptr);
}

size_t newton_problems_bytes(const NewtonProblems& P) { return newton_layout(P).total; }

int newton_problems_run(hipStream_t s, NewtonWork& w, const NewtonProblems& P, int it_max, double tol) {
    const size_t total = newton_problems_bytes(P);
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
    return newton_problems_run_in(s, w.buf, P, it_max, tol);
}

int newton_problems_run_in(hipStream_t s, unsigned char* buf, const NewtonProblems& P, int it_max, double tol) {
    const int nb = P.nb;
    if (nb <= 0) return HQT_OK;
    const NewtonLayout L = newton_layout(P);
    std::vector<double> host(L.img_d);
    for (const auto& kv : L.ops)
        newton_operators(std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), host.data() + kv.second);
    for (int b = 0; b < nb; ++b) {
        const size_t nz = P.nz[b];
        std::memcpy(host.data() + L.vin[b], P.Z1[b], nz * 8);
        std::memcpy(host.data() + L.vin[b] + nz, P.Z2[b], nz * 8);
        std::memcpy(host.data() + L.vin[b] + 2 * nz, P.vy1[b], nz * 8);
    }
    double* dimg = reinterpret_cast<double*>(buf);
    double* dres = reinterpret_cast<double*>(buf + L.res);
    double* dout = reinterpret_cast<double*>(buf + L.out);
    NewtonJob* djobs = reinterpret_cast<NewtonJob*>(buf + L.jobs);
    double* dwork = reinterpret_cast<double*>(buf + L.work);
    // jobs ordered LDS-resident problems first, then those with a global Jacobian: one launch each
    std::vector<int> order;
    int nlds = 0, zlds = 0, zglob = 0;
    for (int pass = 0; pass < 2; ++pass)
        for (int b = 0; b < nb; ++b) {
            const bool lds = newton_lds_resident(P.nz[b]);
            if (lds != (pass == 0)) continue;
            order.push_back(b);
            if (lds) { ++nlds; zlds = std::max(zlds, P.nz[b]); }
            else zglob = std::max(zglob, P.nz[b]);
        }
    std::vector<NewtonJob> jobs(nb);
    for (int j = 0; j < nb; ++j) {
        const int b = order[j];
        const int nz = P.nz[b];
        const size_t nn = (size_t)nz * nz;
        const double* op = dimg + L.ops.at(OpKey{nz, P.small[b], P.D[b]});
        const double* in = dimg + L.vin[b];
        double* rs = dres + L.vres[b];
        NewtonJob& J = jobs[j];
        J.nz = nz;
        J.it_max = it_max;
        J.zh = P.zh[b];
        J.tol = tol;
        J.vy = P.vy[b];
        J.mu_seed = P.mu_seed[b];
        J.dmu = P.dmu[b];
        J.Lz = op;
        J.Lz2 = op + nn;
        J.zz = op + 2 * nn;
        J.Z1 = in;
        J.Z2 = in + nz;
        J.vy1 = in + 2 * nz;
        J.z1 = rs;
        J.z2 = rs + nz;
        J.VY1 = rs + 2 * nz;
        J.work = newton_lds_resident(nz) ? nullptr : dwork + L.vwork[b];
        J.out = dout + 4 * (size_t)b;
    }
    std::vector<double> res(L.res_d + 4 * (size_t)nb);
    HIPC(hipMemcpyAsync(dimg, host.data(), host.size() * 8, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(djobs, jobs.data(), jobs.size() * sizeof(NewtonJob), hipMemcpyHostToDevice, s));
    if (nlds) HIPC(launch_newton(djobs, nlds, zlds, true, s));
    if (nb > nlds) HIPC(launch_newton(djobs + nlds, nb - nlds, zglob, false, s));
    HIPC(hipMemcpyAsync(res.data(), dres, L.res_d * 8, hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(res.data() + L.res_d, dout, 4 * (size_t)nb * 8, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    const double* o = res.data() + L.res_d;
    for (int b = 0; b < nb; ++b) {
        const size_t nz = P.nz[b];
        std::memcpy(P.z1[b], res.data() + L.vres[b], nz * 8);
        std::memcpy(P.z2[b], res.data() + L.vres[b] + nz, nz * 8);
        std::memcpy(P.VY1[b], res.data() + L.vres[b] + 2 * nz, nz * 8);
        P.it[b] = (int32_t)o[4 * b];
        P.err[b] = o[4 * b + 1];
        P.conv[b] = o[4 * b + 2] != 0 ? 1 : 0;
        if (P.steps) P.steps[b] = (int32_t)o[4 * b + 3];
    }
    return HQT_OK;
}

size_t newton_batch_bytes(int nz, int nb) {
    NewtonProblems P;
    P.resize(nb);
    for (int b = 0; b < nb; ++b) P.nz[b] = nz;
    return newton_problems_bytes(P);
}

int newton_batch_run(hipStream_t s, NewtonWork& w, int nz, double zh, double small, double D, int nb,
                     const double* Z1, const double* Z2, const double* vy1, const double* vy, int it_max, double tol,
                     double* z1, double* z2, double* VY1, int32_t* it, double* err, uint8_t* conv,
                     const NewtonCont* cont) {
    return newton_problems_run(s, w, uniform_problems(nz, zh, small, D, nb, Z1, Z2, vy1, vy, z1, z2, VY1, it, err,
                                                      conv, cont), it_max, tol);
}

int newton_batch_run_in(hipStream_t s, unsigned char* buf, int nz, double zh, double small, double D, int nb,
                        const double* Z1, const double* Z2, const double* vy1, const double* vy, int it_max,
                        double tol, double* z1, double* z2, double* VY1, int32_t* it, double* err, uint8_t* conv,
                        const NewtonCont* cont) {
    return newton_problems_run_in(s, buf, uniform_problems(nz, zh, small, D, nb, Z1, Z2, vy1, vy, z1, z2, VY1, it,
                                                           err, conv, cont), it_max, tol);
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

int newton_cont_args(const char* who, int nz, int nb, const double* Z2, const double* mu_seed, const double* dmu) {
    if (!mu_seed || !dmu) return hqt_fail(HQT_ERR_ARG, std::string(who) + ": NULL argument");
    for (int b = 0; b < nb; ++b) {
        const double ms = mu_seed[b], mt = Z2[(size_t)b * nz + nz - 1], h = dmu[b];
        if (std::isnan(ms)) continue;
        if (!std::isfinite(ms) || !std::isfinite(mt) || !(ms * mt > 0) || !std::isfinite(h) || !(h > 0))
            return hqt_fail(HQT_ERR_ARG, std::string(who) + ": problem " + std::to_string(b) +
                                             ": mu_seed and the target A_t(z=0) must be finite, nonzero and of one "
                                             "sign, dmu finite and positive");
    }
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

extern "C" int hqt_newton_vy_continue_batch(int32_t nz, double zh, double small, double D, int32_t nb,
                                            const double* Z1, const double* Z2, const double* vy1, const double* vy,
                                            const double* mu_seed, double dmu, int32_t it_max, double tol, double* z1,
                                            double* z2, double* VY1, int32_t* it, double* err, uint8_t* conv,
                                            int32_t* steps) {
    static const char* who = "hqt_newton_vy_continue_batch";
    if (int rc = newton_batch_args(who, nz, nb, Z1, Z2, vy1, vy, z1, z2, VY1, it, err, conv, it_max, tol)) return rc;
    if (!steps) return hqt_fail(HQT_ERR_ARG, std::string(who) + ": NULL argument");
    std::vector<double> h((size_t)nb, dmu);
    if (int rc = newton_cont_args(who, nz, nb, Z2, mu_seed, h.data())) return rc;
    const NewtonCont cont{mu_seed, h.data(), steps};
    hipStream_t s = nullptr;
    HIPC(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int rc;
    {
        NewtonWork w;
        rc = newton_batch_run(s, w, nz, zh, small, D, nb, Z1, Z2, vy1, vy, it_max, tol, z1, z2, VY1, it, err, conv,
                              &cont);
    }
    (void)hipStreamDestroy(s);
    if (rc) return rc;
    return newton_batch_status(who, nb, conv);
}
