// Checkpoint writer in the reference's format: `check_point_t=<num2str(t)>.mat` holding Z1
// (complex), Z2, Z3, Z4 and rhoi, each [nx, ny, nz] double (tevol_driven_cluster.m:178-181;
// docs/DATA_DICTIONARY.md "Checkpoint files").  Files are MAT level 5 (uncompressed, the
// layout MATLAB writes with save -v6), which MATLAB's load and scipy.io.loadmat read.
//
// The asynchronous form snapshots the state device-to-device on the context's stream (a few
// microseconds at HBM speed), then a host thread copies the snapshot to pinned memory on its
// own stream and writes the file while the evolution goes on.
#include "hqt_internal.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <new>
#include <thread>

namespace {

enum { MI_INT8 = 1, MI_INT32 = 5, MI_UINT32 = 6, MI_DOUBLE = 9, MI_MATRIX = 14 };
constexpr uint32_t MX_DOUBLE_CLASS = 6, MX_COMPLEX = 0x0800;

size_t pad8(size_t n) { return (n + 7) & ~size_t(7); }

struct MatVar {
    const char* name;
    const double* re;       // column-major data
    const double* im;       // NULL for a real array
    int32_t dims[3];
    bool interleaved;       // re holds (re, im) pairs; im unused
};

bool put(FILE* f, const void* p, size_t n) { return n == 0 || fwrite(p, 1, n, f) == n; }
bool tag(FILE* f, uint32_t type, uint32_t bytes) {
    const uint32_t t[2] = {type, bytes};
    return put(f, t, 8);
}
bool zeros(FILE* f, size_t n) {
    static const char z[8] = {0};
    return put(f, z, n);
}

bool write_header(FILE* f) {
    char h[128];
    std::memset(h, ' ', sizeof(h));
    char txt[117];
    const std::time_t now = std::time(nullptr);
    char when[64];
    std::strftime(when, sizeof(when), "%a %b %d %H:%M:%S %Y", std::gmtime(&now));
    const int n = std::snprintf(txt, sizeof(txt), "MATLAB 5.0 MAT-file, Platform: GLNXA64, Created on: %s (libhqt)", when);
    std::memcpy(h, txt, (size_t)std::min(n, 116));
    std::memset(h + 116, 0, 8);                  // no subsystem data
    const uint16_t version = 0x0100;
    std::memcpy(h + 124, &version, 2);
    h[126] = 'I';                                // endian indicator "IM": little-endian writer
    h[127] = 'M';
    return put(f, h, sizeof(h));
}

bool write_var(FILE* f, const MatVar& v, std::vector<double>& buf) {
    const size_t n = (size_t)v.dims[0] * v.dims[1] * v.dims[2];
    const size_t nl = std::strlen(v.name);
    const bool cplx = v.interleaved || v.im;
    const size_t body = 16 + 8 + pad8(12) + 8 + pad8(nl) + (cplx ? 2 : 1) * (8 + n * 8);
    bool ok = tag(f, MI_MATRIX, (uint32_t)body);
    const uint32_t flags[2] = {MX_DOUBLE_CLASS | (cplx ? MX_COMPLEX : 0u), 0u};
    ok = ok && tag(f, MI_UINT32, 8) && put(f, flags, 8);
    ok = ok && tag(f, MI_INT32, 12) && put(f, v.dims, 12) && zeros(f, pad8(12) - 12);
    ok = ok && tag(f, MI_INT8, (uint32_t)nl) && put(f, v.name, nl) && zeros(f, pad8(nl) - nl);
    if (v.interleaved) {
        buf.resize(n);
        for (size_t i = 0; i < n; ++i) buf[i] = v.re[2 * i];
        ok = ok && tag(f, MI_DOUBLE, (uint32_t)(n * 8)) && put(f, buf.data(), n * 8);
        for (size_t i = 0; i < n; ++i) buf[i] = v.re[2 * i + 1];
        ok = ok && tag(f, MI_DOUBLE, (uint32_t)(n * 8)) && put(f, buf.data(), n * 8);
    } else {
        ok = ok && tag(f, MI_DOUBLE, (uint32_t)(n * 8)) && put(f, v.re, n * 8);
        if (v.im) ok = ok && tag(f, MI_DOUBLE, (uint32_t)(n * 8)) && put(f, v.im, n * 8);
    }
    return ok;
}

// Z1 either interleaved complex (Z1i) or split (Z1r, Z1m); Z2..rhoi real.
int write_checkpoint(const char* path, int nx, int ny, int nz, const double* Z1i, const double* Z1r,
                     const double* Z1m, const double* Z2, const double* Z3, const double* Z4, const double* rhoi) {
    const size_t n = (size_t)nx * ny * nz;
    if (n * 8 > 0xFFFFFFF0ull / 2) return hqt_fail(HQT_ERR_ARG, "checkpoint: field too large for a MAT level 5 variable");
    FILE* f = std::fopen(path, "wb");
    if (!f) return hqt_fail(HQT_ERR_ARG, std::string("checkpoint: cannot open ") + path);
    std::vector<double> buf;
    MatVar vars[5] = {{"Z1", Z1i ? Z1i : Z1r, Z1i ? nullptr : Z1m, {nx, ny, nz}, Z1i != nullptr},
                      {"Z2", Z2, nullptr, {nx, ny, nz}, false},
                      {"Z3", Z3, nullptr, {nx, ny, nz}, false},
                      {"Z4", Z4, nullptr, {nx, ny, nz}, false},
                      {"rhoi", rhoi, nullptr, {nx, ny, nz}, false}};
    bool ok = write_header(f);
    for (const MatVar& v : vars) ok = ok && write_var(f, v, buf);
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return hqt_fail(HQT_ERR_ARG, std::string("checkpoint: write failed for ") + path);
    return HQT_OK;
}

}  // namespace

// Per-context asynchronous writer state (defined here, owned by hqt_ctx::ckpt).
struct CheckpointWriter {
    void* snap[6] = {};            // R, I, Ax, Ay (TS), At, rho (FP64) on the device
    void* host = nullptr;          // pinned staging, 6 * npts doubles
    hipStream_t copy = nullptr;
    hipEvent_t ready = nullptr, done = nullptr;
    std::thread th;
    int status = HQT_OK;
    std::string msg;
    ~CheckpointWriter() {
        if (th.joinable()) th.join();
        for (void* p : snap)
            if (p) (void)hipFree(p);
        if (host) (void)hipHostFree(host);
        if (copy) (void)hipStreamDestroy(copy);
        if (ready) (void)hipEventDestroy(ready);
        if (done) (void)hipEventDestroy(done);
    }
};

void checkpoint_writer_release(hqt_ctx* c) {
    delete c->ckpt;
    c->ckpt = nullptr;
}

namespace {

int wait_writer(hqt_ctx* c) {
    CheckpointWriter* w = c->ckpt;
    if (!w || !w->th.joinable()) return HQT_OK;
    w->th.join();
    if (w->status) return hqt_fail(w->status, w->msg);
    return HQT_OK;
}

// The ensemble stream (if any) and the context stream in both directions around the snapshot.
int order_with_owner(hqt_ctx* c, hipEvent_t ev, bool before) {
    if (!c->owner) return HQT_OK;
    if (before) {
        HIPC(hipEventRecord(ev, c->owner));
        HIPC(hipStreamWaitEvent(c->stream, ev, 0));
    } else {
        HIPC(hipEventRecord(ev, c->stream));
        HIPC(hipStreamWaitEvent(c->owner, ev, 0));
    }
    return HQT_OK;
}

}  // namespace

extern "C" {

int hqt_checkpoint_name(double t, char* buf, int32_t len) {
    if (!buf || len < 1) return hqt_fail(HQT_ERR_ARG, "NULL buffer");
    char num[64];
    if (t == std::floor(t) && std::fabs(t) < 1e15) {
        std::snprintf(num, sizeof(num), "%.0f", t);
    } else {
        // num2str of a non-integer scalar: %.<d>g with d = max(floor(log10(|t|)), 0) + 5
        int d = (int)std::floor(std::log10(std::fabs(t)));
        d = (d > 0 ? d : 0) + 5;
        std::snprintf(num, sizeof(num), "%.*g", d, t);
    }
    const int n = std::snprintf(buf, (size_t)len, "check_point_t=%s.mat", num);
    if (n < 0 || n >= len) return hqt_fail(HQT_ERR_ARG, "buffer too short");
    return HQT_OK;
}

int hqt_checkpoint_write_host(const char* path, int32_t nx, int32_t ny, int32_t nz, const double* Z1, const double* Z2,
                              const double* Z3, const double* Z4, const double* rhoi) {
    if (!path || !Z1 || !Z2 || !Z3 || !Z4 || !rhoi || nx < 1 || ny < 1 || nz < 1)
        return hqt_fail(HQT_ERR_ARG, "hqt_checkpoint_write_host: bad argument");
    return write_checkpoint(path, nx, ny, nz, Z1, nullptr, nullptr, Z2, Z3, Z4, rhoi);
}

int hqt_checkpoint_save(hqt_ctx* c, const char* path) {
    if (!c || !path) return hqt_fail(HQT_ERR_ARG, "NULL argument");
    if (!c->have_state) return hqt_fail(HQT_ERR_STATE, "no state uploaded");
    if (int rc = wait_writer(c)) return rc;
    const size_t n = (size_t)c->npts;
    std::vector<double> Z1(2 * n), Z2(n), Z3(n), Z4(n), rho(n);
    const int g = hqt_get_state(c, Z1.data(), Z2.data(), Z3.data(), Z4.data(), rho.data());
    if (g != HQT_OK && g != HQT_ERR_NAN) return g;
    return write_checkpoint(path, c->nx, c->ny, c->nz, Z1.data(), nullptr, nullptr, Z2.data(), Z3.data(), Z4.data(),
                            rho.data());
}

int hqt_checkpoint_save_async(hqt_ctx* c, const char* path) {
    if (!c || !path) return hqt_fail(HQT_ERR_ARG, "NULL argument");
    if (!c->have_state) return hqt_fail(HQT_ERR_STATE, "no state uploaded");
    if (int rc = wait_writer(c)) return rc;
    // snapshots keep the device layout (planes ps apart); the copy to the host drops the padding
    const size_t n = (size_t)c->npts, es = c->es, nv = (size_t)c->nvol, nxy = (size_t)c->nxy, ps = (size_t)c->ps,
                 rp = (size_t)c->rp;
    if (!c->ckpt) {
        CheckpointWriter* w = new (std::nothrow) CheckpointWriter();
        if (!w) return hqt_fail(HQT_ERR_OOM, "host allocation failed");
        c->ckpt = w;
        const size_t sz[6] = {nv * es, nv * es, nv * es, nv * es, nv * 8, nv * 8};
        for (int k = 0; k < 6; ++k)
            if (hipMalloc(&w->snap[k], sz[k]) != hipSuccess) { w->snap[k] = nullptr; checkpoint_writer_release(c); return hqt_fail(HQT_ERR_OOM, "hipMalloc(checkpoint snapshot)"); }
        if (hipHostMalloc(&w->host, 6 * n * 8, hipHostMallocDefault) != hipSuccess) { w->host = nullptr; checkpoint_writer_release(c); return hqt_fail(HQT_ERR_OOM, "hipHostMalloc(checkpoint staging)"); }
        if (hipStreamCreateWithFlags(&w->copy, hipStreamNonBlocking) != hipSuccess || hipEventCreateWithFlags(&w->ready, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&w->done, hipEventDisableTiming) != hipSuccess) {
            checkpoint_writer_release(c);
            return hqt_fail(HQT_ERR_HIP, "checkpoint stream/event");
        }
    }
    CheckpointWriter* w = c->ckpt;
    if (int rc = order_with_owner(c, w->done, true)) return rc;
    const void* src[6] = {c->ZR, c->ZI, c->ZAx, c->ZAy, c->At, c->rho};
    const size_t sz[6] = {nv * es, nv * es, nv * es, nv * es, nv * 8, nv * 8};
    for (int k = 0; k < 6; ++k) HIPC(hipMemcpyAsync(w->snap[k], src[k], sz[k], hipMemcpyDeviceToDevice, c->stream));
    HIPC(hipEventRecord(w->ready, c->stream));
    if (int rc = order_with_owner(c, w->done, false)) return rc;
    w->status = HQT_OK;
    w->msg.clear();
    const int nx = c->nx, ny = c->ny, nz = c->nz, fp32 = c->fp32;
    const std::string file(path);
    w->th = std::thread([w, n, es, nxy, ps, rp, nx, ny, nz, fp32, file] {
        double* h = static_cast<double*>(w->host);
        auto fail_w = [w](int code, const std::string& m) { w->status = code; w->msg = m; };
        if (hipStreamWaitEvent(w->copy, w->ready, 0) != hipSuccess) return fail_w(HQT_ERR_HIP, "checkpoint: wait");
        // TS fields land in the first half of each double slot region, FP64 fields directly
        for (int k = 0; k < 6; ++k) {
            const size_t e = k < 4 ? es : 8;
            for (size_t z = 0; z < (size_t)nz; ++z)     // one pitched copy per plane (rows rp, planes ps apart)
                if (hipMemcpy2DAsync(reinterpret_cast<char*>(h + k * n) + z * nxy * e, (size_t)nx * e,
                                     static_cast<const char*>(w->snap[k]) + z * ps * e, rp * e, (size_t)nx * e,
                                     (size_t)ny, hipMemcpyDeviceToHost, w->copy) != hipSuccess)
                    return fail_w(HQT_ERR_HIP, "checkpoint: copy");
        }
        if (hipStreamSynchronize(w->copy) != hipSuccess) return fail_w(HQT_ERR_HIP, "checkpoint: sync");
        if (fp32) {
            for (int k = 0; k < 4; ++k) {   // widen in place, back to front
                const float* fsrc = reinterpret_cast<const float*>(h + k * n);
                double* dst = h + k * n;
                for (size_t i = n; i-- > 0;) dst[i] = (double)fsrc[i];
            }
        }
        // staging order: R, I, Ax, Ay, At, rho -> Z1 = R + iI, Z2 = At, Z3 = Ax, Z4 = Ay
        const int rc = write_checkpoint(file.c_str(), nx, ny, nz, nullptr, h, h + n, h + 4 * n, h + 2 * n, h + 3 * n,
                                        h + 5 * n);
        if (rc) fail_w(rc, hqt_last_error());
    });
    return HQT_OK;
}

int hqt_checkpoint_wait(hqt_ctx* c) {
    if (!c) return hqt_fail(HQT_ERR_ARG, "NULL context");
    return wait_writer(c);
}

}  // extern "C"
