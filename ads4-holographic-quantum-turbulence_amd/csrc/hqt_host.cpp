// C ABI of the AdS4 stepper (include/hqt.h): context set-up, device residency, the launch
// sequence of one RK4 step, and the stand-alone operators.
//
// One step (tevol_driven_cluster.m:120-142) is queued as:
//   4 x { line_x, line_y, zstage } RHS F1..F4 at the stage state (A_t held fixed in all 4 stages,
//                                  its d_z formed in zstage), RK bookkeeping (:120-128)
//   line_x_cons, line_y_cons       d_x A_x and d_y A_y of the updated state (:129)
//   zpost                          A_t = Lzr(Lzr(cons) - rhoi) + mu, boundary reset, O (:129-136)
//   winding                        removal of the global winding along y (:136-139)
//   line_x_filter, line_y_filter   ff on Re Psi, Im Psi, A_t, A_x, A_y (:142), with the NaN guard (:182-185)
#include "hqt_operators.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "hqt_internal.h"
#include "hqt_spec_core.h"

namespace {
thread_local std::string g_err;
}  // namespace

int hqt_fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

namespace {
inline int fail(int code, const std::string& msg) { return hqt_fail(code, msg); }
}  // namespace

// The folded filter takes the cross term through 4 x 4 slice matrices when both directions have the
// rank-4 imaginary part and a 16-line block of x-lines touches at most two z-slices (ny >= 16).
// Whether the x/y passes of line length N take the factored-DFT form.  HQT_FACTORED_STAGES and
// HQT_FOLDED_STAGES force one form.  By default:
//   * the mixed form takes it at every supported N: an FP32 transform rounds like an FFT (an error of a few
//     ulps of the line's size), while an FP32 matrix product rounds an N-term sum of operator entries times
//     values; on the 64^2 vortex lattice the folded FP32 products left a condensate error of 6e-3 after 100
//     time units, and FP32 transforms in a CPU model of the same step 2e-5;
//   * FP64 passes take it from N = 128: per RK4 step the factored FP64 passes beat the folded FP64 matrix
//     products at 128, 192, 256 and 301 (1.09x, 1.21x, 1.43x, 1.26x) and lose at 64 (0.87x).
bool factored_form(const hqt_params& p, int N) {
    if (spec_fact_id(N) < 0 || (p.flags & HQT_DENSE_LINES) || (p.flags & HQT_FOLDED_STAGES)) return false;
    if (p.flags & HQT_FACTORED_STAGES) return true;
    return p.precision == HQT_MIXED || N >= 128;
}

// The mixed form's filter passes compute w = x + (P_r - I) x (+ the cross term) instead of P_r x: the
// source passes through as the output's addend and only the masked modes' part, which is small, goes
// through the FP32 matrix products or transforms (applying P_r x itself in FP32 rounds a full N-term sum of
// values of the field's size).  After 200 steps of the driven 301^2 x 20 flow the mixed |O| then differs
// from the FP64 one by 8e-6 instead of 2.8e-5.
bool filter_correction(const hqt_ctx* c) { return c->fp32 && c->fold; }

// Data type of the filter's operator tables and intermediate fields (FP64 when the filter computes in FP64),
// and the element size of a data type.
int filt_dt(const hqt_ctx* c) { return c->filt64 ? DT_F64 : c->tdt; }
size_t dt_bytes(int dt) { return dt == DT_F64 ? 8 : 4; }

bool gfilt_on(const hqt_params& p, bool lrx, bool lry) {
    return !(p.flags & HQT_DENSE_LINES) && lrx && lry && p.ny >= 16;
}

int64_t ctx_footprint(const hqt_params& p) {
    const int64_t nxy = (int64_t)p.nx * p.ny, ncol = row_stride(p.nx) * p.ny, np = plane_stride(ncol) * p.nz,
                  nz = p.nz;
    const int64_t es = p.precision == HQT_MIXED ? 4 : 8;
    auto r8 = [](int64_t b) { return b == 0 ? 8 : b; };
    int64_t b = 0;
    b += 42 * np * es;                                   // state, stage, derivative and filter buffers (TS)
    if (p.precision == HQT_MIXED && !(p.flags & HQT_FP32_FILTER)) b += 10 * np * 4;   // the filter's FP64 intermediates
    b += 3 * np * 8;                                     // At, rho, AtR
    b += ncol * 8 + 2 * ncol * 8 + r8(p.ny * 8) + 8;     // arho, O, part, W
    b += 2 * ncol * 8 + r8(p.ny * 8) + 16;               // axb, ayb, yph, flags
    b += (int64_t)NPASS * (int64_t)sizeof(LPass) + (int64_t)sizeof(hqt_ctx::DevDesc);
    b += 2 * nxy * 8 + nxy + 8 + 2 * (int64_t)census_blocks(nxy) * 8;   // diagnostics
    b += (nz * nz + (nz - 1) * (nz - 1) + 5 * nz) * 8;   // radial operators
    // the filter's tables (3 of the 6 per direction, and the rank-4 ones) are FP64 when the mixed form's filter
    // computes in FP64
    const int64_t fes = (p.precision == HQT_MIXED && !(p.flags & HQT_FP32_FILTER)) ? 8 : es;
    b += 3 * ((int64_t)line_geom(p.nx).TL + (int64_t)line_geom(p.ny).TL) * (es + fes);   // x/y operator tables
    const bool fold = !(p.flags & HQT_DENSE_LINES);
    if (fold) b += 3 * ((int64_t)fold_geom(p.nx).TL + (int64_t)fold_geom(p.ny).TL) * (es + fes);   // folded tables
    for (int n : {p.nx, p.ny})                            // trig tables of the factored-DFT passes (+ FP64 copy)
        if (factored_form(p, n)) b += (int64_t)spec_table_len(spec_fact_id(n)) * (es + (fes != es ? 8 : 0));
    bool lr[2] = {false, false};
    for (int d = 0; d < 2; ++d) {                         // rank-4 filter tables, when used
        if (p.flags & HQT_DENSE_FILTER) break;
        const int n = d ? p.ny : p.nx;
        std::vector<uint8_t> mk;
        matlab_masks(n, mk);
        std::vector<double> t;
        if (filter_lowrank(n, mk.data(), line_geom(n), 1.0, t)) {
            lr[d] = true;
            b += 2 * (int64_t)t.size() * fes;
            if (fold) b += 2 * (8 * (int64_t)fold_geom(n).Kp + 8 * (int64_t)fold_geom(n).Mp) * fes;
        }
    }
    if (gfilt_on(p, lr[0], lr[1])) b += 5 * (((int64_t)p.ny * p.nz + 15) / 16) * 32 * 8;   // cross-term partials
    return b;
}

int hqt_sync_streams(hqt_ctx* c) {
    if (c->owner) HIPC(hipStreamSynchronize(c->owner));
    HIPC(hipStreamSynchronize(c->stream));
    return HQT_OK;
}

// Whether a pass of the context runs the FP32 kernels: the mixed form's passes, except its filter when that
// computes in FP64 (filt64).
int pass_fp32(const hqt_ctx* c, int p) { return c->fp32 && !(c->filt64 && (p == PXF || p == PYF)); }

namespace {

int dalloc(hqt_ctx* c, size_t bytes, void** out) {
    void* p = nullptr;
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(HQT_ERR_OOM, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    e = hipMemset(p, 0, bytes);
    if (e != hipSuccess) { (void)hipFree(p); return fail(HQT_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(e)); }
    c->allocs.push_back(p);
    c->bytes += (int64_t)bytes;
    *out = p;
    return HQT_OK;
}

// Volume fields: host arrays are dense [nx, ny, nz]; on the device the planes are c->ps elements apart.
// Host array [nx, ny, nz] (dense) <-> device field (rows rp, planes ps elements apart), one pitched copy per plane.
int copy_planes(const hqt_ctx* c, void* dst, const void* src, size_t es, hipMemcpyKind kind) {
    const size_t nx = (size_t)c->nx, ny = (size_t)c->ny, nxy = (size_t)c->nxy, rp = (size_t)c->rp, ps = (size_t)c->ps;
    const bool up = kind == hipMemcpyHostToDevice;
    for (size_t z = 0; z < (size_t)c->nz; ++z) {
        char* d = static_cast<char*>(dst) + z * (up ? ps : nxy) * es;
        const char* s = static_cast<const char*>(src) + z * (up ? nxy : ps) * es;
        HIPC(hipMemcpy2D(d, (up ? rp : nx) * es, s, (up ? nx : rp) * es, nx * es, ny, kind));
    }
    return HQT_OK;
}

int upload_vol(hqt_ctx* c, void* dst, const double* src, int dt) {
    if (dt == DT_F64) return copy_planes(c, dst, src, 8, hipMemcpyHostToDevice);
    std::vector<float> tmp((size_t)c->npts);
    for (size_t i = 0; i < tmp.size(); ++i) tmp[i] = (float)src[i];
    return copy_planes(c, dst, tmp.data(), 4, hipMemcpyHostToDevice);
}

int download_vol(const hqt_ctx* c, double* dst, const void* src, int dt) {
    if (dt == DT_F64) return copy_planes(c, dst, src, 8, hipMemcpyDeviceToHost);
    std::vector<float> tmp((size_t)c->npts);
    if (int rc = copy_planes(c, tmp.data(), src, 4, hipMemcpyDeviceToHost)) return rc;
    for (size_t i = 0; i < tmp.size(); ++i) dst[i] = tmp[i];
    return HQT_OK;
}

int upload_as(hqt_ctx* c, void* dst, const double* src, size_t n, int dt) {
    if (dt == DT_F64) {
        HIPC(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp(n);
        for (size_t i = 0; i < n; ++i) tmp[i] = (float)src[i];
        HIPC(hipMemcpy(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
    }
    (void)c;
    return HQT_OK;
}

// ---- line-pass descriptor helpers --------------------------------------------------------
// Lines per work-group (16 or 32) and the grid that follows from it.
void set_lb(LPass& P, int lb, int dir, int tpb = 1) {
    P.lb = lb;
    P.tpb = tpb;
    P.nxc = (P.nx + lb - 1) / lb;
    if (dir == 0) P.ntiles = (int)(((int64_t)P.ny * P.nz + lb - 1) / lb);
    else P.ntiles = P.nxc * P.nz;
}

LPass base_pass(int N, int nx, int ny, int nz, int dir, void* const* ops, int32_t* flags, int fold, int es, int64_t ps,
                int64_t rp) {
    LPass P;
    std::memset(&P, 0, sizeof(P));
    const LineGeom lg = line_geom(N);
    P.N = N;
    P.Np = lg.Np;
    P.Mp = lg.Mp;
    P.TL = lg.TL;
    P.S = P.Np + 1;
    if (fold) {
        const FoldGeom fg = fold_geom(N, es);
        P.fold = 1;
        P.Mp = fg.Mp;
        P.TL = fg.TL;
        P.S = fg.S;
        P.Kp = fg.Kp;
        P.Z = fg.Z;
    }
    P.nx = nx; P.ny = ny; P.nz = nz;
    P.ps = (int32_t)ps;
    P.rp = (int32_t)rp;
    P.pad = (int32_t)(ps - rp * ny);
    set_lb(P, 16, dir);
    for (int i = 0; i < OP_COUNT; ++i) P.ops[i] = ops[i];
    P.flags = flags;
    return P;
}

LGroup& add_group(LPass& P, const void* s0, int d0, const void* s1 = nullptr, int d1 = DT_F64) {
    LGroup& g = P.g[P.ngroups++];
    std::memset(&g, 0, sizeof(g));
    g.src[0] = s0; g.src_dt[0] = d0;
    g.ntiles = 1;
    if (s1) { g.src[1] = s1; g.src_dt[1] = d1; g.ntiles = 2; }
    return g;
}

int slot_for(LGroup& g, int op) {
    for (int s = 0; s < g.nslots; ++s)
        if (g.slot_op[s] == op) return s;
    g.slot_op[g.nslots] = op;
    return g.nslots++;
}

// Parity of an operator's circulant column: d_x and the filter's imaginary part are odd, d_xx and
// the filter's real part even.
int op_parity(int op) {
    switch (op) {
        case OP_D2X: case OP_PRX: case OP_D2Y: case OP_QRY: return 1;
        default: return -1;
    }
}

// Row sum of an operator (its value on a constant line) as far as it follows from the operator
// alone: 0 for the derivatives and the filter's imaginary part.  The filter's real part has row
// sum 1 - mask[0] (mode 0 is masked for N <= 4); build_context sets it from the masks.
int op_rowsum(int op) { return (op == OP_PRX || op == OP_QRY) ? 1 : 0; }

// A term with lr = 1 is a rank-4 operator: its `slot` is the ops index of the [V | U] table
// (no LDS slot), applied by the kernel as U (V^T line).
void add_out(LGroup& g, void* out, int odt, int tile0, int op0, int tile1 = -1, int op1 = -1,
             const void* addend = nullptr, int adt = DT_F64, int nan_check = 0, int lr0 = 0, int lr1 = 0) {
    LOut& o = g.out[g.nout++];
    std::memset(&o, 0, sizeof(o));
    o.out = out; o.out_dt = odt; o.addend = addend; o.add_dt = adt; o.nan_check = nan_check;
    o.term[o.nterms++] = LTerm{tile0, lr0 ? op0 : slot_for(g, op0), lr0, lr0 ? 0 : op_parity(op0), lr0 ? 0 : op_rowsum(op0)};
    if (tile1 >= 0)
        o.term[o.nterms++] = LTerm{tile1, lr1 ? op1 : slot_for(g, op1), lr1, lr1 ? 0 : op_parity(op1), lr1 ? 0 : op_rowsum(op1)};
    if (lr0 || lr1) g.has_lr = 1;
}

// Toeplitz tables of the x/y operators in the context precision.
int build_tables(hqt_ctx* c, int N, double L, const uint8_t* mask, int base) {
    std::vector<double> col, row;
    fourdif_colrow(N, 1, col, row);
    const double px = 2.0 * M_PI / L;                  // px = 2.*pi/L0 (:38)
    const LineGeom Np = line_geom(N);
    std::vector<std::vector<double>> tabs(6);
    toeplitz_table(col, row, Np, px, tabs[0]);         // Lx = px*Dx (:47)
    std::vector<double> cc(N), c2;
    for (int m = 0; m < N; ++m) cc[m] = px * col[m];
    circulant_square(cc, c2);                          // Lx*Lx (func_Lx applied twice)
    std::vector<double> r2(N);
    for (int m = 0; m < N; ++m) r2[m] = c2[(N - m) % N];
    toeplitz_table(c2, r2, Np, 1.0, tabs[1]);
    toeplitz_table(col, row, Np, -px, tabs[2]);
    std::vector<double> pr, pi, prr(N), pir(N);
    filter_tables(N, mask, pr, pi);
    if (c->fold) {
        // folded tables of the same six operators (columns and parities as above)
        const FoldGeom fg = fold_geom(N);
        std::vector<double> ft;
        // mixed form: the filter's real part enters as its masked part P_r - I, the pass adding the source
        // itself as the output's addend (filter_correction)
        std::vector<double> prm = pr;
        if (filter_correction(c)) prm[0] -= 1.0;
        const std::vector<double>* fc[6] = {&col, &c2, &col, &prm, &pi, &pi};
        const int fsym[6] = {-1, 1, -1, 1, -1, -1};
        const double fsc[6] = {px, 1.0, -px, 1.0, 1.0, -1.0};
        for (int t = 0; t < 6; ++t) {
            fold_table(*fc[t], fsym[t], fg, fsc[t], ft);
            const int dt = t >= 3 ? filt_dt(c) : c->tdt;     // tables 3-5: the filter's
            int rc = dalloc(c, ft.size() * dt_bytes(dt), &c->fops[base + t]);
            if (rc) return rc;
            rc = upload_as(c, c->fops[base + t], ft.data(), ft.size(), dt);
            if (rc) return rc;
        }
    }
    for (int m = 0; m < N; ++m) { prr[m] = pr[(N - m) % N]; pir[m] = pi[(N - m) % N]; }
    toeplitz_table(pr, prr, Np, 1.0, tabs[3]);
    toeplitz_table(pi, pir, Np, 1.0, tabs[4]);
    toeplitz_table(pi, pir, Np, -1.0, tabs[5]);
    for (int t = 0; t < 6; ++t) {
        const int dt = t >= 3 ? filt_dt(c) : c->tdt;
        int rc = dalloc(c, tabs[t].size() * dt_bytes(dt), &c->ops[base + t]);
        if (rc) return rc;
        rc = upload_as(c, c->ops[base + t], tabs[t].data(), tabs[t].size(), dt);
        if (rc) return rc;
    }
    const int dir = base == OP_D1X ? 0 : 1, lrbase = dir ? OP_UVY : OP_UVX;
    std::vector<double> uv, nuv;
    c->lr[dir] = !(c->p.flags & HQT_DENSE_FILTER) && filter_lowrank(N, mask, Np, 1.0, uv) &&
                 filter_lowrank(N, mask, Np, -1.0, nuv);
    if (c->lr[dir]) {
        const std::vector<double>* tt[2] = {&uv, &nuv};
        for (int t = 0; t < 2; ++t) {
            const int dt = filt_dt(c);
            int rc = dalloc(c, tt[t]->size() * dt_bytes(dt), &c->ops[lrbase + t]);
            if (rc) return rc;
            rc = upload_as(c, c->ops[lrbase + t], tt[t]->data(), tt[t]->size(), dt);
            if (rc) return rc;
            if (c->fold) {
                std::vector<double> ft;
                fold_lowrank(*tt[t], Np, fold_geom(N), ft);
                rc = dalloc(c, ft.size() * dt_bytes(dt), &c->fops[lrbase + t]);
                if (rc) return rc;
                rc = upload_as(c, c->fops[lrbase + t], ft.data(), ft.size(), dt);
                if (rc) return rc;
            }
        }
    }
    return HQT_OK;
}

// cos and sin of 2 pi a / b in long double, exact at the quarter turns.
void cos_sin_2pi(long long a, long long b, double& c, double& sn) {
    a %= b;
    if (a < 0) a += b;
    if (4 * a % b == 0) {
        const int q = (int)(4 * a / b);
        c = q == 0 ? 1.0 : q == 2 ? -1.0 : 0.0;
        sn = q == 1 ? 1.0 : q == 3 ? -1.0 : 0.0;
        return;
    }
    const long double t = 2.0L * 3.14159265358979323846264338327950288L * (long double)a / (long double)b;
    c = (double)cosl(t);
    sn = (double)sinl(t);
}

// Trig table of the factored-DFT passes for line length N (SpecGeom layout), in the context precision;
// none when N has no factored form.
int build_spec_table(hqt_ctx* c, int N, int dir) {
    const int fid = spec_fact_id(N);
    if (fid < 0) return HQT_OK;
    const int N1 = kSpecFact[fid][0], N2 = kSpecFact[fid][1], H = N2 / 2;
    std::vector<double> t((size_t)spec_table_len(fid));
    for (int m = 0; m <= H; ++m) cos_sin_2pi(m, N2, t[m], t[H + 1 + m]);
    for (int m = 0; m < N1; ++m) cos_sin_2pi(m, N1, t[2 * (H + 1) + m], t[2 * (H + 1) + N1 + m]);
    const size_t tw = 2 * (H + 1) + 2 * (size_t)N1;
    for (int k2 = 0; k2 <= H; ++k2)
        for (int n1 = 0; n1 < N1; ++n1)
            cos_sin_2pi((long long)n1 * k2, N, t[tw + 2 * ((size_t)k2 * N1 + n1)], t[tw + 2 * ((size_t)k2 * N1 + n1) + 1]);
    int rc = dalloc(c, t.size() * c->es, &c->stab[dir]);
    if (rc) return rc;
    if ((rc = upload_as(c, c->stab[dir], t.data(), t.size(), c->tdt))) return rc;
    if (!c->filt64) return HQT_OK;
    if ((rc = dalloc(c, t.size() * 8, &c->stab64[dir]))) return rc;
    return upload_as(c, c->stab64[dir], t.data(), t.size(), DT_F64);
}

// An RK-stage pass (groups of d/dx, d2/dx2, -d/dx terms only) in the factored-DFT form, when its line
// length has one; otherwise unchanged.
int make_spec_pass(hqt_ctx* c, LPass& P, int dir) {
    const int N = P.N, fid = spec_fact_id(N);
    if (fid < 0 || !c->stab[dir]) return HQT_OK;
    for (int g = 0; g < P.ngroups; ++g)
        for (int o = 0; o < P.g[g].nout; ++o)
            for (int e = 0; e < P.g[g].out[o].nterms; ++e) {
                LTerm& tm = P.g[g].out[o].term[e];
                const int op = tm.lr ? -1 : P.g[g].slot_op[tm.slot];
                if (op == OP_D1X || op == OP_D1Y) tm.sk = 1;
                else if (op == OP_D2X || op == OP_D2Y) tm.sk = 2;
                else if (op == OP_ND1X || op == OP_ND1Y) tm.sk = 3;
                else return fail(HQT_ERR_ARG, "internal: factored-DFT pass with a term that is not a derivative");
            }
    // One output per group: a group's second output becomes a group of its own on the same sources (the kernel
    // transforms them again), so a work-group needs one LDS line region.  Two outputs sharing one forward
    // transform in a second region lost in both precisions once the kernels fit four work-groups per CU
    // (mixed line_y_stage 64.6 -> 63.2 us with one output per group; FP64 needs 221 VGPRs for the second output:
    // line_x_stage 69.4 -> 81.0 us, line_y_stage 94.1 -> 117.7 us).
    const int ng0 = P.ngroups;
    for (int g = 0; g < ng0; ++g)
        while (P.g[g].nout > 1) {
            if (P.ngroups == LP_MAXG) return fail(HQT_ERR_ARG, "internal: factored-DFT pass with too many outputs");
            LGroup& h = P.g[P.ngroups++];
            h = P.g[g];
            h.nout = 1;
            h.out[0] = P.g[g].out[P.g[g].nout - 1];
            --P.g[g].nout;
        }
    for (int g = 0; g < P.ngroups; ++g)
        for (int o = 0; o < P.g[g].nout; ++o) {
            LOut& O = P.g[g].out[o];
            if (O.nterms != P.g[g].ntiles || O.gmode)
                return fail(HQT_ERR_ARG, "internal: factored-DFT output without one term per source");
            // a two-source output holds its first source's result until the second source is in
            if (P.g[g].ntiles == 2 && O.out == P.g[g].src[1])
                return fail(HQT_ERR_ARG, "internal: factored-DFT two-source output aliases its second source");
            for (int e = 0; e < O.nterms; ++e)
                if (O.term[e].tile != e || O.term[e].lr)
                    return fail(HQT_ERR_ARG, "internal: factored-DFT output without one term per source");
        }
    P.fold = 3;
    P.Mp = 64 * (fid + 1);
    P.S = N | 1;
    P.sp = 2.0 * M_PI / (dir ? c->p.Ly : c->p.Lx);
    P.st = c->stab[dir];
    set_lb(P, 16, dir, 1);
    return HQT_OK;
}

// The filter pass of one direction (the cross term through the slice matrices G) in the factored-DFT form,
// when its line length has one and the mask is one contiguous range of DFT indices whose unpaired modes
// (masked with the conjugate not masked, or the reverse) fall on complex columns of the factorization;
// otherwise unchanged.
int make_spec_filter(hqt_ctx* c, LPass& P, int dir, const std::vector<uint8_t>& mask) {
    const int N = P.N, fid = spec_fact_id(N);
    if (fid < 0 || !c->stab[dir] || !c->gfilt) return HQT_OK;
    const int N2 = kSpecFact[fid][1];
    int lo = -1, hi = -2;
    for (int k = 0; k < N; ++k)
        if (mask[k]) {
            if (lo < 0) lo = k;
            else if (k != hi + 1) return HQT_OK;          // not one range
            hi = k;
        }
    int nkp = 0, kp[2] = {0, 0};
    double ke[2] = {0, 0};
    for (int k = 1; 2 * k < N; ++k)
        if (mask[k] != mask[N - k]) {
            if (nkp == 2) return HQT_OK;
            const int k2 = k % N2;
            if (k2 == 0 || 2 * k2 == N2) return HQT_OK;    // a real column of the factorization
            kp[nkp] = k;
            ke[nkp] = -(double)((mask[k] ? 1 : 0) - (mask[N - k] ? 1 : 0));
            ++nkp;
        }
    for (int g = 0; g < P.ngroups; ++g)
        for (int o = 0; o < P.g[g].nout; ++o) {
            LOut& O = P.g[g].out[o];
            for (int e = 0; e < O.nterms; ++e) {
                LTerm& tm = O.term[e];
                const int op = tm.lr ? -1 : P.g[g].slot_op[tm.slot];
                if (tm.lr && O.gmode == 0)
                    return fail(HQT_ERR_ARG, "internal: factored-DFT filter pass with a rank-4 term outside the cross term");
                if (tm.lr) tm.sk = 0;
                else if (op == OP_PRX || op == OP_QRY) tm.sk = 4;
                else return fail(HQT_ERR_ARG, "internal: factored-DFT filter pass with an unexpected term");
            }
        }
    P.fold = 3;
    P.Mp = 64 * (fid + 1);
    P.S = N | 1;
    P.sp = 2.0 * M_PI / (dir ? c->p.Ly : c->p.Lx);
    P.st = c->stab[dir];
    if (c->filt64) P.st = c->stab64[dir];
    P.mlo = lo;
    P.mhi = hi;
    P.fcorr = filter_correction(c) ? 1 : 0;
    P.nkp = nkp;
    for (int q = 0; q < 2; ++q) { P.kp[q] = kp[q]; P.ke[q] = ke[q]; }
    set_lb(P, 16, dir, 1);
    return HQT_OK;
}

template <class F>
int run(hqt_ctx* c, int name, F&& f) {
    hipEvent_t a = nullptr, b = nullptr;
    if (c->prof) {
        HIPC(hipEventCreate(&a));
        HIPC(hipEventCreate(&b));
        HIPC(hipEventRecord(a, c->stream));
    }
    hipError_t e = f();
    if (e != hipSuccess) return fail(HQT_ERR_HIP, std::string("launch ") + kNames[name] + ": " + hipGetErrorString(e));
    if (c->prof) {
        HIPC(hipEventRecord(b, c->stream));
        c->recs.push_back({name, a, b});
    }
    return HQT_OK;
}

const Members kOne{nullptr, 1};

int line_kind(int name) {
    return (name == KN_LX_FILT || name == KN_LY_FILT) ? 2 : (name == KN_LX_CONS || name == KN_LY_CONS) ? 1 : 0;
}

int line(hqt_ctx* c, int pass, int dir, int name) {
    const LPass& H = c->hp[pass];
    return run(c, name, [&] {
        const int f32 = pass_fp32(c, pass);
        return launch_line_pass(c->dp + pass, kOne, line_blocks(H), H.ngroups, line_pass_lds_bytes(H, f32), H.Mp / 64,
                                H.lb / 16, dir, f32, line_kind(name), c->stream, H.fold, H.xcd);
    });
}

int upload_desc(hqt_ctx* c) {
    hqt_ctx::DevDesc h;
    std::memset(&h, 0, sizeof(h));
    for (int s = 0; s < 4; ++s) h.zs[s] = c->zs[s];
    h.zs[4] = c->zr;
    h.zp = c->zp;
    h.wind = c->wind;
    h.nan = c->nan;
    HIPC(hipMemcpy(c->dd, &h, sizeof(h), hipMemcpyHostToDevice));
    return HQT_OK;
}

// Kernel kind of a pass (the instantiation its launches use): 2 filter, 1 constraint, 0 stage/RHS.
int pass_kind(int p) { return (p == PXF || p == PYF) ? 2 : (p == PXC || p == PYC) ? 1 : 0; }

int pass_dir(int p) { return ((p >= PY1 && p <= PY4) || p == PYC || p == PYF || p == PYR) ? 1 : 0; }


constexpr size_t kLdsCap = 160 * 1024;

// Line tiles of every line pass: lines per tile (16 or 32), tiles per work-group (1, 2 or 4) and, for
// the folded kernels, the block order (launch order or XCD-aware), forced by the params flags (then
// 1 tile per work-group, launch order), else measured once per (pass, precision, grid, form) and
// kept for the process.  Every choice gives bitwise identical results (each output's MFMA
// accumulation order is the same), so tuning never changes numbers.
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int, int, int>, int> g_tune;   // -> 10000 * order + 100 * lb + tpb

int tune_lines(hqt_ctx* c) {
    const int forced = (c->p.flags & HQT_LINES32) ? 32 : (c->p.flags & HQT_LINES16) ? 16 : 0;
    hipEvent_t ea = nullptr, eb = nullptr;
    int rc = HQT_OK;
    for (int p = 0; p < NPASS && !rc; ++p) {
        LPass& H = c->hp[p];
        const int dir = pass_dir(p);
        int choice = forced ? 100 * forced + 1 + ((c->p.flags & HQT_ORDER_XCD) ? 10000 : 0) : 0;
        const int f32 = pass_fp32(c, p);
        const auto key = std::make_tuple(p, f32, c->nx, c->ny, c->nz, (int)H.fold);
        if (!choice) {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tune.find(key);
            if (it != g_tune.end()) choice = it->second;
        }
        if (!choice) {
            if (!ea && (hipEventCreate(&ea) != hipSuccess || hipEventCreate(&eb) != hipSuccess)) {
                rc = fail(HQT_ERR_HIP, "hipEventCreate");
                break;
            }
            float best = 0;
            for (int order : {0, 1}) {
              if (order && !H.fold) continue;                      // the dense kernels keep the launch order
              for (int lb : {16, 32}) {
                if (lb == 32 && H.fold >= 3) continue;              // the factored-DFT kernel takes 16 lines
                for (int tpb : {1, 2, 4}) {
                    if (dir == 1 && tpb != 1 && !H.fold) continue;   // dense y-lines: one tile per work-group
                    set_lb(H, lb, dir, tpb);
                    const size_t lds = line_pass_lds_bytes(H, f32);
                    if (lds > kLdsCap) continue;
                    float ms = 0;
                    hipError_t e = hipMemcpy(c->dp + p, &H, sizeof(H), hipMemcpyHostToDevice);
                    for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) {
                        if (rep == 1) e = hipEventRecord(ea, c->stream);
                        if (e == hipSuccess)
                            e = launch_line_pass(c->dp + p, kOne, line_blocks(H), H.ngroups, lds, H.Mp / 64, lb / 16, dir,
                                                 f32, pass_kind(p), c->stream, H.fold, order);
                    }
                    if (e == hipSuccess) e = hipEventRecord(eb, c->stream);
                    if (e == hipSuccess) e = hipEventSynchronize(eb);
                    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ea, eb);
                    if (e != hipSuccess) { rc = fail(HQT_ERR_HIP, std::string("line-pass tuning: ") + hipGetErrorString(e)); break; }
                    if (!choice || ms < best) { best = ms; choice = 10000 * order + 100 * lb + tpb; }
                }
                if (rc) break;
              }
              if (rc) break;
            }
            if (rc) break;
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune[key] = choice;
        }
        // choice = 10000 order + 100 lines + tiles per work-group; the dense y-line kernel processes one
        // tile per work-group and keeps the launch order
        set_lb(H, H.fold >= 3 ? 16 : (choice % 10000) / 100, dir, (dir == 1 && !H.fold) ? 1 : choice % 100);
        H.xcd = H.fold ? choice / 10000 : 0;
        if (line_pass_lds_bytes(H, f32) > kLdsCap) set_lb(H, 16, dir, 1);
    }
    // occupancy target of the radial stage kernel: fixed, so that results never depend on a timing.
    // Measured on the MI355X (301^2 for nz = 12..20, 256^2 for nz = 21..32, both precisions): 2 waves
    // per SIMD is as fast as 3 up to nz = 16 and 4-55% faster above (3 waves spill).  The params
    // flags can force 3.
    c->zwpe = (c->p.flags & HQT_RADIAL_W3) ? 3 : 2;
    if (ea) (void)hipEventDestroy(ea);
    if (eb) (void)hipEventDestroy(eb);
    if (rc) return rc;
    HIPC(hipMemcpy(c->dp, c->hp, sizeof(c->hp), hipMemcpyHostToDevice));
    return HQT_OK;
}

int enqueue_step(hqt_ctx* c) {
    int rc;
    const int zb = zblocks(c->ncol), nz = c->nz, f32 = c->fp32;
    hipStream_t st = c->stream;
    for (int s = 0; s < 4; ++s) {
        if ((rc = line(c, PX1 + s, 0, KN_LX_STAGE))) return rc;
        if ((rc = line(c, PY1 + s, 1, KN_LY_STAGE))) return rc;
        if ((rc = run(c, KN_ZSTAGE, [&] { return launch_zstage(&c->dd->zs[s], kOne, zb, nz, f32, c->zwpe, st); }))) return rc;
    }
    {
        // the constraint's two passes are independent and each fills only part of the GPU: one
        // launch for both when they have the same shape
        const LPass &X = c->hp[PXC], &Y = c->hp[PYC];
        if (X.fold == 1 && Y.fold == 1 && X.lb == 16 && Y.lb == 16 && X.Mp == Y.Mp && X.ngroups == 1 && Y.ngroups == 1) {
            const size_t lds = std::max(line_pass_lds_bytes(X, f32), line_pass_lds_bytes(Y, f32));
            if ((rc = run(c, KN_CONS, [&] {
                     return launch_cons_pair(c->dp + PXC, c->dp + PYC, line_blocks(X), line_blocks(Y), lds, X.Mp / 64, f32,
                                             st, X.xcd, Y.xcd);
                 })))
                return rc;
        } else if (X.fold == 3 && Y.fold == 3 && X.Mp == Y.Mp && X.ngroups == 1 && Y.ngroups == 1) {
            const size_t lds = std::max(line_pass_lds_bytes(X, f32), line_pass_lds_bytes(Y, f32));
            if ((rc = run(c, KN_CONS, [&] {
                     return launch_cons_pair_spec(c->dp + PXC, c->dp + PYC, line_blocks(X), line_blocks(Y), lds,
                                                  X.Mp / 64 - 1, f32, st, X.xcd, Y.xcd);
                 })))
                return rc;
        } else {
            if ((rc = line(c, PXC, 0, KN_LX_CONS))) return rc;
            if ((rc = line(c, PYC, 1, KN_LY_CONS))) return rc;
        }
    }
    if ((rc = run(c, KN_ZPOST, [&] { return launch_zpost(&c->dd->zp, kOne, zb, nz, f32, st); }))) return rc;
    if (c->p.winding)
        if ((rc = run(c, KN_WIND, [&] {
                 hipError_t e = launch_wind_partial(&c->dd->wind, kOne, c->ny, st);
                 return e != hipSuccess ? e : launch_wind_apply(&c->dd->wind, kOne, wind_apply_blocks(c->ncol, nz), f32, st);
             })))
            return rc;
    if (c->p.filter) {
        if ((rc = line(c, PXF, 0, KN_LX_FILT))) return rc;
        if ((rc = line(c, PYF, 1, KN_LY_FILT))) return rc;
    } else {
        if ((rc = run(c, KN_NAN, [&] { return launch_nancheck(&c->dd->nan, kOne, nancheck_blocks(c->nvol), f32, st); })))
            return rc;
    }
    return HQT_OK;
}

int resolve_profile(hqt_ctx* c) {
    for (auto& r : c->recs) {
        float ms = 0;
        HIPC(hipEventElapsedTime(&ms, r.a, r.b));
        c->counts[r.name] += 1;
        c->totals[r.name] += ms;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    c->recs.clear();
    return HQT_OK;
}

int build_context(hqt_ctx* c) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const size_t es = c->es, np = (size_t)c->nvol, nxy = (size_t)c->nxy, ncol = (size_t)c->ncol;
    int rc;
#define ALLOC(ptr, bytes) do { void* q_; if ((rc = dalloc(c, (bytes), &q_))) return rc; ptr = (decltype(ptr))q_; } while (0)
    ALLOC(c->ZR, np * es); ALLOC(c->ZI, np * es); ALLOC(c->ZAx, np * es); ALLOC(c->ZAy, np * es);
    for (int i = 0; i < 4; ++i) { ALLOC(c->SA[i], np * es); ALLOC(c->SB[i], np * es); ALLOC(c->AC[i], np * es); }
    ALLOC(c->dxR, np * es); ALLOC(c->dxI, np * es); ALLOC(c->dxxR, np * es); ALLOC(c->dxxI, np * es);
    ALLOC(c->dxAx, np * es); ALLOC(c->dxAy, np * es); ALLOC(c->dxxAy, np * es);
    ALLOC(c->dyR, np * es); ALLOC(c->dyI, np * es); ALLOC(c->lapR, np * es); ALLOC(c->lapI, np * es);
    ALLOC(c->divA, np * es); ALLOC(c->curl3, np * es); ALLOC(c->curl4, np * es);
    ALLOC(c->dxAt, np * es); ALLOC(c->dyAt, np * es);
    const size_t fes = dt_bytes(filt_dt(c));
    for (int i = 0; i < 5; ++i) { ALLOC(c->fa[i], np * fes); ALLOC(c->fb[i], np * fes); }
    ALLOC(c->At, np * 8); ALLOC(c->rho, np * 8); ALLOC(c->AtR, np * 8);
    // per-column arrays of the radial kernels: one entry per column of the padded plane
    ALLOC(c->arho, ncol * 8); ALLOC(c->O, 2 * ncol * 8); ALLOC(c->part, (size_t)ny * 8); ALLOC(c->W, 8);
    ALLOC(c->axb, ncol * 8); ALLOC(c->ayb, ncol * 8); ALLOC(c->yph, (size_t)ny * 8);
    ALLOC(c->flags, 4 * sizeof(int32_t));
    ALLOC(c->dp, NPASS * sizeof(LPass));
    ALLOC(c->dd, sizeof(hqt_ctx::DevDesc));
    ALLOC(c->Od, 2 * nxy * 8); ALLOC(c->cmap, nxy); ALLOC(c->ccount, 2 * sizeof(int32_t));
    ALLOC(c->cpart, 2 * (size_t)census_blocks((int64_t)nxy) * 8);

    // ---- radial operators (tevol_driven_cluster.m:15,36-52,70-71) ----
    std::vector<double> zc, DM;
    chebdif(nz, 1, zc, DM);
    const double small = 0.0, D = 0.0, zi = D, zf = 1 - small, dlz = zf - zi, pz = 2 / dlz, zh = 1.0;
    std::vector<double> Lz((size_t)nz * nz), zz(nz);
    for (int i = 0; i < nz; ++i)
        for (int j = 0; j < nz; ++j) Lz[(size_t)i * nz + j] = pz * DM[i + (size_t)nz * j];
    for (int k = 0; k < nz; ++k) zz[k] = (zc[k] - 1) * dlz / 2 + zf * 1.0;
    const int m1 = nz - 1;
    std::vector<double> A((size_t)m1 * m1), Lz1((size_t)m1 * m1, 0.0);
    for (int i = 0; i < m1; ++i)
        for (int j = 0; j < m1; ++j) A[(size_t)i * m1 + j] = Lz[(size_t)i * nz + j];
    for (int i = 0; i < m1; ++i) Lz1[(size_t)i * m1 + i] = 1.0;
    if (!lu_solve(m1, A, Lz1.data(), m1)) return fail(HQT_ERR_ARG, "Lz(1:nz-1,1:nz-1) is singular");
    std::vector<double> zbuf;
    zbuf.insert(zbuf.end(), Lz.begin(), Lz.end());
    zbuf.insert(zbuf.end(), Lz1.begin(), Lz1.end());
    const size_t off_hf = zbuf.size();
    for (int k = 0; k < nz; ++k) zbuf.push_back(0.5 * (1 - (zz[k] / zh) * (zz[k] / zh) * (zz[k] / zh)));
    const size_t off_cd = zbuf.size();
    for (int k = 0; k < nz; ++k) zbuf.push_back(-0.5 * 3 * (zz[k] * zz[k]) / (zh * zh * zh));
    const size_t off_cp = zbuf.size();
    for (int k = 0; k < nz; ++k) zbuf.push_back(-0.5 * zz[k]);
    const size_t off_f1 = zbuf.size();
    for (int k = 0; k < nz; ++k) zbuf.push_back(1 - (zz[k] / zh) * (zz[k] / zh) * (zz[k] / zh));
    const size_t off_f2 = zbuf.size();
    for (int k = 0; k < nz; ++k) zbuf.push_back(1 - zz[k] * zz[k] * zz[k]);
    ALLOC(c->zops, zbuf.size() * 8);
    HIPC(hipMemcpy(c->zops, zbuf.data(), zbuf.size() * 8, hipMemcpyHostToDevice));
    c->zo.Lz = c->zops;
    c->zo.Lz1 = c->zops + (size_t)nz * nz;
    c->zo.hf = c->zops + off_hf;
    c->zo.cd = c->zops + off_cd;
    c->zo.cp = c->zops + off_cp;
    c->zo.f1 = c->zops + off_f1;
    c->zo.f2 = c->zops + off_f2;

    // ---- boundary arrays and the winding phase (:43, :133, :139) ----
    std::vector<double> bx(ncol, 0.0), by(ncol, 0.0), yph(ny);
    for (size_t i = 0; i < nxy; ++i) {
        const size_t col = i % nx + (size_t)c->rp * (i / nx);
        bx[col] = c->p.ax_b ? c->p.ax_b[i] : c->p.ax_b_const;
        by[col] = c->p.ay_b ? c->p.ay_b[i] : c->p.ay_b_const;
    }
    for (int j = 0; j < ny; ++j) {
        const double ycheb = 2 * M_PI * j / ny;
        const double yy = (ycheb * c->p.Ly) / 2 / M_PI;
        yph[j] = 2 * M_PI * yy / c->p.Ly;
    }
    HIPC(hipMemcpy(c->axb, bx.data(), ncol * 8, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->ayb, by.data(), ncol * 8, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->yph, yph.data(), (size_t)ny * 8, hipMemcpyHostToDevice));

    // ---- x/y operator tables (:13-14, :27-31, :38-39, :47-48) ----
    std::vector<uint8_t> fx, fy;
    matlab_masks(nx, fx);
    matlab_masks(ny, fy);
    if ((rc = build_tables(c, nx, c->p.Lx, fx.data(), OP_D1X))) return rc;
    if ((rc = build_tables(c, ny, c->p.Ly, fy.data(), OP_D1Y))) return rc;
    c->gfilt = gfilt_on(c->p, c->lr[0], c->lr[1]);
    if (c->gfilt) {
        c->gblk = ((int64_t)ny * nz + 15) / 16;
        ALLOC(c->Gp, 5 * (size_t)c->gblk * 32 * 8);
    }

    // ---- line-pass descriptors ----
    const int T = c->tdt;
    void* const* ops = c->fold ? c->fops : c->ops;
    const int fo = c->fold;
    void* Sin[4][4] = {{c->ZR, c->ZI, c->ZAx, c->ZAy},
                       {c->SA[0], c->SA[1], c->SA[2], c->SA[3]},
                       {c->SB[0], c->SB[1], c->SB[2], c->SB[3]},
                       {c->SA[0], c->SA[1], c->SA[2], c->SA[3]}};
    auto stage_x = [&](LPass& P, void* const* s, const double* at) {
        P = base_pass(nx, nx, ny, nz, 0, ops, c->flags, fo, (int)c->es, c->ps, c->rp);
        LGroup* g = &add_group(P, s[0], T);
        add_out(*g, c->dxR, T, 0, OP_D1X); add_out(*g, c->dxxR, T, 0, OP_D2X);
        g = &add_group(P, s[1], T);
        add_out(*g, c->dxI, T, 0, OP_D1X); add_out(*g, c->dxxI, T, 0, OP_D2X);
        g = &add_group(P, s[2], T);
        add_out(*g, c->dxAx, T, 0, OP_D1X);
        g = &add_group(P, s[3], T);
        add_out(*g, c->dxAy, T, 0, OP_D1X); add_out(*g, c->dxxAy, T, 0, OP_D2X);
        if (at) { g = &add_group(P, at, DT_F64); add_out(*g, c->dxAt, T, 0, OP_D1X); }
    };
    auto stage_y = [&](LPass& P, void* const* s, const double* at) {
        P = base_pass(ny, nx, ny, nz, 1, ops, c->flags, fo, (int)c->es, c->ps, c->rp);
        LGroup* g = &add_group(P, s[0], T);
        add_out(*g, c->dyR, T, 0, OP_D1Y); add_out(*g, c->lapR, T, 0, OP_D2Y, -1, -1, c->dxxR, T);
        g = &add_group(P, s[1], T);
        add_out(*g, c->dyI, T, 0, OP_D1Y); add_out(*g, c->lapI, T, 0, OP_D2Y, -1, -1, c->dxxI, T);
        g = &add_group(P, s[3], T);
        add_out(*g, c->divA, T, 0, OP_D1Y, -1, -1, c->dxAx, T);              // dxAx + dyAy
        g = &add_group(P, s[2], T, c->dxAy, T);
        add_out(*g, c->curl3, T, 0, OP_D2Y, 1, OP_ND1Y);                     // dyy Ax - dy(dx Ay)
        g = &add_group(P, c->dxAx, T);
        add_out(*g, c->curl4, T, 0, OP_ND1Y, -1, -1, c->dxxAy, T);           // dxx Ay - dy(dx Ax)
        if (at) { g = &add_group(P, at, DT_F64); add_out(*g, c->dyAt, T, 0, OP_D1Y); }
    };
    for (int s = 0; s < 4; ++s) {
        stage_x(c->hp[PX1 + s], Sin[s], s == 0 ? c->At : nullptr);
        stage_y(c->hp[PY1 + s], Sin[s], s == 0 ? c->At : nullptr);
    }
    void* SR_[4] = {c->SB[0], c->SB[1], c->SB[2], c->SB[3]};
    stage_x(c->hp[PXR], SR_, c->AtR);
    stage_y(c->hp[PYR], SR_, c->AtR);
    // the RK stages' (and the RHS entry's) derivative passes of each direction in the factored-DFT form
    // where factored_form picks it
    const bool fac[2] = {c->fold && factored_form(c->p, nx), c->fold && factored_form(c->p, ny)};
    if (fac[0]) {
        if ((rc = build_spec_table(c, nx, 0))) return rc;
        for (int p : {PX1, PX2, PX3, PX4, PXR})
            if ((rc = make_spec_pass(c, c->hp[p], 0))) return rc;
    }
    if (fac[1]) {
        if ((rc = build_spec_table(c, ny, 1))) return rc;
        for (int p : {PY1, PY2, PY3, PY4, PYR})
            if ((rc = make_spec_pass(c, c->hp[p], 1))) return rc;
    }
    {
        LPass& P = c->hp[PXC];
        P = base_pass(nx, nx, ny, nz, 0, ops, c->flags, fo, (int)c->es, c->ps, c->rp);
        LGroup& g = add_group(P, c->ZAx, T);
        add_out(g, c->dxAx, T, 0, OP_D1X);
        LPass& Q = c->hp[PYC];
        Q = base_pass(ny, nx, ny, nz, 1, ops, c->flags, fo, (int)c->es, c->ps, c->rp);
        LGroup& h = add_group(Q, c->ZAy, T);
        // d_y A_y into curl3 (a stage temporary, free after the last stage); the post-step kernel
        // adds d_x A_x, so the two passes are independent
        add_out(h, c->curl3, T, 0, OP_D1Y);
        // in the factored form where the direction takes it (one launch for both when both do, as folded)
        if (fac[0] && (rc = make_spec_pass(c, P, 0))) return rc;
        if (fac[1] && (rc = make_spec_pass(c, Q, 1))) return rc;
    }
    {
        void* f[5] = {c->ZR, c->ZI, c->At, c->ZAx, c->ZAy};
        int fdt[5] = {T, T, DT_F64, T, T};
        // FT: the filter's intermediates (FP64 when the filter computes in FP64)
        const int FT = filt_dt(c), fes = (int)dt_bytes(FT);
        LPass& P = c->hp[PXF];
        P = base_pass(nx, nx, ny, nz, 0, ops, c->flags, fo, fes, c->ps, c->rp);
        LPass& Q = c->hp[PYF];
        Q = base_pass(ny, nx, ny, nz, 1, ops, c->flags, fo, fes, c->ps, c->rp);
        // Im(P) has rank 4 for the reference's masks: applied as U (V^T x) instead of a dense product.
        // Folded with both directions rank 4 (and ny >= 16): the cross term (Im P_x (x) Im P_y) x is
        // U_x G U_y^T per z-slice, so the x-pass writes partials of G (4 x 4 per slice) instead of the
        // intermediate field Im(P_x) x, and the y-pass has the one source Re(P_x) x.
        const bool corr = filter_correction(c);
        for (int i = 0; i < 5; ++i) {
            LGroup& g = add_group(P, f[i], fdt[i]);
            add_out(g, c->fa[i], FT, 0, OP_PRX, -1, -1, corr ? f[i] : nullptr, corr ? fdt[i] : DT_F64);
            if (c->gfilt) {
                double* gp = c->Gp + (size_t)i * 32 * c->gblk;
                add_out(g, gp, DT_F64, 0, OP_UVX, -1, -1, nullptr, DT_F64, 0, 1);
                g.out[1].gmode = 1;
                g.out[1].gp = gp;
                g.out[1].gtab = c->ops[OP_UVY];                                        // V_y: ny x 4
                LGroup& h = add_group(Q, c->fa[i], FT);
                add_out(h, f[i], fdt[i], 0, OP_QRY, 0, OP_NUVY, corr ? c->fa[i] : nullptr, corr ? FT : DT_F64, 1, 0, 1);
                h.out[0].gmode = 2;
                h.out[0].gp = gp;
                h.out[0].gtab = static_cast<const char*>(c->ops[OP_UVX]) + 4 * (size_t)line_geom(nx).Np * fes;  // U_x
                continue;
            }
            if (c->lr[0]) add_out(g, c->fb[i], FT, 0, OP_UVX, -1, -1, nullptr, DT_F64, 0, 1);
            else add_out(g, c->fb[i], FT, 0, OP_PIX);
            LGroup& h = add_group(Q, c->fa[i], FT, c->fb[i], FT);
            if (c->lr[1]) add_out(h, f[i], fdt[i], 0, OP_QRY, 1, OP_NUVY, corr ? c->fa[i] : nullptr, corr ? FT : DT_F64, 1, 0, 1);
            else add_out(h, f[i], fdt[i], 0, OP_QRY, 1, OP_NQIY, corr ? c->fa[i] : nullptr, corr ? FT : DT_F64, 1);
        }
        // the filter of a direction whose stage passes are factored, too
        if (!(c->p.flags & HQT_FOLDED_FILTER)) {
            if (fac[0] && (rc = make_spec_filter(c, P, 0, fx))) return rc;
            if (fac[1] && (rc = make_spec_filter(c, Q, 1, fy))) return rc;
        }
    }
    // the folded kernel keeps one output's accumulators: a group has one source or one output;
    // an FP32 pass with a double source (A_t) stages through double registers; the filter's real
    // part has row sum 1 - mask[0] (ff.m:7-11: the mean is masked when round(round(N/3)) = 1)
    for (int p = 0; p < NPASS; ++p)
        for (int g = 0; g < c->hp[p].ngroups; ++g) {
            LPass& H = c->hp[p];
            for (int o = 0; o < H.g[g].nout; ++o)
                for (int e = 0; e < H.g[g].out[o].nterms; ++e) {
                    LTerm& tm = H.g[g].out[o].term[e];
                    if (tm.lr) continue;
                    const int op = H.g[g].slot_op[tm.slot];
                    // row sums: of P_r, 1 - mask[0]; of P_r - I (the correction form), -mask[0]
                    const int corr1 = filter_correction(c) ? 1 : 0;
                    if (op == OP_PRX) tm.rsum = (fx[0] ? 0 : 1) - corr1;
                    if (op == OP_QRY) tm.rsum = (fy[0] ? 0 : 1) - corr1;
                }
            if (H.fold && H.g[g].ntiles > 1 && H.g[g].nout > 1)
                return fail(HQT_ERR_ARG, "internal: folded line pass group with two sources and two outputs");
            for (int t = 0; t < H.g[g].ntiles; ++t)
                if (pass_fp32(c, p) && H.g[g].src_dt[t] == DT_F64) {
                    if (H.fold == 1) H.fold = 2;          // folded, double prefetch registers
                    else if (H.fold == 3) H.fold = 4;     // factored-DFT, double prefetch registers
                }
        }
    HIPC(hipMemcpy(c->dp, c->hp, sizeof(c->hp), hipMemcpyHostToDevice));

    // ---- radial kernel descriptors ----
    for (int s = 0; s < 4; ++s) {
        ZStage& Z = c->zs[s];
        std::memset(&Z, 0, sizeof(Z));
        Z.nxy = c->ncol; Z.ps = c->ps; Z.nz = nz; Z.stage = s + 1; Z.mode = 0; Z.h = c->p.dt; Z.zo = c->zo;
        Z.sR = Sin[s][0]; Z.sI = Sin[s][1]; Z.sAx = Sin[s][2]; Z.sAy = Sin[s][3];
        Z.At = c->At; Z.dxAt = c->dxAt; Z.dyAt = c->dyAt;
        Z.dxR = c->dxR; Z.dxI = c->dxI; Z.dyR = c->dyR; Z.dyI = c->dyI; Z.lapR = c->lapR; Z.lapI = c->lapI;
        Z.divA = c->divA; Z.curl3 = c->curl3; Z.curl4 = c->curl4;
        Z.zR = c->ZR; Z.zI = c->ZI; Z.zAx = c->ZAx; Z.zAy = c->ZAy;
        Z.aR = c->AC[0]; Z.aI = c->AC[1]; Z.aAx = c->AC[2]; Z.aAy = c->AC[3];
        Z.rho = c->rho; Z.arho = c->arho;
        void* const* nxt = (s == 0 || s == 2) ? c->SA : c->SB;
        Z.nR = nxt[0]; Z.nI = nxt[1]; Z.nAx = nxt[2]; Z.nAy = nxt[3];
        Z.flags = c->flags;
    }
    c->zr = c->zs[0];
    c->zr.mode = 1;
    c->zr.sR = c->SB[0]; c->zr.sI = c->SB[1]; c->zr.sAx = c->SB[2]; c->zr.sAy = c->SB[3];
    c->zr.At = c->AtR;
    std::memset(&c->zp, 0, sizeof(c->zp));
    c->zp.nxy = c->ncol; c->zp.ps = c->ps; c->zp.nz = nz; c->zp.mu = c->p.mu; c->zp.zo = c->zo;
    c->zp.R = c->ZR; c->zp.I = c->ZI; c->zp.Ax = c->ZAx; c->zp.Ay = c->ZAy;
    c->zp.divA = c->dxAx; c->zp.divB = c->curl3;
    c->zp.rho = c->rho; c->zp.At = c->At; c->zp.axb = c->axb; c->zp.ayb = c->ayb; c->zp.O = c->O;
    c->zp.flags = c->flags;
    std::memset(&c->wind, 0, sizeof(c->wind));
    c->wind.nx = nx; c->wind.ny = ny; c->wind.nz = nz; c->wind.nxy = c->ncol; c->wind.ps = c->ps; c->wind.rp = c->rp;
    c->wind.O = c->O; c->wind.part = c->part; c->wind.R = c->ZR; c->wind.I = c->ZI;
    c->wind.yph = c->yph; c->wind.W = c->W; c->wind.flags = c->flags;
    c->nan.R = c->ZR; c->nan.I = c->ZI; c->nan.Ax = c->ZAx; c->nan.Ay = c->ZAy; c->nan.At = c->At;
    c->nan.n = (int64_t)np; c->nan.flags = c->flags;
#undef ALLOC
    return upload_desc(c);
}

int check_params(const hqt_params* p) {
    if (!p) return fail(HQT_ERR_ARG, "params is NULL");
    if (p->nx < 2 || p->ny < 2 || p->nx > 512 || p->ny > 512)
        return fail(HQT_ERR_ARG, "nx, ny must lie in [2, 512]");
    if (!zkernels_support_nz(p->nz)) return fail(HQT_ERR_ARG, "nz must lie in [4, 32]");
    if (!(p->Lx > 0) || !(p->Ly > 0)) return fail(HQT_ERR_ARG, "Lx, Ly must be positive");
    if (p->precision != HQT_FP64 && p->precision != HQT_MIXED) return fail(HQT_ERR_ARG, "unknown precision");
    return HQT_OK;
}

int read_flags(hqt_ctx* c, int32_t* nan_step) {
    int32_t f[2];
    HIPC(hipMemcpy(f, c->flags, sizeof(f), hipMemcpyDeviceToHost));
    if (nan_step) *nan_step = f[0];
    if (f[0] != 0) return fail(HQT_ERR_NAN, "NaN in the state at step " + std::to_string(f[0]));
    return HQT_OK;
}

void split_complex(const double* Z, size_t n, std::vector<double>& re, std::vector<double>& im) {
    re.resize(n); im.resize(n);
    for (size_t i = 0; i < n; ++i) { re[i] = Z[2 * i]; im[i] = Z[2 * i + 1]; }
}

}  // namespace

extern "C" {

const char* hqt_last_error(void) { return g_err.c_str(); }
const char* hqt_version(void) { return "hqt 0.1 (gfx950)"; }

int hqt_ctx_create(const hqt_params* p, hqt_ctx** out) {
    if (!out) return fail(HQT_ERR_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_params(p);
    if (rc) return rc;
    hqt_ctx* c = new (std::nothrow) hqt_ctx();
    if (!c) return fail(HQT_ERR_OOM, "host allocation failed");
    c->p = *p;
    c->p.ax_b = nullptr;   // copied to the device below; the caller's arrays are borrowed
    c->p.ay_b = nullptr;
    c->nx = p->nx; c->ny = p->ny; c->nz = p->nz;
    c->nxy = (int64_t)p->nx * p->ny;
    c->npts = c->nxy * p->nz;
    c->rp = row_stride(p->nx);
    c->ncol = c->rp * p->ny;
    c->ps = plane_stride(c->ncol);
    c->nvol = c->ps * p->nz;
    c->fp32 = p->precision == HQT_MIXED;
    c->tdt = c->fp32 ? DT_F32 : DT_F64;
    c->fold = (p->flags & HQT_DENSE_LINES) ? 0 : 1;
    c->es = c->fp32 ? 4 : 8;
    c->filt64 = c->fp32 && !(p->flags & HQT_FP32_FILTER);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return fail(HQT_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
    hqt_params tmp = *p;
    c->p = tmp;     // keep the caller's boundary pointers only during build_context
    rc = build_context(c);
    if (!rc) rc = tune_lines(c);
    c->p.ax_b = nullptr;
    c->p.ay_b = nullptr;
    if (rc) { hqt_ctx_destroy(c); return rc; }
    *out = c;
    return HQT_OK;
}

int hqt_ctx_destroy(hqt_ctx* c) {
    if (!c) return HQT_OK;
    if (c->owner) return fail(HQT_ERR_STATE, "context is an ensemble member; destroy the ensemble first");
    checkpoint_writer_release(c);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    for (auto& r : c->recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (void* q : c->allocs) (void)hipFree(q);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return HQT_OK;
}

int hqt_set_state(hqt_ctx* c, const double* Z1, const double* Z2, const double* Z3, const double* Z4,
                  const double* rhoi) {
    if (!c || !Z1 || !Z2 || !Z3 || !Z4 || !rhoi) return fail(HQT_ERR_ARG, "NULL argument");
    if (int rs_ = hqt_sync_streams(c)) return rs_;
    const size_t n = (size_t)c->npts;
    std::vector<double> re, im;
    split_complex(Z1, n, re, im);
    int rc;
    if ((rc = upload_vol(c, c->ZR, re.data(), c->tdt))) return rc;
    if ((rc = upload_vol(c, c->ZI, im.data(), c->tdt))) return rc;
    if ((rc = upload_vol(c, c->ZAx, Z3, c->tdt))) return rc;
    if ((rc = upload_vol(c, c->ZAy, Z4, c->tdt))) return rc;
    if ((rc = upload_vol(c, c->At, Z2, DT_F64))) return rc;
    if ((rc = upload_vol(c, c->rho, rhoi, DT_F64))) return rc;
    HIPC(hipMemset(c->flags, 0, 4 * sizeof(int32_t)));
    c->have_state = true;
    return HQT_OK;
}

int hqt_get_state(hqt_ctx* c, double* Z1, double* Z2, double* Z3, double* Z4, double* rhoi) {
    if (!c) return fail(HQT_ERR_ARG, "NULL context");
    if (!c->have_state) return fail(HQT_ERR_STATE, "no state uploaded");
    if (int rs_ = hqt_sync_streams(c)) return rs_;
    const size_t n = (size_t)c->npts;
    int rc;
    if (Z1) {
        std::vector<double> re(n), im(n);
        if ((rc = download_vol(c, re.data(), c->ZR, c->tdt))) return rc;
        if ((rc = download_vol(c, im.data(), c->ZI, c->tdt))) return rc;
        for (size_t i = 0; i < n; ++i) { Z1[2 * i] = re[i]; Z1[2 * i + 1] = im[i]; }
    }
    if (Z2 && (rc = download_vol(c, Z2, c->At, DT_F64))) return rc;
    if (Z3 && (rc = download_vol(c, Z3, c->ZAx, c->tdt))) return rc;
    if (Z4 && (rc = download_vol(c, Z4, c->ZAy, c->tdt))) return rc;
    if (rhoi && (rc = download_vol(c, rhoi, c->rho, DT_F64))) return rc;
    return read_flags(c, nullptr);
}

// Capture one step's launches on the context stream into a graph (stream capture; the step
// has no memset or copy nodes, only kernels, and all kernel arguments are fixed per context).
static int build_graph(hqt_ctx* c) {
    hipGraph_t g = nullptr;
    HIPC(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(c);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(HQT_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { c->graph = nullptr; return fail(HQT_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e)); }
    return HQT_OK;
}

int hqt_step_n(hqt_ctx* c, int32_t nsteps) {
    if (!c) return fail(HQT_ERR_ARG, "NULL context");
    if (!c->have_state) return fail(HQT_ERR_STATE, "hqt_step_n before hqt_set_state");
    if (nsteps < 0) return fail(HQT_ERR_ARG, "nsteps < 0");
    if (c->owner) HIPC(hipStreamSynchronize(c->owner));   // ensemble steps queued earlier come first
    const bool use_graph = !(c->p.flags & HQT_NO_GRAPH) && !c->prof;
    if (use_graph && nsteps > 0 && !c->graph) {
        int rc = build_graph(c);
        if (rc) return rc;
    }
    for (int i = 0; i < nsteps; ++i) {
        if (use_graph) {
            HIPC(hipGraphLaunch(c->graph, c->stream));
        } else {
            int rc = enqueue_step(c);
            if (rc) return rc;
        }
    }
    return HQT_OK;
}

int hqt_sync(hqt_ctx* c, int32_t* nan_step) {
    if (!c) return fail(HQT_ERR_ARG, "NULL context");
    if (int rs_ = hqt_sync_streams(c)) return rs_;
    int rc = resolve_profile(c);
    if (rc) return rc;
    return read_flags(c, nan_step);
}

int hqt_last_winding(hqt_ctx* c, double* w) {
    if (!c || !w) return fail(HQT_ERR_ARG, "NULL argument");
    if (int rs_ = hqt_sync_streams(c)) return rs_;
    HIPC(hipMemcpy(w, c->W, 8, hipMemcpyDeviceToHost));
    return HQT_OK;
}

int hqt_rhs(hqt_ctx* c, const double* Z1, const double* Z2, const double* Z3, const double* Z4, double* F1,
            double* F2, double* F3, double* F4) {
    if (!c || !Z1 || !Z2 || !Z3 || !Z4) return fail(HQT_ERR_ARG, "NULL argument");
    const size_t n = (size_t)c->npts;
    int rc;
    if (!c->F1R) {
#define ALLOC(ptr, bytes) do { void* q_; if ((rc = dalloc(c, (bytes), &q_))) return rc; ptr = (decltype(ptr))q_; } while (0)
        const size_t nv = (size_t)c->nvol;
        ALLOC(c->F1R, nv * 8); ALLOC(c->F1I, nv * 8); ALLOC(c->F2, nv * 8); ALLOC(c->F3, nv * 8); ALLOC(c->F4, nv * 8);
#undef ALLOC
        c->zr.F1R = c->F1R; c->zr.F1I = c->F1I; c->zr.F2 = c->F2; c->zr.F3 = c->F3; c->zr.F4 = c->F4;
        if ((rc = upload_desc(c))) return rc;
    }
    if (int rs_ = hqt_sync_streams(c)) return rs_;
    std::vector<double> re, im;
    split_complex(Z1, n, re, im);
    if ((rc = upload_vol(c, c->SB[0], re.data(), c->tdt))) return rc;
    if ((rc = upload_vol(c, c->SB[1], im.data(), c->tdt))) return rc;
    if ((rc = upload_vol(c, c->SB[2], Z3, c->tdt))) return rc;
    if ((rc = upload_vol(c, c->SB[3], Z4, c->tdt))) return rc;
    if ((rc = upload_vol(c, c->AtR, Z2, DT_F64))) return rc;
    int32_t saved[2];
    HIPC(hipMemcpy(saved, c->flags, sizeof(saved), hipMemcpyDeviceToHost));
    HIPC(hipMemset(c->flags, 0, sizeof(saved)));
    const int zb = zblocks(c->ncol);
    if ((rc = line(c, PXR, 0, KN_LX_RHS))) return rc;
    if ((rc = line(c, PYR, 1, KN_LY_RHS))) return rc;
    if ((rc = run(c, KN_ZRHS, [&] { return launch_zstage(&c->dd->zs[4], kOne, zb, c->nz, c->fp32, c->zwpe, c->stream); }))) return rc;
    if (int rs_ = hqt_sync_streams(c)) return rs_;
    HIPC(hipMemcpy(c->flags, saved, sizeof(saved), hipMemcpyHostToDevice));
    if (F1) {
        std::vector<double> r(n), i(n);
        if ((rc = download_vol(c, r.data(), c->F1R, DT_F64))) return rc;
        if ((rc = download_vol(c, i.data(), c->F1I, DT_F64))) return rc;
        for (size_t k = 0; k < n; ++k) { F1[2 * k] = r[k]; F1[2 * k + 1] = i[k]; }
    }
    if (F2 && (rc = download_vol(c, F2, c->F2, DT_F64))) return rc;
    if (F3 && (rc = download_vol(c, F3, c->F3, DT_F64))) return rc;
    if (F4 && (rc = download_vol(c, F4, c->F4, DT_F64))) return rc;
    return resolve_profile(c);
}

namespace {
// O of the current state into c->Od; with `census`, also the plaquette charges and counts.
int enqueue_census(hqt_ctx* c, bool census) {
    // the state may still be written by an ensemble's queued steps on its own stream
    if (c->owner) HIPC(hipStreamSynchronize(c->owner));
    Census A{};
    A.nxy = c->nxy; A.ps = c->ps; A.rp = c->rp; A.nx = c->nx; A.ny = c->ny; A.nz = c->nz;
    A.R = c->ZR; A.I = c->ZI;
    A.Lzb = c->zo.Lz + (size_t)(c->nz - 1) * c->nz;
    A.O = c->Od; A.charge = c->cmap; A.counts = c->ccount; A.part = c->cpart;
    HIPC(launch_order_param(A, c->tdt == DT_F32, c->stream));
    if (census) {
        HIPC(hipMemsetAsync(c->ccount, 0, 2 * sizeof(int32_t), c->stream));
        HIPC(launch_census(A, c->stream));
    }
    return HQT_OK;
}
}  // namespace

int hqt_order_parameter(hqt_ctx* c, double* O) {
    if (!c || !O) return fail(HQT_ERR_ARG, "NULL argument");
    if (!c->have_state) return fail(HQT_ERR_STATE, "no state uploaded");
    int rc;
    if ((rc = enqueue_census(c, false))) return rc;
    HIPC(hipMemcpyAsync(O, c->Od, 2 * (size_t)c->nxy * 8, hipMemcpyDeviceToHost, c->stream));
    if (int rs_ = hqt_sync_streams(c)) return rs_;
    return HQT_OK;
}

int hqt_vortex_census(hqt_ctx* c, int32_t* counts, int8_t* charge, double* Oabs) {
    if (!c || !counts) return fail(HQT_ERR_ARG, "NULL argument");
    if (!c->have_state) return fail(HQT_ERR_STATE, "no state uploaded");
    int rc;
    if ((rc = enqueue_census(c, true))) return rc;
    int32_t pm[2];
    HIPC(hipMemcpyAsync(pm, c->ccount, sizeof(pm), hipMemcpyDeviceToHost, c->stream));
    const int nb = census_blocks(c->nxy);
    std::vector<double> part(2 * (size_t)nb);
    if (Oabs) HIPC(hipMemcpyAsync(part.data(), c->cpart, part.size() * 8, hipMemcpyDeviceToHost, c->stream));
    if (charge) HIPC(hipMemcpyAsync(charge, c->cmap, (size_t)c->nxy, hipMemcpyDeviceToHost, c->stream));
    if (int rs_ = hqt_sync_streams(c)) return rs_;
    counts[0] = pm[0] + pm[1];
    counts[1] = pm[0];
    counts[2] = pm[1];
    if (Oabs) {
        double mx = 0.0, sum = 0.0;
        for (int b = 0; b < nb; ++b) {
            mx = std::fmax(mx, part[2 * b]);
            sum += part[2 * b + 1];
        }
        Oabs[0] = mx;
        Oabs[1] = sum / (double)c->nxy;
    }
    return HQT_OK;
}

int hqt_ff(int32_t nx, int32_t ny, int32_t nslices, const double* x_re, const double* x_im, const uint8_t* fx,
           const uint8_t* fy, double* w) {
    if (!x_re || !fx || !fy || !w) return fail(HQT_ERR_ARG, "NULL argument");
    if (nx < 1 || ny < 1 || nx > 512 || ny > 512 || nslices < 1) return fail(HQT_ERR_ARG, "bad shape");
    hqt_ctx tmp;   // scratch holder for allocations only
    tmp.es = 8; tmp.tdt = DT_F64;
    const size_t n = (size_t)nx * ny * nslices;
    int rc = HQT_OK;
    void *u = nullptr, *v = nullptr, *a = nullptr, *b = nullptr, *out = nullptr;
    int32_t* flags = nullptr;
    LPass* dp = nullptr;
    void* ops[OP_COUNT] = {};
    hipStream_t s = nullptr;
    do {
        if ((rc = dalloc(&tmp, n * 8, &u))) break;
        if (x_im && (rc = dalloc(&tmp, n * 8, &v))) break;
        if ((rc = dalloc(&tmp, n * 8, &a))) break;
        if ((rc = dalloc(&tmp, n * 8, &b))) break;
        if ((rc = dalloc(&tmp, n * 8, &out))) break;
        void* fp;
        if ((rc = dalloc(&tmp, 16, &fp))) break;
        flags = (int32_t*)fp;
        void* dpp;
        if ((rc = dalloc(&tmp, 2 * sizeof(LPass), &dpp))) break;
        dp = (LPass*)dpp;
        // filter tables only (D1/D2 unused by this entry point, the table slots stay NULL)
        for (int dir = 0; dir < 2; ++dir) {
            const int N = dir ? ny : nx;
            const uint8_t* mk = dir ? fy : fx;
            const LineGeom Np = line_geom(N);
            std::vector<double> pr, pi, prr(N), pir(N), t0, t1, t2;
            filter_tables(N, mk, pr, pi);
            for (int m = 0; m < N; ++m) { prr[m] = pr[(N - m) % N]; pir[m] = pi[(N - m) % N]; }
            toeplitz_table(pr, prr, Np, 1.0, t0);
            toeplitz_table(pi, pir, Np, 1.0, t1);
            toeplitz_table(pi, pir, Np, -1.0, t2);
            const int base = dir ? OP_D1Y : OP_D1X;
            std::vector<double>* tt[3] = {&t0, &t1, &t2};
            for (int k = 0; k < 3; ++k) {
                if ((rc = dalloc(&tmp, tt[k]->size() * 8, &ops[base + 3 + k]))) break;
                if (hipMemcpy(ops[base + 3 + k], tt[k]->data(), tt[k]->size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
                    rc = fail(HQT_ERR_HIP, "hipMemcpy table");
                    break;
                }
            }
            if (rc) break;
        }
        if (rc) break;
        if (hipMemcpy(u, x_re, n * 8, hipMemcpyHostToDevice) != hipSuccess) { rc = fail(HQT_ERR_HIP, "hipMemcpy x"); break; }
        if (x_im && hipMemcpy(v, x_im, n * 8, hipMemcpyHostToDevice) != hipSuccess) { rc = fail(HQT_ERR_HIP, "hipMemcpy x_im"); break; }
        LPass hp[2];
        hp[0] = base_pass(nx, nx, ny, nslices, 0, ops, flags, 0, 8, (int64_t)nx * ny, nx);
        LGroup& g = add_group(hp[0], u, DT_F64, v, DT_F64);
        if (x_im) {
            // Px (u + i v) = (Pr u - Pi v) + i (Pi u + Pr v)
            add_out(g, a, DT_F64, 0, OP_PRX, 1, OP_NPIX);
            add_out(g, b, DT_F64, 0, OP_PIX, 1, OP_PRX);
        } else {
            g.ntiles = 1;
            add_out(g, a, DT_F64, 0, OP_PRX);
            add_out(g, b, DT_F64, 0, OP_PIX);
        }
        hp[1] = base_pass(ny, nx, ny, nslices, 1, ops, flags, 0, 8, (int64_t)nx * ny, nx);
        LGroup& h = add_group(hp[1], a, DT_F64, b, DT_F64);
        add_out(h, out, DT_F64, 0, OP_QRY, 1, OP_NQIY);     // Re(Py (a + i b)) = Qr a - Qi b
        if (hipMemcpy(dp, hp, sizeof(hp), hipMemcpyHostToDevice) != hipSuccess) { rc = fail(HQT_ERR_HIP, "hipMemcpy pass"); break; }
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) { rc = fail(HQT_ERR_HIP, "stream"); break; }
        hipError_t e = launch_line_pass(dp, kOne, line_blocks(hp[0]), hp[0].ngroups, line_pass_lds_bytes(hp[0], 0), hp[0].Mp / 64,
                                        1, 0, 0, 2, s);
        if (e == hipSuccess)
            e = launch_line_pass(dp + 1, kOne, line_blocks(hp[1]), hp[1].ngroups, line_pass_lds_bytes(hp[1], 0), hp[1].Mp / 64, 1,
                                 1, 0, 2, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { rc = fail(HQT_ERR_HIP, std::string("hqt_ff: ") + hipGetErrorString(e)); break; }
        if (hipMemcpy(w, out, n * 8, hipMemcpyDeviceToHost) != hipSuccess) { rc = fail(HQT_ERR_HIP, "hipMemcpy w"); break; }
    } while (0);
    if (s) (void)hipStreamDestroy(s);
    for (void* q : tmp.allocs) (void)hipFree(q);
    tmp.allocs.clear();
    return rc;
}

int hqt_ff_mask(int32_t nx, int32_t ny, int32_t nslices, const double* x_re, const double* x_im, const uint8_t* mask,
                int32_t mask_slices, double* w) {
    if (!x_re || !mask || !w) return fail(HQT_ERR_ARG, "NULL argument");
    if (nx < 1 || ny < 1 || nx > 4096 || ny > 4096 || nslices < 1) return fail(HQT_ERR_ARG, "bad shape");
    if (mask_slices != 1 && mask_slices != nslices) return fail(HQT_ERR_ARG, "mask_slices must be 1 or nslices");
    hqt_ctx tmp;   // scratch holder for allocations only
    const size_t n = (size_t)nx * ny * nslices, nm = (size_t)nx * ny * mask_slices;
    int rc = HQT_OK;
    void *re = nullptr, *im = nullptr, *mk = nullptr, *twx = nullptr, *twy = nullptr, *a = nullptr, *b = nullptr,
         *out = nullptr;
    hipStream_t s = nullptr;
    do {
        if ((rc = dalloc(&tmp, n * 8, &re))) break;
        if (x_im && (rc = dalloc(&tmp, n * 8, &im))) break;
        if ((rc = dalloc(&tmp, nm, &mk))) break;
        if ((rc = dalloc(&tmp, (size_t)nx * 16, &twx))) break;
        if ((rc = dalloc(&tmp, (size_t)ny * 16, &twy))) break;
        if ((rc = dalloc(&tmp, n * 16, &a))) break;
        if ((rc = dalloc(&tmp, n * 16, &b))) break;
        if ((rc = dalloc(&tmp, n * 8, &out))) break;
        // cos, sin (2 pi m / N), m < N, in long double
        std::vector<double> tx(2 * (size_t)nx), ty(2 * (size_t)ny);
        for (int m = 0; m < nx; ++m) {
            const long double t = 2.0L * (long double)M_PI * m / nx;
            tx[2 * m] = (double)cosl(t); tx[2 * m + 1] = (double)sinl(t);
        }
        for (int m = 0; m < ny; ++m) {
            const long double t = 2.0L * (long double)M_PI * m / ny;
            ty[2 * m] = (double)cosl(t); ty[2 * m + 1] = (double)sinl(t);
        }
        std::vector<uint8_t> mb(nm);
        for (size_t i = 0; i < nm; ++i) mb[i] = mask[i] ? 1 : 0;
        if (hipMemcpy(re, x_re, n * 8, hipMemcpyHostToDevice) != hipSuccess ||
            (x_im && hipMemcpy(im, x_im, n * 8, hipMemcpyHostToDevice) != hipSuccess) ||
            hipMemcpy(mk, mb.data(), nm, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(twx, tx.data(), tx.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(twy, ty.data(), ty.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
            rc = fail(HQT_ERR_HIP, "hqt_ff_mask: hipMemcpy");
            break;
        }
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) { rc = fail(HQT_ERR_HIP, "stream"); break; }
        hipError_t e = launch_ff_mask(static_cast<const double*>(re), static_cast<const double*>(im),
                                      static_cast<const uint8_t*>(mk), mask_slices == 1 ? 0 : (int64_t)nx * ny,
                                      static_cast<const double*>(twx), static_cast<const double*>(twy),
                                      static_cast<double2*>(a), static_cast<double2*>(b), static_cast<double*>(out), nx,
                                      ny, nslices, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { rc = fail(HQT_ERR_HIP, std::string("hqt_ff_mask: ") + hipGetErrorString(e)); break; }
        if (hipMemcpy(w, out, n * 8, hipMemcpyDeviceToHost) != hipSuccess) { rc = fail(HQT_ERR_HIP, "hipMemcpy w"); break; }
    } while (0);
    if (s) (void)hipStreamDestroy(s);
    for (void* q : tmp.allocs) (void)hipFree(q);
    tmp.allocs.clear();
    return rc;
}

int hqt_fourdif(int32_t N, int32_t m, double* x, double* DM) {
    if (N < 1 || m < 0 || !x || !DM) return fail(HQT_ERR_ARG, "bad argument");
    fourdif_dense(N, m, x, DM);
    return HQT_OK;
}

int hqt_chebdif(int32_t N, int32_t M, double* x, double* DM) {
    if (N < 2 || M < 1 || M > N - 1 || !x || !DM) return fail(HQT_ERR_ARG, "need N >= 2 and 0 < M <= N-1");
    std::vector<double> xv, dm;
    chebdif(N, M, xv, dm);
    std::memcpy(x, xv.data(), N * sizeof(double));
    std::memcpy(DM, dm.data(), dm.size() * sizeof(double));
    return HQT_OK;
}

int hqt_newton_vy(int32_t nz, double zh, double small, double D, const double* Z1, const double* Z2,
                  const double* vy1, double vy, double* z1, double* z2, double* VY1, int32_t* it, double* err) {
    if (nz < 3 || !Z1 || !Z2 || !vy1 || !z1 || !z2 || !VY1 || !it || !err) return fail(HQT_ERR_ARG, "bad argument");
    int iti = 0;
    const int r = newton_vy(nz, zh, small, D, Z1, Z2, vy1, vy, z1, z2, VY1, &iti, err);
    *it = iti;
    if (r) return fail(HQT_ERR_NOCONV, "NewtonRaphson_vy: no convergence (it=" + std::to_string(iti) + ")");
    return HQT_OK;
}

int hqt_debug_field(hqt_ctx* c, const char* name, double* out) {
    if (!c || !name || !out) return fail(HQT_ERR_ARG, "NULL argument");
    struct { const char* n; const void* p; int dt; } tab[] = {
        {"dxR", c->dxR, c->tdt}, {"dxI", c->dxI, c->tdt}, {"dxxR", c->dxxR, c->tdt}, {"dxxI", c->dxxI, c->tdt},
        {"dxAx", c->dxAx, c->tdt}, {"dxAy", c->dxAy, c->tdt}, {"dxxAy", c->dxxAy, c->tdt},
        {"dyR", c->dyR, c->tdt}, {"dyI", c->dyI, c->tdt}, {"lapR", c->lapR, c->tdt}, {"lapI", c->lapI, c->tdt},
        {"divA", c->divA, c->tdt}, {"curl3", c->curl3, c->tdt}, {"curl4", c->curl4, c->tdt},
        {"dxAt", c->dxAt, c->tdt}, {"dyAt", c->dyAt, c->tdt}};
    for (auto& t : tab) {
        if (std::strcmp(t.n, name) == 0) {
            if (int rs_ = hqt_sync_streams(c)) return rs_;
            return download_vol(c, out, t.p, t.dt);
        }
    }
    return fail(HQT_ERR_ARG, std::string("unknown field ") + name);
}

int hqt_calibration_copy(int64_t bytes, double* ms) {
    if (bytes < 32 || !ms) return fail(HQT_ERR_ARG, "bytes must be >= 32");
    bytes -= bytes % 32;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = HQT_OK;
    hipError_t e = hipMalloc(&a, (size_t)bytes);
    if (e == hipSuccess) e = hipMalloc(&b, (size_t)bytes);
    if (e == hipSuccess) e = hipMemset(a, 0, (size_t)bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_copy_calib(a, b, bytes, nullptr);     // untimed: first touch of both buffers
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) e = launch_copy_calib(a, b, bytes, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e != hipSuccess) rc = fail(HQT_ERR_HIP, std::string("hqt_calibration_copy: ") + hipGetErrorString(e));
    *ms = t;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int hqt_profile(hqt_ctx* c, int32_t enable) {
    if (!c) return fail(HQT_ERR_ARG, "NULL context");
    if (int rs_ = hqt_sync_streams(c)) return rs_;
    int rc = resolve_profile(c);
    if (rc) return rc;
    c->prof = enable != 0;
    for (int i = 0; i < KN_COUNT; ++i) { c->counts[i] = 0; c->totals[i] = 0; }
    return HQT_OK;
}

int hqt_kernel_stats(hqt_ctx* c, int32_t idx, char* name, int32_t name_len, int64_t* count, double* total_ms) {
    if (!c) return fail(HQT_ERR_ARG, "NULL context");
    if (idx < 0 || idx >= KN_COUNT) return fail(HQT_ERR_ARG, "index past the last kernel");
    if (int rs_ = hqt_sync_streams(c)) return rs_;
    int rc = resolve_profile(c);
    if (rc) return rc;
    if (name && name_len > 0) std::snprintf(name, name_len, "%s", kNames[idx]);
    if (count) *count = c->counts[idx];
    if (total_ms) *total_ms = c->totals[idx];
    return HQT_OK;
}

}  // extern "C"
