/*
 * hqt_mex.c -- MATLAB MEX gateway over the C ABI of include/hqt.h.
 *
 * Lets the reference's driver scripts (scripts/simulation/tevol_driven_cluster.m,
 * tevol_lattice.m) replace the inline RK4 loop body (:120-142) and the F1..F4 closures
 * (:72-76) with calls into libhqt.  Uses the interleaved-complex API (mex -R2018a).
 * Requires MATLAB's mex.h; it is not built in this repository (no MATLAB here).
 *
 *   h = hqt_mex('create', nx, ny, nz, Lx, Ly, mu, dt, winding, filter, precision, ax_b, ay_b)
 *         precision 0 = FP64, 1 = mixed; ax_b / ay_b scalars or [nx,ny] arrays
 *   hqt_mex('set', h, Z1, Z2, Z3, Z4, rhoi)
 *   [Z1,Z2,Z3,Z4,rhoi] = hqt_mex('step', h, nsteps)      % :120-142 iterated nsteps times
 *   [F1,F2,F3,F4] = hqt_mex('rhs', h, Z1, Z2, Z3, Z4)     % the closures of :72-76
 *   w = hqt_mex('ff', x, y)                               % src/numerics/ff.m, any logical mask y
 *   w = hqt_mex('ff', x, fx, fy)                          % the same with the separable mask fx|fy
 *   hqt_mex('init', h, Z1, Z2, Vy1, rhoi, amp_noise, phase_noise, seed)
 *         % :81-102 on the device from the nz0-point profile (1e-2 amplitude noise as :101)
 *   hqt_mex('checkpoint', h, fileName)                    % save(fileName,'Z1',...,'rhoi') of :178-181,
 *                                                         % written by a host thread while stepping goes on
 *   [n, np, nm] = hqt_mex('countN', h)                    % src/diagnostics/countN.m on the current state
 *   hqt_mex('destroy', h)
 *
 * Errors (bad arguments, the NaN guard of :182-185, HIP failures) are raised with
 * mexErrMsgIdAndTxt("hqt:<code>", ...); no C++ exception crosses the boundary.
 */
#include "mex.h"
#include "../../include/hqt.h"

#include <stdint.h>
#include <string.h>

#define MAX_CTX 64
static hqt_ctx* g_ctx[MAX_CTX];

static void cleanup(void) {
    for (int i = 0; i < MAX_CTX; ++i)
        if (g_ctx[i]) { hqt_ctx_destroy(g_ctx[i]); g_ctx[i] = NULL; }
}

static void fail(int rc) {
    static const char* ids[] = {"hqt:ok", "hqt:arg", "hqt:nan", "hqt:hip", "hqt:oom", "hqt:state", "hqt:noconv",
                                "hqt:timeout"};
    mexErrMsgIdAndTxt(rc >= 0 && rc < 8 ? ids[rc] : "hqt:error", "%s", hqt_last_error());
}

static hqt_ctx* handle(const mxArray* a) {
    if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt("hqt:arg", "expected a uint64 handle");
    uint64_t k = *(const uint64_t*)mxGetData(a);
    if (k == 0 || k > MAX_CTX || !g_ctx[k - 1]) mexErrMsgIdAndTxt("hqt:arg", "invalid handle");
    return g_ctx[k - 1];
}

static const double* real_field(const mxArray* a, size_t n, const char* what) {
    if (!mxIsDouble(a) || mxIsComplex(a) || mxGetNumberOfElements(a) != n)
        mexErrMsgIdAndTxt("hqt:arg", "%s must be a real double array with %zu elements", what, n);
    return mxGetDoubles(a);
}

/* Z1 may be stored real by MATLAB when its imaginary part is zero: widen it. */
static const double* complex_field(const mxArray* a, size_t n, mxArray** tmp) {
    if (!mxIsDouble(a) || mxGetNumberOfElements(a) != n) mexErrMsgIdAndTxt("hqt:arg", "Z1 must be double [nx,ny,nz]");
    if (mxIsComplex(a)) { *tmp = NULL; return (const double*)mxGetComplexDoubles(a); }
    *tmp = mxDuplicateArray(a);
    if (!mxMakeArrayComplex(*tmp)) mexErrMsgIdAndTxt("hqt:oom", "mxMakeArrayComplex failed");
    return (const double*)mxGetComplexDoubles(*tmp);
}

static mxArray* field_out(int32_t nx, int32_t ny, int32_t nz, int cplx) {
    mwSize d[3] = {(mwSize)nx, (mwSize)ny, (mwSize)nz};
    return mxCreateNumericArray(3, d, mxDOUBLE_CLASS, cplx ? mxCOMPLEX : mxREAL);
}

/* fx / fy as 0/1 bytes from a logical or double vector (the reference builds them with false(...)). */
static void mask_bytes(const mxArray* a, uint8_t* out, size_t n) {
    if (mxIsLogical(a)) {
        const mxLogical* v = mxGetLogicals(a);
        for (size_t i = 0; i < n; ++i) out[i] = v[i] ? 1 : 0;
    } else if (mxIsDouble(a) && !mxIsComplex(a)) {
        const double* v = mxGetDoubles(a);
        for (size_t i = 0; i < n; ++i) out[i] = v[i] != 0.0 ? 1 : 0;
    } else {
        mexErrMsgIdAndTxt("hqt:arg", "masks must be logical or double arrays");
    }
}

typedef struct { int32_t nx, ny, nz; } dims_t;
static dims_t g_dims[MAX_CTX];

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    char cmd[16];
    if (nrhs < 1 || mxGetString(prhs[0], cmd, sizeof(cmd))) mexErrMsgIdAndTxt("hqt:arg", "first argument is a command");
    mexAtExit(cleanup);
    if (!strcmp(cmd, "create")) {
        if (nrhs != 13) mexErrMsgIdAndTxt("hqt:arg", "create needs 12 arguments");
        hqt_params p;
        memset(&p, 0, sizeof(p));
        p.nx = (int32_t)mxGetScalar(prhs[1]); p.ny = (int32_t)mxGetScalar(prhs[2]); p.nz = (int32_t)mxGetScalar(prhs[3]);
        p.Lx = mxGetScalar(prhs[4]); p.Ly = mxGetScalar(prhs[5]); p.mu = mxGetScalar(prhs[6]); p.dt = mxGetScalar(prhs[7]);
        p.winding = (int32_t)mxGetScalar(prhs[8]); p.filter = (int32_t)mxGetScalar(prhs[9]);
        p.precision = (int32_t)mxGetScalar(prhs[10]);
        size_t nxy = (size_t)p.nx * p.ny;
        if (mxGetNumberOfElements(prhs[11]) == 1) p.ax_b_const = mxGetScalar(prhs[11]);
        else p.ax_b = real_field(prhs[11], nxy, "ax_b");
        if (mxGetNumberOfElements(prhs[12]) == 1) p.ay_b_const = mxGetScalar(prhs[12]);
        else p.ay_b = real_field(prhs[12], nxy, "ay_b");
        int slot = -1;
        for (int i = 0; i < MAX_CTX; ++i) if (!g_ctx[i]) { slot = i; break; }
        if (slot < 0) mexErrMsgIdAndTxt("hqt:oom", "too many open contexts");
        int rc = hqt_ctx_create(&p, &g_ctx[slot]);
        if (rc) fail(rc);
        g_dims[slot].nx = p.nx; g_dims[slot].ny = p.ny; g_dims[slot].nz = p.nz;
        plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
        *(uint64_t*)mxGetData(plhs[0]) = (uint64_t)slot + 1;
        return;
    }
    if (!strcmp(cmd, "ff")) {
        if (nrhs != 3 && nrhs != 4) mexErrMsgIdAndTxt("hqt:arg", "ff needs x and a mask y, or x, fx, fy");
        const mwSize* d = mxGetDimensions(prhs[1]);
        int32_t nx = (int32_t)d[0], ny = (int32_t)d[1];
        int32_t ns = mxGetNumberOfDimensions(prhs[1]) > 2 ? (int32_t)d[2] : 1;
        size_t n = (size_t)nx * ny * ns;
        uint8_t *fx = NULL, *fy = NULL, *mk = NULL;
        int32_t mslices = 1;
        if (nrhs == 3) {
            /* ff(x, y): any logical mask of nx*ny (every slice) or nx*ny*ns entries */
            size_t nm = mxGetNumberOfElements(prhs[2]);
            if (nm != (size_t)nx * ny && nm != n) mexErrMsgIdAndTxt("hqt:arg", "the mask must have nx*ny or numel(x) entries");
            mslices = nm == (size_t)nx * ny ? 1 : ns;
            mk = (uint8_t*)mxMalloc(nm);
            mask_bytes(prhs[2], mk, nm);
        } else {
            if (mxGetNumberOfElements(prhs[2]) != (size_t)nx || mxGetNumberOfElements(prhs[3]) != (size_t)ny)
                mexErrMsgIdAndTxt("hqt:arg", "fx, fy must have nx, ny entries");
            fx = (uint8_t*)mxMalloc(nx);
            fy = (uint8_t*)mxMalloc(ny);
            mask_bytes(prhs[2], fx, nx);
            mask_bytes(prhs[3], fy, ny);
        }
        double* re = (double*)mxMalloc(n * sizeof(double));
        double* im = NULL;
        if (mxIsComplex(prhs[1])) {
            const mxComplexDouble* z = mxGetComplexDoubles(prhs[1]);
            im = (double*)mxMalloc(n * sizeof(double));
            for (size_t i = 0; i < n; ++i) { re[i] = z[i].real; im[i] = z[i].imag; }
        } else {
            memcpy(re, mxGetDoubles(prhs[1]), n * sizeof(double));
        }
        plhs[0] = mxCreateNumericArray(mxGetNumberOfDimensions(prhs[1]), d, mxDOUBLE_CLASS, mxREAL);
        int rc = mk ? hqt_ff_mask(nx, ny, ns, re, im, mk, mslices, mxGetDoubles(plhs[0]))
                    : hqt_ff(nx, ny, ns, re, im, fx, fy, mxGetDoubles(plhs[0]));
        mxFree(re); if (im) mxFree(im); if (fx) mxFree(fx); if (fy) mxFree(fy); if (mk) mxFree(mk);
        if (rc) fail(rc);
        return;
    }
    if (nrhs < 2) mexErrMsgIdAndTxt("hqt:arg", "missing handle");
    hqt_ctx* c = handle(prhs[1]);
    uint64_t k = *(const uint64_t*)mxGetData(prhs[1]);
    dims_t dm = g_dims[k - 1];
    size_t n = (size_t)dm.nx * dm.ny * dm.nz;
    if (!strcmp(cmd, "init")) {
        if (nrhs != 9) mexErrMsgIdAndTxt("hqt:arg", "init needs Z1, Z2, Vy1, rhoi, amp_noise, phase_noise, seed");
        const size_t nz0 = mxGetNumberOfElements(prhs[2]);
        hqt_init in;
        memset(&in, 0, sizeof(in));
        in.nz0 = (int32_t)nz0;
        in.Z1 = real_field(prhs[2], nz0, "Z1");
        in.Z2 = real_field(prhs[3], nz0, "Z2");
        in.Vy1 = real_field(prhs[4], nz0, "Vy1");
        in.rhoi = real_field(prhs[5], nz0, "rhoi");
        in.amp_noise = mxGetScalar(prhs[6]);
        in.phase_noise = mxGetScalar(prhs[7]);
        in.seed = (uint64_t)mxGetScalar(prhs[8]);
        int rc = hqt_init_state(c, &in);
        if (rc) fail(rc);
        return;
    }
    if (!strcmp(cmd, "checkpoint")) {
        char path[4096];
        if (nrhs != 3 || mxGetString(prhs[2], path, sizeof(path))) mexErrMsgIdAndTxt("hqt:arg", "checkpoint needs a file name");
        int rc = hqt_checkpoint_save_async(c, path);
        if (rc) fail(rc);
        return;
    }
    if (!strcmp(cmd, "countN")) {
        int32_t counts[3];
        int rc = hqt_vortex_census(c, counts, NULL, NULL);
        if (rc) fail(rc);
        for (int i = 0; i < 3 && i < (nlhs > 0 ? nlhs : 1); ++i) plhs[i] = mxCreateDoubleScalar((double)counts[i]);
        return;
    }
    if (!strcmp(cmd, "destroy")) {
        hqt_checkpoint_wait(c);
        hqt_ctx_destroy(c);
        g_ctx[k - 1] = NULL;
        return;
    }
    if (!strcmp(cmd, "set")) {
        if (nrhs != 7) mexErrMsgIdAndTxt("hqt:arg", "set needs Z1, Z2, Z3, Z4, rhoi");
        mxArray* tmp = NULL;
        const double* z1 = complex_field(prhs[2], n, &tmp);
        int rc = hqt_set_state(c, z1, real_field(prhs[3], n, "Z2"), real_field(prhs[4], n, "Z3"),
                               real_field(prhs[5], n, "Z4"), real_field(prhs[6], n, "rhoi"));
        if (tmp) mxDestroyArray(tmp);
        if (rc) fail(rc);
        return;
    }
    if (!strcmp(cmd, "step")) {
        int32_t ns = nrhs > 2 ? (int32_t)mxGetScalar(prhs[2]) : 1;
        int rc = hqt_step_n(c, ns);
        if (rc) fail(rc);
        plhs[0] = field_out(dm.nx, dm.ny, dm.nz, 1);
        for (int i = 1; i < 5; ++i) plhs[i] = field_out(dm.nx, dm.ny, dm.nz, 0);
        rc = hqt_get_state(c, (double*)mxGetComplexDoubles(plhs[0]), mxGetDoubles(plhs[1]), mxGetDoubles(plhs[2]),
                           mxGetDoubles(plhs[3]), mxGetDoubles(plhs[4]));
        if (rc) fail(rc);
        return;
    }
    if (!strcmp(cmd, "rhs")) {
        if (nrhs != 6) mexErrMsgIdAndTxt("hqt:arg", "rhs needs Z1, Z2, Z3, Z4");
        mxArray* tmp = NULL;
        const double* z1 = complex_field(prhs[2], n, &tmp);
        plhs[0] = field_out(dm.nx, dm.ny, dm.nz, 1);
        for (int i = 1; i < 4; ++i) plhs[i] = field_out(dm.nx, dm.ny, dm.nz, 0);
        int rc = hqt_rhs(c, z1, real_field(prhs[3], n, "Z2"), real_field(prhs[4], n, "Z3"), real_field(prhs[5], n, "Z4"),
                         (double*)mxGetComplexDoubles(plhs[0]), mxGetDoubles(plhs[1]), mxGetDoubles(plhs[2]),
                         mxGetDoubles(plhs[3]));
        if (tmp) mxDestroyArray(tmp);
        if (rc) fail(rc);
        return;
    }
    mexErrMsgIdAndTxt("hqt:arg", "unknown command %s", cmd);
}
