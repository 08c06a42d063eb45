// Internal definitions shared by the host-side translation units of libhqt: the context
// structure, status helpers and the launch sequence of one step.
#pragma once
#include "hqt.h"
#include "hqt_kernels.h"

#include <string>
#include <vector>

using namespace hqt;

struct CheckpointWriter;

// Sets the thread-local message returned by hqt_last_error() and returns code.
int hqt_fail(int code, const std::string& msg);

#define HIPC(call)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) return hqt_fail(HQT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

enum KName {
    KN_LX_STAGE, KN_LY_STAGE, KN_ZSTAGE, KN_LX_CONS, KN_LY_CONS, KN_ZPOST, KN_WIND,
    KN_LX_FILT, KN_LY_FILT, KN_NAN, KN_LX_RHS, KN_LY_RHS, KN_ZRHS, KN_CONS, KN_COUNT
};
inline constexpr const char* kNames[KN_COUNT] = {
    "line_x_stage", "line_y_stage", "zstage", "line_x_cons", "line_y_cons", "zpost", "winding",
    "line_x_filter", "line_y_filter", "nancheck", "line_x_rhs", "line_y_rhs", "zrhs", "line_cons"};

enum Op {
    OP_D1X, OP_D2X, OP_ND1X, OP_PRX, OP_PIX, OP_NPIX, OP_D1Y, OP_D2Y, OP_ND1Y, OP_QRY, OP_QIY, OP_NQIY,
    OP_UVX, OP_NUVX, OP_UVY, OP_NUVY,   // rank-4 forms of +-Im(P_x), +-Im(P_y) ([V | U] tables)
    OP_COUNT
};

enum Pass { PX1, PX2, PX3, PX4, PY1, PY2, PY3, PY4, PXC, PYC, PXF, PYF, PXR, PYR, NPASS };

struct hqt_ctx {
    hqt_params p{};
    int nx = 0, ny = 0, nz = 0;
    int64_t nxy = 0, npts = 0;
    // rows rp elements apart (nx rounded up to 16: every 16-point row tile is one aligned 128-byte segment in
    // FP64); ncol = rp * ny columns per plane, the padding columns included (radial kernels run over them)
    int64_t rp = 0, ncol = 0;
    // volume fields: planes ps elements apart (nxy rounded up to a multiple of 32, so every z-plane
    // starts on a 128-byte boundary); nvol = ps * nz elements per field
    int64_t ps = 0, nvol = 0;
    int fp32 = 0;
    int tdt = DT_F64;
    size_t es = 8;
    // mixed form: the filter's passes compute in FP64 (FP32 fields in and out, FP64 intermediates fa / fb and
    // FP64 operator and trig tables): an FP32 filter leaves its masked modes at the rounding level of the whole
    // field every step, and the dynamics grow them
    int filt64 = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    int64_t bytes = 0;               // device bytes allocated by this context
    int zwpe = 2;                    // occupancy target of the radial stage kernel (tuned)
    bool lr[2] = {false, false};     // rank-4 tables of the filter's imaginary part exist (x, y)

    void *ZR, *ZI, *ZAx, *ZAy;
    void *SA[4], *SB[4], *AC[4];
    void *dxR, *dxI, *dxxR, *dxxI, *dxAx, *dxAy, *dxxAy;
    void *dyR, *dyI, *lapR, *lapI, *divA, *curl3, *curl4;
    void *dxAt, *dyAt;
    void *fa[5], *fb[5];
    double *At, *rho, *arho, *O, *part, *W, *axb, *ayb, *yph, *zops, *AtR;
    double *F1R = nullptr, *F1I = nullptr, *F2 = nullptr, *F3 = nullptr, *F4 = nullptr;
    int32_t* flags = nullptr;
    void* ops[OP_COUNT];
    void* fops[OP_COUNT];            // folded tables of the same operators (NULL where unused)
    int fold = 1;                    // line passes in the even/odd folded form
    void* stab[2] = {nullptr, nullptr};   // trig tables of the factored-DFT passes (x, y), or NULL
    void* stab64[2] = {nullptr, nullptr}; // the same in FP64 for the filter's passes (filt64), or NULL
    int gfilt = 0;                   // folded filter with the cross term through the 4 x 4 slice matrices
    int64_t gblk = 0;                // 16-line blocks of the x-lines (partials of the cross term)
    double* Gp = nullptr;            // [5 fields][gblk][2][16] cross-term partials
    LPass hp[NPASS];
    LPass* dp = nullptr;
    ZOps zo{};
    ZStage zs[4];
    ZStage zr;
    ZPost zp;
    Wind wind;
    NanCheck nan{};
    // device copies of the radial/winding/guard descriptors (kernels read them from memory)
    struct DevDesc { ZStage zs[5]; ZPost zp; Wind wind; NanCheck nan; };
    DevDesc* dd = nullptr;
    hipStream_t owner = nullptr;     // stream of an ensemble that steps this context, if any
    CheckpointWriter* ckpt = nullptr;  // asynchronous checkpoint writer (created on first use)
    double* Od = nullptr;            // [2*nxy] condensate of the current state (diagnostics)
    int8_t* cmap = nullptr;          // [nxy] plaquette charges
    int32_t* ccount = nullptr;       // [2] +1 / -1 plaquette counts
    double* cpart = nullptr;         // [2*census_blocks] per-block max|O|, sum|O|
    bool have_state = false;

    hipGraphExec_t graph = nullptr;  // one captured step (built on first use)

    bool prof = false;
    struct Rec { int name; hipEvent_t a, b; };
    std::vector<Rec> recs;
    int64_t counts[KN_COUNT] = {};
    double totals[KN_COUNT] = {};
};

// Waits for the context's own stream and, while it is an ensemble member, the ensemble's.
int hqt_sync_streams(hqt_ctx* c);

// Batched stationary Newton solves on a stream, with a device workspace that grows on demand
// and is reused across calls (no allocation once it is large enough).  Inputs and outputs are
// host arrays as in hqt_newton_vy_batch; returns HQT_OK or an error status (non-convergence is
// reported per problem in conv, not as a status).
struct NewtonWork {
    unsigned char* buf = nullptr;
    size_t cap = 0;
    ~NewtonWork();
};
// mu-continuation of a batch (all three arrays [nb]): problem b starts from its guess rescaled to
// A_t(z=0) = mu_seed[b] (NaN: one plain solve) and steps mu by at most dmu[b]; steps[b] receives the
// number of solves (0 for a plain solve).
struct NewtonCont {
    const double* mu_seed;
    const double* dmu;
    int32_t* steps;
};
int newton_batch_run(hipStream_t s, NewtonWork& w, int nz, double zh, double small, double D, int nb,
                     const double* Z1, const double* Z2, const double* vy1, const double* vy, int it_max, double tol,
                     double* z1, double* z2, double* VY1, int32_t* it, double* err, uint8_t* conv,
                     const NewtonCont* cont = nullptr);

// A Newton launch over problems of any radial sizes and operator parameters: per problem nz, zh, small, D,
// vy, continuation seed and step, and host pointers to its [nz] guesses and results; it, err, conv (and
// steps, if not NULL) are [nb].  One work-group per problem, in at most two kernel launches (problems with
// the Jacobian in LDS, then the rest).
struct NewtonProblems {
    int nb = 0;
    std::vector<int> nz;
    std::vector<double> zh, small, D, vy, mu_seed, dmu;
    std::vector<const double*> Z1, Z2, vy1;
    std::vector<double*> z1, z2, VY1;
    int32_t* it = nullptr;
    double* err = nullptr;
    uint8_t* conv = nullptr;
    int32_t* steps = nullptr;
    void resize(int n);   // n problems with zh = 1, small = D = 0, no continuation, NULL pointers
};
size_t newton_problems_bytes(const NewtonProblems& P);
int newton_problems_run(hipStream_t s, NewtonWork& w, const NewtonProblems& P, int it_max, double tol);
int newton_problems_run_in(hipStream_t s, unsigned char* buf, const NewtonProblems& P, int it_max, double tol);

// Device bytes newton_batch_run needs for nb problems of radial size nz, and the same run in a caller-given
// device buffer of at least that many bytes (no allocation).
size_t newton_batch_bytes(int nz, int nb);
int newton_batch_run_in(hipStream_t s, unsigned char* buf, int nz, double zh, double small, double D, int nb,
                        const double* Z1, const double* Z2, const double* vy1, const double* vy, int it_max,
                        double tol, double* z1, double* z2, double* VY1, int32_t* it, double* err, uint8_t* conv,
                        const NewtonCont* cont = nullptr);

// Argument checks of the batched Newton entry points (it_max <= 0 and tol <= 0 select the reference's
// 1000 and 1e-6), and their return status from the per-problem convergence flags.
int newton_batch_args(const char* who, int nz, int nb, const void* Z1, const void* Z2, const void* vy1, const void* vy,
                      const void* z1, const void* z2, const void* VY1, const void* it, const void* err,
                      const void* conv, int32_t& it_max, double& tol);
int newton_batch_status(const char* who, int nb, const uint8_t* conv);
// Checks the continuation arguments of nb problems: each non-NaN seed finite, nonzero and of the sign of
// the problem's target A_t(z=0) (Z2's last entry), its step finite and positive.
int newton_cont_args(const char* who, int nz, int nb, const double* Z2, const double* mu_seed, const double* dmu);

// Device bytes a context with these parameters allocates (state, stage buffers, operator
// tables, descriptors), as hqt_ctx_create would; used for admission by HBM budget.
int64_t ctx_footprint(const hqt_params& p);

// Joins a pending asynchronous checkpoint write and frees the writer's buffers.
void checkpoint_writer_release(hqt_ctx* c);

// Work-groups of a line pass along grid.x (line tiles / tiles per work-group, rounded up).
// Whether pass p of the context runs the FP32 kernels (hqt_host.cpp).
int pass_fp32(const hqt_ctx* c, int p);
inline int line_blocks(const LPass& P) { return (P.ntiles + P.tpb - 1) / P.tpb; }
