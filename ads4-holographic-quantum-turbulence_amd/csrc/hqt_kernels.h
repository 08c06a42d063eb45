// Internal interface between the host library (hqt_host.cpp) and the gfx950 kernels
// (hqt_kernels.hip).  Plain structs passed by value as kernel arguments.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hqt {

enum Dtype : int32_t { DT_F64 = 0, DT_F32 = 1 };

// ---------------------------------------------------------------------------------------
// Members of a launch.  A kernel launch serves one or more independent evolutions (members)
// of one kernel class; its descriptors are an array in device memory, one per member, and
// member m owns blocks [boff[m], boff[m+1]) of grid.x.  nmem = 1 (boff unused) is the
// single-evolution case.
// ---------------------------------------------------------------------------------------
struct Members {
    const int32_t* boff;   // [nmem + 1] block prefix offsets
    int32_t nmem;
};
struct MemberBlock { int m, bx, nb; };   // member, block within the member, member's block count

// XCD-aware block order.  The dispatcher hands linear block i of a launch to XCD i mod 8 (8 XCDs, each
// with its own L2), so consecutive block indices land on different L2s.  xcd_block(i, n) is a
// permutation of [0, n) that gives the blocks one XCD receives consecutive logical indices: tiles
// next to each other in memory (which share cache lines at their edges) then run on one L2.
__device__ __forceinline__ int xcd_block(int i, int n) {
    const int q = n >> 3, rem = n & 7, x = i & 7, local = i >> 3;
    return x * q + (x < rem ? x : rem) + local;
}

// member_block for an explicit (logical) block index bx of a grid of nb blocks along x.
__device__ __forceinline__ MemberBlock member_block_at(const Members M, int bx, int nb) {
    MemberBlock r{0, bx, nb};
    if (M.nmem > 1) {
        int lo = 0, hi = M.nmem - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (M.boff[mid] <= r.bx) lo = mid;
            else hi = mid - 1;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);
        r.m = lo;
        r.bx -= M.boff[lo];
        r.nb = M.boff[lo + 1] - M.boff[lo];
    }
    return r;
}

__device__ __forceinline__ MemberBlock member_block(const Members M) {
    MemberBlock r{0, (int)blockIdx.x, (int)gridDim.x};
    if (M.nmem > 1) {
        int lo = 0, hi = M.nmem - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (M.boff[mid] <= r.bx) lo = mid;
            else hi = mid - 1;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);
        r.m = lo;
        r.bx -= M.boff[lo];
        r.nb = M.boff[lo + 1] - M.boff[lo];
    }
    return r;
}

// ---------------------------------------------------------------------------------------
// Line pass: apply Toeplitz/circulant line operators along x (dir 0) or y (dir 1).
// One work-group owns a tile of 16 lines and one "group" of outputs.  An output is a sum
// of up to two terms op(tile) plus an optional point-wise addend field.
// ---------------------------------------------------------------------------------------
constexpr int LP_MAXG = 8;      // groups per pass
constexpr int LP_MAXOPS = 16;   // operator table entries per pass
constexpr int LP_THREADS = 256;

// A term of an output: op(tile).  lr = 0: dense Toeplitz operator in LDS slot `slot`;
// lr = 1: rank-4 operator U V^T with the [V (Np x 4) | U (Mp x 4)] row-major table at P.ops[slot].
// pad: parity of a dense operator's column (+1 even, -1 odd); rsum: the operator's row sum
// (C 1 = rsum 1: 1 for the filter's real part, 0 for the derivatives).  Both used by the folded form.
// sk (factored-DFT passes): the term's spectral multiplier, 1 d/dx, 2 d2/dx2, 3 -d/dx, 4 the filter's
// real part (the even part of the mask's complement, LPass::mlo .. mhi); 0 for a rank-4 term (its
// output's gmode carries it).
struct LTerm { int32_t tile; int32_t slot; int32_t lr; int32_t pad; int32_t rsum; int32_t sk; };
// gmode (folded filter, the rank-4 x rank-4 cross term of Re((P_x (x) P_y) x) as U_x G U_y^T per z-slice):
//   1 (x-pass): instead of storing the output, write per-16-line-block partials of G = D_x V_y, D_x the
//     output's rank-4 dots, to gp ([block][2 slices][4 x 4] doubles; gtab: V_y, ny x 4);
//   2 (y-pass): the rank-4 term's dots are U_x G (gtab: U_x, nx x 4), G summed from gp's partials.
struct LOut {
    void* out; const void* addend;
    int32_t out_dt, add_dt;
    int32_t nterms; LTerm term[2];
    int32_t nan_check;
    int32_t gmode;
    const double* gp; const void* gtab;
};
struct LGroup {
    int32_t ntiles; int32_t has_lr; int32_t src_dt[2]; const void* src[2];
    int32_t nslots; int32_t slot_op[4];
    int32_t nout; LOut out[2];
};
struct LPass {
    int32_t N, Np, S;          // line length, padded length (multiple of 16), LDS row stride
    int32_t Mp, TL;            // output positions padded to 64 (4 waves x Q tiles), Toeplitz table length
    int32_t nx, ny, nz;
    int32_t lb;                // lines per line tile: 16 or 32
    int32_t tpb;               // line tiles per work-group (grid.x = ceil(ntiles / tpb))
    int32_t ntiles;            // line tiles of lb lines
    int32_t nxc;               // dir 1: x chunks of lb
    int32_t ngroups;           // grid.y
    LGroup g[LP_MAXG];
    const void* ops[LP_MAXOPS]; // Toeplitz tables: tab[i - j + Np + 4] = T[i][j], TL entries
                                // (folded form: tab[d + Z] = c[d mod N] / 2, TL entries)
    int32_t fold;              // even/odd folded form (hqt_fold_core.h; Mp = 64 QF, S its LDS row stride):
                               // 1, or 2 when a source is double in an FP32 pass (wide prefetch registers);
                               // 3: the factored-DFT form (hqt_spec_core.h), 4 the same with double prefetch
    int32_t Kp, Z;             // folded form: padded folded positions, table offset
    int32_t xcd;               // folded form: block order of the launch (1: XCD-aware), set by the autotuner
    int32_t* flags;            // [0] = NaN step (0: none), [1] = step counter
    int32_t ps, pad;           // plane stride of the volume fields (elements) and ps - rp*ny
    int32_t rp;                // row stride of the volume fields (elements): nx rounded up to 16
    // factored-DFT form (fold = 3 or 4, hqt_spec_core.h): Mp = 64 (factorization index + 1), S the LDS row
    // stride, sp = 2 pi / L (the derivative scale of the direction), st its trig table (pass precision)
    double sp;
    const void* st;
    // factored-DFT filter: the masked DFT indices are mlo .. mhi; the unpaired ones (masked with the
    // conjugate index not masked, or the reverse) are kp[0 .. nkp), ke = -1 or +1 the sign
    // -(mask[kp] - mask[N - kp]) of the rank-4 imaginary part they leave
    int32_t mlo, mhi, nkp, kp[2];
    double ke[2];
    int32_t fcorr;             // filter term in the correction form: the multiplier is -(mask(k) + mask(-k)) / 2
                               // (the masked part, the pass adding the source as the output's addend)
};

// Volume fields are [nx, ny, nz] with rows rp >= nx elements apart (rp a multiple of 16, so that a
// 16-point tile of a row is one aligned 128-byte segment in FP64 and half of one in FP32) and planes
// ps >= rp*ny elements apart (a multiple of 32).  Element offset of x-line line0 + d (line = y + ny*z;
// d >= 0, a few tiles at most) given z0 = line0 / ny, without a division: line*rp + z*(ps - rp*ny).
__host__ __device__ inline int xline_base_at(const LPass& P, int line0, int z0, int d) {
    int r = line0 + d - z0 * P.ny, z = z0;
    while (r >= P.ny) { r -= P.ny; ++z; }
    return (line0 + d) * P.rp + z * P.pad;
}

// ---------------------------------------------------------------------------------------
// Radial (z) kernels: one thread per boundary-plane column.
// ---------------------------------------------------------------------------------------
struct ZOps {
    const double* Lz;     // nz*nz row-major (row k, column m)
    const double* Lz1;    // (nz-1)*(nz-1) row-major
    const double* hf;     // 0.5*f(z)           (F1 term 0.5*(1-(zz/zh)^3))
    const double* cd;     // -0.5*3*zz^2/zh^3   (F1 coefficient of d_z Psi)
    const double* cp;     // -0.5*zz            (F1 coefficient of Psi)
    const double* f1;     // 1-(zz/zh)^3         (F3/F4)
    const double* f2;     // 1-zz^3              (F2)
};

struct ZStage {
    int64_t nxy; int32_t nz; int32_t stage; int32_t mode; // mode 0: RK stage, 1: write F1..F4
    int64_t ps;                               // plane stride of the volume fields (elements)
    double h;
    ZOps zo;
    const void *sR, *sI, *sAx, *sAy;          // stage state (TS)
    const double* At;                         // frozen A_t (its d_z is formed in the kernel)
    const void *dxAt, *dyAt;                  // frozen x/y derivatives of A_t (TS)
    const void *dxR, *dxI, *dyR, *dyI, *lapR, *lapI, *divA, *curl3, *curl4;  // (TS)
    void *zR, *zI, *zAx, *zAy;                // step-start state (TS), updated at stage 4
    void *aR, *aI, *aAx, *aAy;                // RK accumulators (TS)
    double *rho, *arho;                       // rho [ps*nz], boundary-flux accumulator [nxy]
    void *nR, *nI, *nAx, *nAy;                // next stage state (TS)
    double *F1R, *F1I, *F2, *F3, *F4;         // mode 1 outputs
    int32_t* flags;
};

struct ZPost {
    int64_t nxy; int32_t nz; double mu;
    int64_t ps;                   // plane stride of the volume fields (elements)
    ZOps zo;
    void *R, *I, *Ax, *Ay;        // state (TS): read Psi, reset z=0 rows
    const void *divA, *divB;      // d_x A_x and d_y A_y of the updated state (TS); summed here
    const double* rho;            // [ps*nz]
    double* At;                   // out
    const double *axb, *ayb;      // [nxy] boundary values
    double* O;                    // [2*nxy] boundary order parameter (re, im)
    int32_t* flags;
};

struct Wind {
    int32_t nx, ny, nz; int64_t nxy; // nxy: columns of a plane including the row padding (rp * ny)
    int64_t ps;                      // plane stride of the volume fields (elements)
    int64_t rp;                      // row stride (elements); O is [2 * rp * ny] (column (x, y) at x + rp y)
    const double* O; double* part;   // [ny] partial sums
    void *R, *I;                     // Psi (TS)
    const double* yph;               // [ny] 2*pi*yy/L
    double* W;                       // result
    int32_t* flags;
};

// NaN guard over the state when the filter (which carries the guard) is off.
struct NanCheck {
    const void *R, *I, *Ax, *Ay;     // (TS)
    const double* At;
    int64_t n;
    int32_t* flags;
};

// One stationary Newton problem (NewtonRaphson_vy.m) of a batch; all pointers are device memory.
struct NewtonJob {
    int32_t nz, it_max;
    double zh, tol, vy;
    const double* Lz;    // [nz*nz] row-major pz*Dz
    const double* Lz2;   // [nz*nz] row-major Lz*Lz
    const double* zz;    // [nz] radial nodes
    const double *Z1, *Z2, *vy1;   // [nz] initial guesses
    double *z1, *z2, *VY1;         // [nz] results
    double mu_seed, dmu;           // mu-continuation seed and largest step (NaN seed: one solve)
    double* work;                  // [(3nz)^2] Jacobian / LU factors when they do not fit in LDS
    double* out;                   // [4]: it, err, converged, solves of the continuation
};

// Boundary condensate and plaquette census of one state.
struct Census {
    int64_t nxy; int32_t nx, ny, nz;   // nxy = nx * ny points (O, charge: [nx, ny] without row padding)
    int64_t ps;               // plane stride of the volume fields (elements)
    int64_t rp;               // row stride of the volume fields (elements)
    const void *R, *I;        // Psi (TS), [nx*ny*nz]
    const double* Lzb;        // [nz] last row of Lz (the z=0 row)
    double* O;                // [2*nx*ny] interleaved complex
    int8_t* charge;           // [nx*ny] plaquette charge, or NULL
    int32_t* counts;          // [2]: number of +1 and -1 plaquettes (zeroed by the caller)
    double* part;             // [2*census_blocks]: per-block max|O|, sum|O|
};

// Launchers.  All return the hipError_t of the launch.  d: descriptor array in device memory
// (one per member), M: members, blocks: total grid.x over the members.
// Line passes: grid.y = ngroups (members with fewer groups skip the extra rows); lds bytes
// must cover every member; Q = Mp/64 is shared by the members of one launch.
// kind: 0 RHS stage, 1 constraint, 2 filter (separate instantiations for profiling).
// fold: the even/odd folded kernels (descriptors built with fold = 1 or 2; Q is then QF).
// order (folded kernels only): 0 launch order, 1 XCD-aware block order; results are the same.
// The constraint's folded x- and y-passes (one group each, 16 lines per work-group, the same Q)
// in one launch of nb0 + nb1 work-groups; one context.
hipError_t launch_cons_pair(const LPass* d0, const LPass* d1, int nb0, int nb1, size_t lds, int Q, int fp32,
                            hipStream_t s, int order0, int order1);
// The same pair in the factored-DFT form (both passes of factorization fid, descriptors with fold = 3).
hipError_t launch_cons_pair_spec(const LPass* d0, const LPass* d1, int nb0, int nb1, size_t lds, int fid, int fp32,
                                 hipStream_t s, int order0, int order1);
hipError_t launch_line_pass(const LPass* d, Members M, int blocks, int ngroups, size_t lds, int Q, int nt, int dir,
                            int fp32, int kind, hipStream_t s, int fold = 0, int order = 0);
size_t line_pass_lds_bytes(const LPass& P, int fp32);
// Radial kernels: grid.x per member = zblocks(nxy); nz is shared by the members of one launch.
inline int zblocks(int64_t nxy) { return (int)((nxy + 63) / 64); }
// Plane stride of the volume fields for nxy columns: a multiple of 32 elements (128 bytes in FP32).
inline int64_t plane_stride(int64_t nxy) { return (nxy + 31) & ~(int64_t)31; }
// Row stride of the volume fields for rows of nx points: a multiple of 16 elements.
inline int64_t row_stride(int64_t nx) { return (nx + 15) & ~(int64_t)15; }
// wpe: 2 or 3 waves per SIMD targeted by the register allocation (see k_zstage).
hipError_t launch_zstage(const ZStage* d, Members M, int blocks, int nz, int fp32, int wpe, hipStream_t s);
hipError_t launch_zpost(const ZPost* d, Members M, int blocks, int nz, int fp32, hipStream_t s);
// Winding: partial sums (grid.x per member = ny), then the removal (wind_apply_blocks per member).
int wind_apply_blocks(int64_t nxy, int nz);
hipError_t launch_wind_partial(const Wind* d, Members M, int blocks, hipStream_t s);
hipError_t launch_wind_apply(const Wind* d, Members M, int blocks, int fp32, hipStream_t s);
int nancheck_blocks(int64_t n);
hipError_t launch_nancheck(const NanCheck* d, Members M, int blocks, int fp32, hipStream_t s);
bool zkernels_support_nz(int nz);
// Plain streaming copy of `bytes` (multiple of 32) used to calibrate traffic counters.
hipError_t launch_copy_calib(const void* src, void* dst, int64_t bytes, hipStream_t s);
// The Fourier filter for any mask (hqt_ffmask.hip): re/im [nx, ny, nslices] (im NULL: real), mask 0/1 bytes
// [nx, ny] per slice at stride mstride (0: one mask for every slice), twiddle tables twx, twy (cos, sin of
// 2 pi m / N, m < N), work buffers a, b of nx ny nslices complex values, output w.
hipError_t launch_ff_mask(const double* re, const double* im, const uint8_t* mask, int64_t mstride, const double* twx,
                          const double* twy, double2* a, double2* b, double* w, int nx, int ny, int nslices,
                          hipStream_t s);
// Batched Newton (one work-group per problem); every job of one launch has the same nz <= 64.
// njobs problems of radial size at most maxnz; lds: every Jacobian in LDS (each nz resident), else in
// NewtonJob::work.
hipError_t launch_newton(const NewtonJob* jobs, int njobs, int maxnz, bool lds, hipStream_t s);
bool newton_lds_resident(int nz);   // Jacobian held in LDS (else in NewtonJob::work)
// O from the state, then the census of O.
hipError_t launch_order_param(const Census& A, int fp32, hipStream_t s);
hipError_t launch_census(const Census& A, hipStream_t s);
int census_blocks(int64_t nxy);

}  // namespace hqt
