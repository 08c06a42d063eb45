// Line-operator passes: dispatch over the (precision, direction) families whose kernels are
// compiled in hqt_line_*.hip (the kernel itself is in hqt_line_core.h).
#include "hqt_fold_core.h"
#include "hqt_spec1_core.h"

namespace hqt {

extern template hipError_t line_launch_td<double, 0>(const LPass*, Members, dim3, int, int, int, size_t, hipStream_t);
extern template hipError_t line_launch_td<double, 1>(const LPass*, Members, dim3, int, int, int, size_t, hipStream_t);
extern template hipError_t line_launch_td<float, 0>(const LPass*, Members, dim3, int, int, int, size_t, hipStream_t);
extern template hipError_t line_launch_td<float, 1>(const LPass*, Members, dim3, int, int, int, size_t, hipStream_t);
extern template hipError_t fold_launch_td<double, 0>(const LPass*, Members, dim3, int, int, int, size_t, hipStream_t, int, int);
extern template hipError_t fold_launch_td<double, 1>(const LPass*, Members, dim3, int, int, int, size_t, hipStream_t, int, int);
extern template hipError_t fold_launch_td<float, 0>(const LPass*, Members, dim3, int, int, int, size_t, hipStream_t, int, int);
extern template hipError_t fold_launch_td<float, 1>(const LPass*, Members, dim3, int, int, int, size_t, hipStream_t, int, int);
extern template hipError_t spec1_launch_td<double, 0, true>(const LPass*, Members, dim3, int, size_t, hipStream_t, int, int);
extern template hipError_t spec1_launch_td<double, 1, true>(const LPass*, Members, dim3, int, size_t, hipStream_t, int, int);
extern template hipError_t spec1_launch_td<float, 0, true>(const LPass*, Members, dim3, int, size_t, hipStream_t, int, int);
extern template hipError_t spec1_launch_td<float, 1, true>(const LPass*, Members, dim3, int, size_t, hipStream_t, int, int);
extern template hipError_t spec1_launch_td<double, 0, false>(const LPass*, Members, dim3, int, size_t, hipStream_t, int, int);
extern template hipError_t spec1_launch_td<double, 1, false>(const LPass*, Members, dim3, int, size_t, hipStream_t, int, int);
extern template hipError_t spec1_launch_td<float, 0, false>(const LPass*, Members, dim3, int, size_t, hipStream_t, int, int);
extern template hipError_t spec1_launch_td<float, 1, false>(const LPass*, Members, dim3, int, size_t, hipStream_t, int, int);
extern template hipError_t fold_pair_launch_cons<double>(const LPass*, const LPass*, int, int, int, size_t, hipStream_t, int, int);
extern template hipError_t fold_pair_launch_cons<float>(const LPass*, const LPass*, int, int, int, size_t, hipStream_t, int, int);

// A factored-DFT pass is the filter's (the filter's real part, or the cross term) or an RK-stage /
// constraint pass of derivative terms with one output per group (hqt_spec1_core.h).
bool spec_filter_pass(const LPass& P) {
    for (int g = 0; g < P.ngroups; ++g)
        for (int o = 0; o < P.g[g].nout; ++o) {
            if (P.g[g].out[o].gmode != 0) return true;
            for (int e = 0; e < P.g[g].out[o].nterms; ++e)
                if (P.g[g].out[o].term[e].sk == 4 || P.g[g].out[o].term[e].lr) return true;
        }
    return false;
}

size_t line_pass_lds_bytes(const LPass& P, int fp32) {
    if (P.fold == 3 || P.fold == 4) {
        const int fid = P.Mp / 64 - 1;
        const int es = fp32 ? (int)sizeof(float) : (int)sizeof(double);
        // the RK-stage and constraint passes (hqt_spec1_core.h): one line region (lb x S)
        if (!spec_filter_pass(P)) return (size_t)es * (size_t)P.lb * P.S;
        // the filter: the line region and the cross-term area (G, its partials and the lines' coefficients)
        (void)fid;
        return (size_t)spec1_cross_offset(P.S, es) + (size_t)spec_cross_bytes(es);
    }
    int maxslots = 0;
    for (int g = 0; g < P.ngroups; ++g) maxslots = P.g[g].nslots > maxslots ? P.g[g].nslots : maxslots;
    const size_t es = fp32 ? sizeof(float) : sizeof(double);
    // folded: f and g rows, and each source's per-line reference values
    const size_t lines = P.fold ? 2 * (size_t)P.lb * P.S + 2 * (size_t)P.lb : (size_t)P.lb * P.S;
    // + rank-4 dots; folded: + the 4 x 4 cross-term matrix and its 4 per-wave partials (80 doubles,
    // 8-byte aligned)
    return es * (lines + (size_t)maxslots * P.TL + 6 * 4 * (size_t)P.lb) + (P.fold ? 8 + 80 * sizeof(double) : 0);
}

hipError_t launch_line_pass(const LPass* d, Members M, int blocks, int ngroups, size_t lds, int Q, int nt, int dir,
                            int fp32, int kind, hipStream_t s, int fold, int order) {
    if (Q < 1 || Q > 8 || blocks <= 0 || (nt != 1 && nt != 2) || lds > 160 * 1024) return hipErrorInvalidValue;
    const dim3 grid((unsigned)blocks, (unsigned)ngroups);
    if (fold == 3 || fold == 4) {
        if (nt != 1 || Q > kSpecNF) return hipErrorInvalidValue;
        const int wide = fold == 4;
        if (kind == 2) {           // the filter
            if (fp32) return dir ? spec1_launch_td<float, 1, true>(d, M, grid, Q - 1, lds, s, order, wide)
                                 : spec1_launch_td<float, 0, true>(d, M, grid, Q - 1, lds, s, order, wide);
            return dir ? spec1_launch_td<double, 1, true>(d, M, grid, Q - 1, lds, s, order, wide)
                       : spec1_launch_td<double, 0, true>(d, M, grid, Q - 1, lds, s, order, wide);
        }
        // the RK-stage and constraint passes
        if (fp32) return dir ? spec1_launch_td<float, 1, false>(d, M, grid, Q - 1, lds, s, order, wide)
                             : spec1_launch_td<float, 0, false>(d, M, grid, Q - 1, lds, s, order, wide);
        return dir ? spec1_launch_td<double, 1, false>(d, M, grid, Q - 1, lds, s, order, wide)
                   : spec1_launch_td<double, 0, false>(d, M, grid, Q - 1, lds, s, order, wide);
    }
    if (fold) {
        if (Q > 5) return hipErrorInvalidValue;
        const int wide = fold == 2;
        if (fp32) return dir ? fold_launch_td<float, 1>(d, M, grid, Q, nt, kind, lds, s, wide, order)
                             : fold_launch_td<float, 0>(d, M, grid, Q, nt, kind, lds, s, wide, order);
        return dir ? fold_launch_td<double, 1>(d, M, grid, Q, nt, kind, lds, s, wide, order)
                   : fold_launch_td<double, 0>(d, M, grid, Q, nt, kind, lds, s, wide, order);
    }
    if (fp32) return dir ? line_launch_td<float, 1>(d, M, grid, Q, nt, kind, lds, s)
                         : line_launch_td<float, 0>(d, M, grid, Q, nt, kind, lds, s);
    return dir ? line_launch_td<double, 1>(d, M, grid, Q, nt, kind, lds, s)
               : line_launch_td<double, 0>(d, M, grid, Q, nt, kind, lds, s);
}

extern template hipError_t spec_pair_launch<double>(const LPass*, const LPass*, int, int, int, size_t, hipStream_t, int,
                                                    int);
extern template hipError_t spec_pair_launch<float>(const LPass*, const LPass*, int, int, int, size_t, hipStream_t, int,
                                                   int);

hipError_t launch_cons_pair_spec(const LPass* d0, const LPass* d1, int nb0, int nb1, size_t lds, int fid, int fp32,
                                 hipStream_t s, int order0, int order1) {
    if (fid < 0 || fid >= kSpecNF || nb0 <= 0 || nb1 <= 0 || lds > 160 * 1024) return hipErrorInvalidValue;
    return fp32 ? spec_pair_launch<float>(d0, d1, nb0, nb1, fid, lds, s, order0, order1)
                : spec_pair_launch<double>(d0, d1, nb0, nb1, fid, lds, s, order0, order1);
}

hipError_t launch_cons_pair(const LPass* d0, const LPass* d1, int nb0, int nb1, size_t lds, int Q, int fp32,
                            hipStream_t s, int order0, int order1) {
    if (Q < 1 || Q > 5 || nb0 <= 0 || nb1 <= 0 || lds > 160 * 1024) return hipErrorInvalidValue;
    return fp32 ? fold_pair_launch_cons<float>(d0, d1, nb0, nb1, Q, lds, s, order0, order1)
                : fold_pair_launch_cons<double>(d0, d1, nb0, nb1, Q, lds, s, order0, order1);
}

}  // namespace hqt
