// Host-side operator construction and the host Newton solver (hqt_operators.cpp).
#pragma once
#include <stdint.h>
#include <vector>

namespace hqt {

// First column / first row of fourdif(N, m) (dmsuite fourdif.m:30-71).
void fourdif_colrow(int N, int m, std::vector<double>& col, std::vector<double>& row);
// x (N) and DM (N*N, column-major) of fourdif(N, m).
void fourdif_dense(int N, int m, double* x, double* DM);
// x (N) and DM (N*N*M, column-major) of chebdif(N, M) (dmsuite chebdif.m:25-53).
void chebdif(int N, int M, std::vector<double>& x, std::vector<double>& DM);
// Solve A X = B in place (A row-major n*n by value, B row-major n*nrhs); false if singular.
bool lu_solve(int n, std::vector<double> A, double* B, int nrhs);
// 0/1 mask fx(round(round(n/3)):round(n/3*2)) = 1 (tevol_driven_cluster.m:27-30).
void matlab_masks(int n, std::vector<uint8_t>& f);
// First column (pr + i pi) of the complex circulant F^-1 diag(1-mask) F.
void filter_tables(int N, const uint8_t* mask, std::vector<double>& pr, std::vector<double>& pi);
// First column of C*C for the circulant C with first column c (long-double accumulation).
void circulant_square(const std::vector<double>& c, std::vector<double>& c2);
// Line-pass geometry: K padded to 16 (Np), outputs padded to 64 (Mp), table length TL.
struct LineGeom { int N, Np, Mp, TL; };
LineGeom line_geom(int N);
// Toeplitz lookup table tab[u] = scale * T[i][j] with u = i - j + Np + 4 (zero outside |i-j| < N).
void toeplitz_table(const std::vector<double>& col, const std::vector<double>& row, const LineGeom& g, double scale,
                    std::vector<double>& tab);
// Folded (even/odd) form of a circulant line operator: rows i = 0 .. N/2 (R of them, in QF
// position tiles of 16 per wave, Mp = 64 QF), folded positions j = 0 .. N/2 padded to Kp (a
// multiple of 8), LDS row stride S, table tab[d + Z] = c[d mod N] / 2 of TL entries.
struct FoldGeom { int N, R, Kp, QF, Mp, Z, TL, S; };
// es: element bytes (8 or 4), which sets the LDS row stride S (see fold_geom).
FoldGeom fold_geom(int N, int es = 8);
// Folded table of the circulant with first column c (C[i][j] = c[(i - j) mod N]) and parity
// sym (+1: c[-k] = c[k], -1: c[-k] = -c[k]): tab[t] = scale * c[(t - Z) mod N] / 2, where c[d]
// for d > N/2 is taken as sym * c[N - d] (the table is exactly even or odd).
void fold_table(const std::vector<double>& c, int sym, const FoldGeom& g, double scale, std::vector<double>& tab);
// Folded form of a rank-4 [V (Np x 4) | U (Mp x 4)] table (filter_lowrank): Vf | Vg (Kp x 4 each),
// UE | UO (Mp x 4 each) with x.V = f.Vf + g.Vg and y_i = E_i + O_i, y_{N-i} = E_i - O_i.
void fold_lowrank(const std::vector<double>& dense, const LineGeom& lg, const FoldGeom& g, std::vector<double>& tab);
// NewtonRaphson_vy; returns 0 when the residual test passed.
int newton_vy(int nz, double zh, double small, double D, const double* Z1, const double* Z2, const double* vy1,
              double vy, double* z1, double* z2, double* VY1, int* it, double* err, int it_max = 1000,
              double tol = 1e-6);

// Rank-4 factorisation of the imaginary part of the filter circulant: Im P[i][j] = sum_k
// U[i][k] V[j][k] when at most two masked modes have an unmasked conjugate (every other masked
// mode pairs with a masked conjugate and cancels).  tab gets Np x 4 V then Mp x 4 U (row-major,
// zero past N), U scaled by `scale`.  Returns false (tab untouched) for more asymmetric pairs.
bool filter_lowrank(int N, const uint8_t* mask, const LineGeom& g, double scale, std::vector<double>& tab);

}  // namespace hqt
