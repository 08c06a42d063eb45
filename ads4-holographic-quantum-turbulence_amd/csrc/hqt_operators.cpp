// Host-side operator construction (FP64, once per context) and the stationary Newton
// solver.  Restated from the reference's MATLAB sources:
//   third_party/dmsuite/fourdif.m:30-71   -> fourdif_colrow / fourdif_dense
//   third_party/dmsuite/chebdif.m:25-53   -> chebdif
//   tevol_driven_cluster.m:27-31,47-52    -> filter tables, Lx/Ly/Lz scaling, Lz1 = inv(Lz(1:nz-1,1:nz-1))
//   src/solvers/NewtonRaphson_vy.m:8-105  -> newton_vy
#include "hqt_operators.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace hqt {

static const double PI = 3.14159265358979323846;

void fourdif_colrow(int N, int m, std::vector<double>& col, std::vector<double>& row) {
    col.assign(N, 0.0);
    row.assign(N, 0.0);
    const double h = 2 * PI / N;
    const int n2 = N / 2;  // ceil((N-1)/2)
    auto sgn = [](int k) { return (k % 2) ? -1.0 : 1.0; };
    if (m == 0) {
        col[0] = 1.0;
        row = col;
    } else if (m == 1) {
        // topc = cot(k h/2) (even N) or csc(k h/2) (odd N), k = 1..n2; col1 = [0; 0.5*(-1)^k*[topc; -+flip(topc(1:n1))]]
        std::vector<double> topc(n2 + 1);
        for (int k = 1; k <= n2; ++k) {
            const double t = k * h / 2;
            topc[k] = (N % 2 == 0) ? 1.0 / std::tan(t) : 1.0 / std::sin(t);
        }
        for (int k = 1; k <= N - 1; ++k) {
            double v;
            if (k <= n2) v = topc[k];
            else {
                const int kk = N - k;  // flipud(topc(1:n1)): entry n2+p -> topc(n1+1-p)
                v = (N % 2 == 0) ? -topc[kk] : topc[kk];
            }
            col[k] = 0.5 * sgn(k) * v;
        }
        for (int k = 0; k < N; ++k) row[k] = -col[k];
    } else if (m == 2) {
        std::vector<double> topc(n2 + 1);
        for (int k = 1; k <= n2; ++k) {
            const double t = k * h / 2;
            topc[k] = (N % 2 == 0) ? std::pow(1.0 / std::sin(t), 2) : (1.0 / std::sin(t)) * (1.0 / std::tan(t));
        }
        col[0] = (N % 2 == 0) ? -PI * PI / 3 / (h * h) - 1.0 / 6 : -PI * PI / 3 / (h * h) + 1.0 / 12;
        for (int k = 1; k <= N - 1; ++k) {
            double v;
            if (k <= n2) v = topc[k];
            else {
                const int kk = N - k;
                v = (N % 2 == 0) ? topc[kk] : -topc[kk];
            }
            col[k] = -0.5 * sgn(k) * v;
        }
        row = col;
    } else {
        // col1 = real(ifft((mwave.^m) .* fft([1 0 ... 0]))), mwave = i*[0..N1, N2, -N1..-1]
        const int N1 = (N - 1) / 2;
        std::vector<long double> wave(N);
        for (int k = 0; k <= N1; ++k) wave[k] = k;
        if (N % 2 == 0) wave[N1 + 1] = (-N / 2.0L) * ((m + 1) % 2);
        for (int k = 1; k <= N1; ++k) wave[N - k] = -k;
        for (int j = 0; j < N; ++j) {
            long double s = 0;
            for (int k = 0; k < N; ++k) {
                // (i w)^m e^{2 pi i jk/N}: real part
                const long double mag = std::pow(wave[k], (long double)m);
                const long double ph = (m % 4) * PI / 2 + 2 * PI * (long double)(((long long)j * k) % N) / N;
                s += mag * std::cos(ph);
            }
            col[j] = (double)(s / N);
        }
        if (m % 2 == 0) row = col;
        else {
            col[0] = 0.0;
            for (int k = 0; k < N; ++k) row[k] = -col[k];
        }
    }
}

void fourdif_dense(int N, int m, double* x, double* DM) {
    std::vector<double> col, row;
    fourdif_colrow(N, m, col, row);
    for (int i = 0; i < N; ++i) x[i] = 2 * PI * i / N;
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) DM[i + (size_t)N * j] = (i >= j) ? col[i - j] : row[j - i];
}

void chebdif(int N, int M, std::vector<double>& x, std::vector<double>& DM) {
    const int n1 = N / 2;
    x.resize(N);
    std::vector<double> th(N), DX((size_t)N * N), C((size_t)N * N), Z((size_t)N * N), D((size_t)N * N, 0.0);
    for (int k = 0; k < N; ++k) th[k] = k * PI / (N - 1);
    for (int k = 0; k < N; ++k) x[k] = std::sin(PI * (double)(N - 1 - 2 * k) / (2.0 * (N - 1)));
    // DX(i,j) = 2 sin((th_j + th_i)/2) sin((th_j - th_i)/2): T = repmat(th/2,1,N), DX = 2*sin(T'+T).*sin(T'-T)
    auto dx = [&](int i, int j) { return 2 * std::sin(th[j] / 2 + th[i] / 2) * std::sin(th[j] / 2 - th[i] / 2); };
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double v;
            if (i < n1) v = dx(i, j);
            else v = -dx(N - 1 - i, N - 1 - j);   // -flipud(fliplr(DX(1:n2,:)))
            DX[(size_t)i * N + j] = v;
        }
    for (int i = 0; i < N; ++i) DX[(size_t)i * N + i] = 1.0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double c = ((i - j) % 2 == 0) ? 1.0 : -1.0;   // toeplitz((-1).^k)
            if (i == 0 || i == N - 1) c *= 2;
            if (j == 0 || j == N - 1) c /= 2;
            C[(size_t)i * N + j] = c;
        }
    for (size_t t = 0; t < (size_t)N * N; ++t) Z[t] = 1.0 / DX[t];
    for (int i = 0; i < N; ++i) Z[(size_t)i * N + i] = 0.0;
    for (int i = 0; i < N; ++i) D[(size_t)i * N + i] = 1.0;
    DM.assign((size_t)N * N * M, 0.0);
    std::vector<double> Dn((size_t)N * N);
    for (int ell = 1; ell <= M; ++ell) {
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j)
                Dn[(size_t)i * N + j] = ell * Z[(size_t)i * N + j] * (C[(size_t)i * N + j] * D[(size_t)i * N + i] - D[(size_t)i * N + j]);
        for (int i = 0; i < N; ++i) {          // D(L) = -sum(D'): row sums of the off-diagonal entries
            double s = 0.0;
            for (int j = 0; j < N; ++j)
                if (j != i) s += Dn[(size_t)i * N + j];
            Dn[(size_t)i * N + i] = -s;
        }
        D = Dn;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) DM[i + (size_t)N * j + (size_t)N * N * (ell - 1)] = D[(size_t)i * N + j];
    }
}

bool lu_solve(int n, std::vector<double> A, double* B, int nrhs) {
    // Gaussian elimination with partial pivoting (the algorithm of MATLAB's mldivide for a
    // square dense matrix); A row-major n*n, B row-major n*nrhs, overwritten with the solution.
    std::vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::fabs(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[(size_t)i * n + k]) > best) { best = std::fabs(A[(size_t)i * n + k]); p = i; }
        if (best == 0.0 || !std::isfinite(best)) return false;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
            for (int j = 0; j < nrhs; ++j) std::swap(B[(size_t)k * nrhs + j], B[(size_t)p * nrhs + j]);
        }
        for (int i = k + 1; i < n; ++i) {
            const double l = A[(size_t)i * n + k] / A[(size_t)k * n + k];
            A[(size_t)i * n + k] = l;
            for (int j = k + 1; j < n; ++j) A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
            for (int j = 0; j < nrhs; ++j) B[(size_t)i * nrhs + j] -= l * B[(size_t)k * nrhs + j];
        }
    }
    for (int k = n - 1; k >= 0; --k)
        for (int j = 0; j < nrhs; ++j) {
            double s = B[(size_t)k * nrhs + j];
            for (int i = k + 1; i < n; ++i) s -= A[(size_t)k * n + i] * B[(size_t)i * nrhs + j];
            B[(size_t)k * nrhs + j] = s / A[(size_t)k * n + k];
        }
    return true;
}

void matlab_masks(int n, std::vector<uint8_t>& f) {
    // fx(round(round(n/3)):round(n/3*2)) = 1, MATLAB round = half away from zero
    auto mround = [](double v) { return (double)std::floor(std::fabs(v) + 0.5) * (v < 0 ? -1 : 1); };
    f.assign(n, 0);
    const int a = (int)mround(mround(n / 3.0));
    const int b = (int)mround(n / 3.0 * 2);
    for (int i = a; i <= b; ++i)
        if (i >= 1 && i <= n) f[i - 1] = 1;
}

void filter_tables(int N, const uint8_t* mask, std::vector<double>& pr, std::vector<double>& pi) {
    // P = F^{-1} diag(1 - mask) F, a complex circulant with first column
    // p[m] = delta_m0 - (1/N) sum_{k masked} exp(2 pi i k m / N)  (ff.m:7-11, real part taken later)
    pr.assign(N, 0.0);
    pi.assign(N, 0.0);
    for (int m = 0; m < N; ++m) {
        long double sr = 0, si = 0;
        for (int k = 0; k < N; ++k) {
            if (!mask[k]) continue;
            const long double ang = 2 * (long double)PI * (long double)(((long long)k * m) % N) / N;
            sr += std::cos(ang);
            si += std::sin(ang);
        }
        pr[m] = (double)((m == 0 ? 1.0L : 0.0L) - sr / N);
        pi[m] = (double)(-si / N);
    }
}

bool filter_lowrank(int N, const uint8_t* mask, const LineGeom& g, double scale, std::vector<double>& tab) {
    // Im P[d] = -(1/N) sum_{k masked} sin(2 pi k d / N): a masked k whose conjugate N-k is also
    // masked cancels; an asymmetric pair leaves c sin(theta d), c = -(mask[k] - mask[N-k]) / N,
    // and sin(theta (i - j)) = sin(theta i) cos(theta j) - cos(theta i) sin(theta j).
    std::vector<int> ks;
    std::vector<long double> cs;
    for (int k = 1; 2 * k < N; ++k) {
        const int a = mask[k] ? 1 : 0, b = mask[N - k] ? 1 : 0;
        if (a != b) { ks.push_back(k); cs.push_back(-(long double)(a - b) / N); }
    }
    if (ks.size() > 2) return false;
    tab.assign(4 * (size_t)g.Np + 4 * (size_t)g.Mp, 0.0);
    for (size_t p = 0; p < ks.size(); ++p) {
        for (int j = 0; j < N; ++j) {
            const long double th = 2 * (long double)PI * (long double)(((long long)ks[p] * j) % N) / N;
            tab[4 * (size_t)j + 2 * p] = (double)std::cos(th);                                        // V
            tab[4 * (size_t)j + 2 * p + 1] = (double)std::sin(th);
            tab[4 * (size_t)g.Np + 4 * (size_t)j + 2 * p] = (double)(scale * cs[p] * std::sin(th));      // U
            tab[4 * (size_t)g.Np + 4 * (size_t)j + 2 * p + 1] = (double)(-scale * cs[p] * std::cos(th));
        }
    }
    return true;
}

void circulant_square(const std::vector<double>& c, std::vector<double>& c2) {
    const int N = (int)c.size();
    c2.assign(N, 0.0);
    for (int m = 0; m < N; ++m) {
        long double s = 0;
        for (int p = 0; p < N; ++p) s += (long double)c[p] * (long double)c[((m - p) % N + N) % N];
        c2[m] = (double)s;
    }
}

LineGeom line_geom(int N) {
    LineGeom g;
    g.N = N;
    g.Np = 16 * ((N + 15) / 16);
    g.Mp = 64 * ((N + 63) / 64);
    g.TL = 4 * ((g.Np + g.Mp + 4 + 3) / 4);
    return g;
}

FoldGeom fold_geom(int N, int es) {
    FoldGeom g;
    g.N = N;
    g.R = N / 2 + 1;
    g.Kp = 8 * ((g.R + 7) / 8);
    g.QF = ((g.R + 15) / 16 + 3) / 4;
    g.Mp = 64 * g.QF;
    g.Z = g.Kp + 4;
    g.TL = 4 * ((g.Z + g.Mp + g.Kp + 4 + 3) / 4);
    // Row stride of the folded lines in LDS, at least Kp + 4.  A k-step
    // reads element (line ln, k lk) in lane ln + 16 lk: with S = 2 (mod 32) doubles or 4 (mod 64)
    // floats, each group of lanes served together (32 of 8 bytes or 64 of 4 bytes over 64 banks of
    // 4 bytes) hits distinct banks.
    const int m = es == 8 ? 32 : 64, r = es == 8 ? 2 : 4;
    g.S = r + m * ((g.Kp + 4 - r + m - 1) / m);
    return g;
}

void fold_table(const std::vector<double>& c, int sym, const FoldGeom& g, double scale, std::vector<double>& tab) {
    const int N = g.N;
    tab.assign((size_t)g.TL, 0.0);
    for (int t = 0; t < g.TL; ++t) {
        const int d = ((t - g.Z) % N + N) % N;
        const double v = 2 * d <= N ? c[d] : sym * c[N - d];
        tab[t] = 0.5 * scale * v;
    }
}

void fold_lowrank(const std::vector<double>& dense, const LineGeom& lg, const FoldGeom& g, std::vector<double>& tab) {
    const int N = g.N;
    const double* V = dense.data();
    const double* U = dense.data() + 4 * (size_t)lg.Np;
    tab.assign(8 * (size_t)g.Kp + 8 * (size_t)g.Mp, 0.0);
    double* vf = tab.data();
    double* vg = vf + 4 * (size_t)g.Kp;
    double* ue = vg + 4 * (size_t)g.Kp;
    double* uo = ue + 4 * (size_t)g.Mp;
    for (int j = 0; 2 * j <= N; ++j)
        for (int r = 0; r < 4; ++r) {
            const size_t a = 4 * (size_t)j + r, b = 4 * (size_t)((N - j) % N) + r;
            if (j == 0 || 2 * j == N) {
                vf[a] = V[a];
                ue[a] = U[a];
            } else {
                vf[a] = 0.5 * (V[a] + V[b]);
                vg[a] = 0.5 * (V[a] - V[b]);
                ue[a] = 0.5 * (U[a] + U[b]);
                uo[a] = 0.5 * (U[a] - U[b]);
            }
        }
}

void toeplitz_table(const std::vector<double>& col, const std::vector<double>& row, const LineGeom& g, double scale,
                    std::vector<double>& tab) {
    const int N = (int)col.size();
    tab.assign((size_t)g.TL, 0.0);
    for (int t = 0; t < g.TL; ++t) {
        const int d = t - g.Np - 4;
        if (d >= 0 && d < N) tab[t] = scale * col[d];
        else if (d < 0 && -d < N) tab[t] = scale * row[-d];
    }
}

// ---------------------------------------------------------------------------------------
// NewtonRaphson_vy (src/solvers/NewtonRaphson_vy.m), FP64 on the host.
// ---------------------------------------------------------------------------------------
int newton_vy(int nz, double zh, double small, double D, const double* Z1in, const double* Z2in,
              const double* vy1in, double vy, double* z1o, double* z2o, double* VY1o, int* ito, double* erro,
              int it_max, double tol) {
    std::vector<double> zc, DM;
    chebdif(nz, 1, zc, DM);
    const double zi = D, zf = 1 - small, dlz = zf - zi, pz = 2 / dlz;
    std::vector<double> zz(nz), Lz((size_t)nz * nz), Lz2((size_t)nz * nz, 0.0), f(nz), df(nz);
    for (int i = 0; i < nz; ++i) zz[i] = (zc[i] - 1) * dlz / 2 + zf * 1.0;
    for (int i = 0; i < nz; ++i)
        for (int j = 0; j < nz; ++j) Lz[(size_t)i * nz + j] = pz * DM[i + (size_t)nz * j];
    for (int i = 0; i < nz; ++i)
        for (int j = 0; j < nz; ++j) {
            double s = 0;
            for (int k = 0; k < nz; ++k) s += Lz[(size_t)i * nz + k] * Lz[(size_t)k * nz + j];
            Lz2[(size_t)i * nz + j] = s;
        }
    for (int i = 0; i < nz; ++i) {
        f[i] = 1 - std::pow(zz[i] / zh, 3);
        df[i] = -3 * zz[i] * zz[i] / std::pow(zh, 3);
    }
    auto mv = [&](const std::vector<double>& M, const std::vector<double>& v) {
        std::vector<double> r(nz, 0.0);
        for (int i = 0; i < nz; ++i) {
            double s = 0;
            for (int j = 0; j < nz; ++j) s += M[(size_t)i * nz + j] * v[j];
            r[i] = s;
        }
        return r;
    };
    std::vector<double> Z1(Z1in, Z1in + nz), Z2(Z2in, Z2in + nz), vy1(vy1in, vy1in + nz), Vy1(nz), Z2_f(nz, 0.0);
    auto z2f = [&](const std::vector<double>& Z2v, std::vector<double>& out) {
        const std::vector<double> p = mv(Lz, Z2v);
        for (int i = 1; i < nz; ++i) out[i] = Z2v[i] / (1 - std::pow(zz[i] / zh, 3));
        out[0] = -p[0] / 3;
    };
    z2f(Z2, Z2_f);
    const double vlast = vy1[nz - 1];
    for (int i = 0; i < nz; ++i) Vy1[i] = vy + (vy1[i] - vlast);
    auto fun = [&](const std::vector<double>& a, const std::vector<double>& b, const std::vector<double>& v,
                   const std::vector<double>& bf) {
        std::vector<double> F(3 * (size_t)nz);
        const std::vector<double> a2 = mv(Lz2, a), a1 = mv(Lz, a), b2 = mv(Lz2, b), v2 = mv(Lz2, v), v1 = mv(Lz, v);
        for (int i = 0; i < nz; ++i) {
            F[i] = f[i] * a2[i] + df[i] * a1[i] - v[i] * v[i] * a[i] + b[i] * bf[i] * a[i] - zz[i] * a[i];
            F[nz + i] = f[i] * b2[i] - 2 * a[i] * a[i] * b[i];
            F[2 * nz + i] = -f[i] * v2[i] - df[i] * v1[i] + 2 * v[i] * a[i] * a[i];
        }
        return F;
    };
    const int kidx[4] = {nz - 1, nz, 2 * nz - 1, 3 * nz - 1};
    const int n3 = 3 * nz;
    std::vector<double> Fun, Fun_next;
    int K = 0;
    double err = NAN;
    bool converged = false;
    while (K < it_max) {
        if (K == 0) {
            Fun = fun(Z1, Z2, Vy1, Z2_f);
            for (int r : kidx) Fun[r] = 0;
        } else {
            Fun = Fun_next;
        }
        std::vector<double> J((size_t)n3 * n3, 0.0);
        for (int i = 0; i < nz; ++i) {
            for (int j = 0; j < nz; ++j) {
                const double lzz = Lz2[(size_t)i * nz + j], lz = Lz[(size_t)i * nz + j];
                J[(size_t)i * n3 + j] = f[i] * lzz + df[i] * lz;
                J[(size_t)(nz + i) * n3 + nz + j] = f[i] * lzz;
                J[(size_t)(2 * nz + i) * n3 + 2 * nz + j] = -f[i] * lzz - df[i] * lz;
            }
            J[(size_t)i * n3 + i] += Z2[i] * Z2_f[i] - Vy1[i] * Vy1[i] - zz[i];
            J[(size_t)i * n3 + nz + i] = 2 * Z2_f[i] * Z1[i];
            J[(size_t)i * n3 + 2 * nz + i] = -2 * Vy1[i] * Z1[i];
            J[(size_t)(nz + i) * n3 + i] = -4 * Z1[i] * Z2[i];
            J[(size_t)(nz + i) * n3 + nz + i] -= 2 * Z1[i] * Z1[i];
            J[(size_t)(2 * nz + i) * n3 + i] = 4 * Vy1[i] * Z1[i];
            J[(size_t)(2 * nz + i) * n3 + 2 * nz + i] += 2 * Z1[i] * Z1[i];
        }
        for (int r : kidx) {   // BCz rows (:50-52): unit entry on the diagonal of each fixed row
            for (int j = 0; j < n3; ++j) J[(size_t)r * n3 + j] = 0.0;
            J[(size_t)r * n3 + r] = 1.0;
        }
        std::vector<double> y(Fun);
        if (!lu_solve(n3, J, y.data(), 1)) { err = NAN; break; }
        for (auto& v : y) v = -v;
        std::vector<double> a(nz), b(nz), v(nz), bf(nz, 0.0);
        for (int i = 0; i < nz; ++i) { a[i] = Z1[i] + y[i]; b[i] = Z2[i] + y[nz + i]; v[i] = Vy1[i] + y[2 * nz + i]; }
        z2f(b, bf);
        Fun_next = fun(a, b, v, bf);
        for (int r : kidx) Fun_next[r] = 0;
        double err2 = NAN;   // max(abs(Fun_next)) skipping NaN, like MATLAB's max
        for (double q : Fun_next) err2 = std::fmax(err2, std::fabs(q));
        Z1 = a; Z2 = b; Vy1 = v; Z2_f = bf;
        err = 0;
        for (double q : y) {
            if (std::isnan(q)) { err = NAN; break; }
            err = std::max(err, std::fabs(q));
        }
        if (std::isnan(err)) break;
        if (err2 < tol) { converged = true; break; }
        ++K;
    }
    std::memcpy(z1o, Z1.data(), nz * sizeof(double));
    std::memcpy(z2o, Z2.data(), nz * sizeof(double));
    std::memcpy(VY1o, Vy1.data(), nz * sizeof(double));
    *ito = K;
    *erro = err;
    return converged ? 0 : 1;
}

}  // namespace hqt
