"""CPU oracle for the AdS4 holographic-superfluid RK4 stepper.

TEST INFRASTRUCTURE ONLY.  This module is the checker: a NumPy FP64 restatement
of the reference's MATLAB algorithm, written op-for-op in the reference's own
dense-matrix form.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only to check or time the
product path; nothing in the product (the HIP library and the ``hqt`` package)
imports, links or calls it.

Reference files restated here (paths relative to the reference repository):

* ``third_party/dmsuite/fourdif.m:30-47,71``  -> :func:`fourdif`
* ``third_party/dmsuite/chebdif.m:25-53``     -> :func:`chebdif`
* ``scripts/simulation/tevol_driven_cluster.m:11-77`` (operator set-up,
  ``func_Lx/Ly/Lz/Lzr``, ``F1..F4``, ``Constraint``) -> :class:`Operators`
* ``scripts/simulation/tevol_driven_cluster.m:120-142`` (RK4 loop body, A_t
  constraint, boundary reset, winding removal, filter) -> :func:`step`
* ``src/numerics/ff.m:7-11``                   -> :func:`ff`
* ``src/solvers/NewtonRaphson_vy.m:8-105``      -> :func:`NewtonRaphson_vy`
* ``scripts/legacy/Statinary_state_with_v.m:11-17,50-52`` (continuation in mu
  by rescaled seeds)                            -> :func:`newton_mu_continuation`
* ``src/diagnostics/countN.m:7-21``             -> :func:`countN`
* ``scripts/simulation/tevol_driven_cluster.m:81-102`` and
  ``scripts/simulation/tevol_lattice.m:101-135`` (initial fields) ->
  :func:`stationary_fields`, :func:`imprint`

How each function is pinned (checked against something other than itself) is
listed in ``tests/test_oracle.py``; the FFT-form twin (:class:`Operators` with
``form="fft"``) is an independent second restatement of the x/y operators used
to bound the oracle's own rounding floor.

Array convention: a field is a NumPy array of shape ``(nx, ny, nz)`` indexed
``[i, j, k]`` = MATLAB ``(i+1, j+1, k+1)``; ``k = nz-1`` is the AdS boundary
z = 0 and ``k = 0`` the horizon z = 1 (``docs/DATA_DICTIONARY.md:7-8``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "matlab_round", "fourdif", "chebdif", "filter_masks", "ff", "Operators",
    "Params", "step", "rhs", "countN", "NewtonRaphson_vy", "spline_interp",
    "stationary_fields", "imprint", "boundary_order_parameter",
]


# --------------------------------------------------------------------------------------
# dmsuite
# --------------------------------------------------------------------------------------

def matlab_round(x):
    """MATLAB ``round``: halves away from zero (Python's ``round`` is banker's)."""
    x = np.asarray(x, dtype=np.float64)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def _toeplitz(col, row):
    """MATLAB ``toeplitz(col,row)``: T[i,j] = col[i-j] for i>=j, row[j-i] for j>i."""
    n = len(col)
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    d = i - j
    return np.where(d >= 0, np.asarray(col)[np.clip(d, 0, n - 1)],
                    np.asarray(row)[np.clip(-d, 0, n - 1)])


def fourdif_col_row(N: int, m: int):
    """First column and row of the Fourier differentiation matrix.

    Follows ``third_party/dmsuite/fourdif.m:30-71`` branch by branch (cot/csc
    flip trick for m=1,2; FFT construction for m>2)."""
    h = 2 * math.pi / N
    kk = np.arange(1, N)
    n1 = (N - 1) // 2
    n2 = -(-(N - 1) // 2)  # ceil((N-1)/2)
    sgn = (-1.0) ** kk
    if m == 0:
        col1 = np.zeros(N); col1[0] = 1.0
        row1 = col1.copy()
    elif m == 1:
        t = np.arange(1, n2 + 1) * h / 2
        if N % 2 == 0:
            topc = 1.0 / np.tan(t)
            col1 = np.concatenate(([0.0], 0.5 * sgn * np.concatenate((topc, -topc[:n1][::-1]))))
        else:
            topc = 1.0 / np.sin(t)
            col1 = np.concatenate(([0.0], 0.5 * sgn * np.concatenate((topc, topc[:n1][::-1]))))
        row1 = -col1
    elif m == 2:
        t = np.arange(1, n2 + 1) * h / 2
        if N % 2 == 0:
            topc = (1.0 / np.sin(t)) ** 2
            col1 = np.concatenate(([-math.pi ** 2 / 3 / h ** 2 - 1.0 / 6],
                                   -0.5 * sgn * np.concatenate((topc, topc[:n1][::-1]))))
        else:
            topc = (1.0 / np.sin(t)) * (1.0 / np.tan(t))
            col1 = np.concatenate(([-math.pi ** 2 / 3 / h ** 2 + 1.0 / 12],
                                   -0.5 * sgn * np.concatenate((topc, -topc[:n1][::-1]))))
        row1 = col1.copy()
    else:
        N1 = (N - 1) // 2
        mid = [(-N / 2) * ((m + 1) % 2)] if (N + 1) % 2 else []
        mwave = 1j * np.array(list(range(0, N1 + 1)) + mid + list(range(-N1, 0)), dtype=np.complex128)
        e0 = np.zeros(N); e0[0] = 1.0
        col1 = np.real(np.fft.ifft((mwave ** m) * np.fft.fft(e0)))
        if m % 2 == 0:
            row1 = col1.copy()
        else:
            col1 = np.concatenate(([0.0], col1[1:]))
            row1 = -col1
    return col1, row1


def fourdif(N: int, m: int):
    """``[x, DM] = fourdif(N, m)`` (``third_party/dmsuite/fourdif.m:30-71``)."""
    x = 2 * math.pi * np.arange(N) / N
    col1, row1 = fourdif_col_row(N, m)
    return x, _toeplitz(col1, row1)


def chebdif(N: int, M: int):
    """``[x, DM] = chebdif(N, M)`` (``third_party/dmsuite/chebdif.m:25-53``).

    Returns ``x`` (descending Chebyshev points) and ``DM`` of shape (M, N, N)."""
    I = np.eye(N)
    L = I.astype(bool)
    n1 = N // 2
    n2 = -(-N // 2)
    k = np.arange(N, dtype=np.float64)
    th = k * math.pi / (N - 1)
    x = np.sin(math.pi * np.arange(N - 1, -N, -2, dtype=np.float64) / (2 * (N - 1)))
    T = np.tile((th / 2)[:, None], (1, N))
    DX = 2 * np.sin(T.T + T) * np.sin(T.T - T)
    DX = np.concatenate((DX[:n1, :], -np.flipud(np.fliplr(DX[:n2, :]))), axis=0)
    DX[L] = 1.0
    C = _toeplitz((-1.0) ** k, (-1.0) ** k)
    C[0, :] *= 2; C[N - 1, :] *= 2
    C[:, 0] /= 2; C[:, N - 1] /= 2
    Z = 1.0 / DX
    Z[L] = 0.0
    D = np.eye(N)
    DM = np.zeros((M, N, N))
    for ell in range(1, M + 1):
        D = ell * Z * (C * np.tile(np.diag(D)[:, None], (1, N)) - D)
        D[L] = -np.sum(D.T, axis=0)
        DM[ell - 1] = D
    return x, DM


# --------------------------------------------------------------------------------------
# Filter (src/numerics/ff.m) and its mask (tevol_driven_cluster.m:27-31)
# --------------------------------------------------------------------------------------

def filter_masks(nx: int, ny: int):
    """``fx``/``fy`` of ``tevol_driven_cluster.m:27-30`` as 0-based boolean vectors."""
    fx = np.zeros(nx, dtype=bool)
    a = int(matlab_round(matlab_round(nx / 3))); b = int(matlab_round(nx / 3 * 2))
    fx[a - 1:b] = True
    fy = np.zeros(ny, dtype=bool)
    a = int(matlab_round(matlab_round(ny / 3))); b = int(matlab_round(ny / 3 * 2))
    fy[a - 1:b] = True
    return fx, fy


def ff(x, mask):
    """``w = ff(x, y)`` (``src/numerics/ff.m:7-11``): zero masked 2-D Fourier modes of
    every slice, return the real part.  ``mask`` broadcasts against ``x``'s first two
    axes (a ``[nx,ny]`` or ``[nx,ny,nz]`` boolean array)."""
    x = np.asarray(x)
    b = np.fft.fft2(x, axes=(0, 1))
    m = np.broadcast_to(mask if mask.ndim == x.ndim else mask.reshape(mask.shape + (1,) * (x.ndim - mask.ndim)), b.shape)
    b = np.where(m, 0, b)
    return np.real(np.fft.ifft2(b, axes=(0, 1)))


# --------------------------------------------------------------------------------------
# Operators and right-hand sides (tevol_driven_cluster.m:11-77)
# --------------------------------------------------------------------------------------

class Operators:
    """Grid, operator matrices and closures of ``tevol_driven_cluster.m:11-77``.

    ``Lx0``/``Ly0`` are the periodic box sides (the reference uses
    ``L=L0=(nx-1)/2``, ``:23-24,85``).  ``form="dense"`` applies the x/y derivatives as
    the reference does (dense ``fourdif`` matrix products, ``:60-61``);
    ``form="fft"`` is the independent twin that applies the same operators as FFT
    multipliers (exact eigen-decomposition of the circulant matrices)."""

    def __init__(self, nx: int, ny: int, nz: int, Lx0: float | None = None,
                 Ly0: float | None = None, form: str = "dense"):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.Lx0 = (nx - 1) / 2 if Lx0 is None else float(Lx0)
        self.Ly0 = (ny - 1) / 2 if Ly0 is None else float(Ly0)
        self.form = form
        x, Dx = fourdif(nx, 1)
        y, Dy = fourdif(ny, 1)
        z, Dzm = chebdif(nz, 1)
        Dz = Dzm[0]
        small = 0.0; D = 0.0
        zi = D; zf = 1 - small
        dlz = zf - zi
        px = 2.0 * math.pi / self.Lx0
        py = 2.0 * math.pi / self.Ly0
        pz = 2 / dlz
        self.x, self.y, self.zcheb = x, y, z
        self.xx1 = x * self.Lx0 / 2 / math.pi             # :42 (1-D factor)
        self.yy1 = (y * self.Ly0) / 2 / math.pi           # :43
        self.zz1 = (z - 1) * dlz / 2 + zf                 # :44
        self.Lx = px * Dx                                 # :47
        self.Ly = py * Dy                                 # :48
        self.Lz = pz * Dz                                 # :49
        self.Lz1 = np.linalg.solve(self.Lz[:nz - 1, :nz - 1], np.eye(nz - 1))  # :52
        zh = 1.0
        zz = self.zz1
        self.f1 = 1 - (zz / zh) ** 3                      # :71, (1-(zz/zh).^3)
        self.c_dpsi = -0.5 * 3 * zz ** 2 / zh ** 3        # :72, -0.5.*3.*zz.^2./zh^3
        self.c_psi = -0.5 * zz                            # :72, -0.5.*zz
        self.f2 = 1 - zz ** 3                             # :76, (1-zz.^3)
        self.fx, self.fy = filter_masks(nx, ny)
        self.mask2d = self.fx[:, None] | self.fy[None, :]
        if form == "fft":
            # eigenvalues of the circulant fourdif matrices: i*k with k_{N/2}=0 (even N)
            self.kx = px * self._wavenumbers(nx)
            self.ky = py * self._wavenumbers(ny)

    @staticmethod
    def _wavenumbers(N):
        k = np.fft.fftfreq(N, d=1.0 / N)
        if N % 2 == 0:
            k[N // 2] = 0.0
        return k

    # -- closures func_Lx / func_Ly / func_Lz / func_Lzr (:60-65) ----------------------
    def Lx_(self, X):
        if self.form == "fft":
            return self._fft_deriv(X, 0)
        if np.iscomplexobj(X):
            return self.Lx_(X.real) + 1j * self.Lx_(X.imag)
        nx, ny, nz = X.shape
        return (self.Lx @ X.reshape(nx, ny * nz)).reshape(nx, ny, nz)

    def Ly_(self, X):
        if self.form == "fft":
            return self._fft_deriv(X, 1)
        if np.iscomplexobj(X):
            return self.Ly_(X.real) + 1j * self.Ly_(X.imag)
        nx, ny, nz = X.shape
        P = np.ascontiguousarray(X.transpose(1, 0, 2)).reshape(ny, nx * nz)
        return (self.Ly @ P).reshape(ny, nx, nz).transpose(1, 0, 2)

    def Lz_(self, X):
        if np.iscomplexobj(X):
            return self.Lz_(X.real) + 1j * self.Lz_(X.imag)
        nx, ny, nz = X.shape
        return (X.reshape(nx * ny, nz) @ self.Lz.T).reshape(nx, ny, nz)

    def Lzr_(self, X):
        """``func_Lzr`` (:63-65): rows 1..nz-1 = Lz1 * x(1:nz-1), row nz = 0."""
        if np.iscomplexobj(X):
            return self.Lzr_(X.real) + 1j * self.Lzr_(X.imag)
        nx, ny, nz = X.shape
        W = np.zeros_like(X)
        W[:, :, :nz - 1] = (X[:, :, :nz - 1].reshape(nx * ny, nz - 1) @ self.Lz1.T).reshape(nx, ny, nz - 1)
        return W

    def _fft_deriv(self, X, axis):
        k = self.kx if axis == 0 else self.ky
        shape = [1, 1, 1]; shape[axis] = -1
        mult = (1j * k).reshape(shape)
        out = np.fft.ifft(mult * np.fft.fft(X, axis=axis), axis=axis)
        return out if np.iscomplexobj(X) else np.real(out)

    # -- right-hand sides (:72-77) ------------------------------------------------------
    def F1(self, y1, y2, y3, y4):
        Lx, Ly, Lz, Lzr = self.Lx_, self.Ly_, self.Lz_, self.Lzr_
        f = self.f1[None, None, :]
        cd = self.c_dpsi[None, None, :]
        cp = self.c_psi[None, None, :]
        dpsi = Lz(y1)
        return Lzr(1j * y2 * dpsi + 0.5j * y1 * Lz(y2) + 0.5 * f * Lz(dpsi) + cd * dpsi + cp * y1
                   - 0.5 * y3 * y3 * y1 - 0.5 * y4 * y4 * y1 + 1. * 0.5 * Lx(Lx(y1))
                   + 1. * 0.5 * Ly(Ly(y1)) - 1.0 * 0.5j * y1 * (Lx(y3) + Ly(y4))
                   - 1j * (y3 * Lx(y1) + y4 * Ly(y1)))

    def F3(self, y1, y2, y3, y4):
        Lx, Ly, Lz, Lzr = self.Lx_, self.Ly_, self.Lz_, self.Lzr_
        f = self.f1[None, None, :]
        cy1 = np.conj(y1)
        return np.real(Lzr(0.5 * Lz(Lx(y2) + f * Lz(y3)) + 0.5 * Ly(Ly(y3) - Lx(y4))
                           - 0.5j * (cy1 * Lx(y1) - y1 * Lx(cy1)) - y3 * y1 * cy1))

    def F4(self, y1, y2, y3, y4):
        Lx, Ly, Lz, Lzr = self.Lx_, self.Ly_, self.Lz_, self.Lzr_
        f = self.f1[None, None, :]
        cy1 = np.conj(y1)
        return np.real(Lzr(0.5 * Lz(Ly(y2) + f * Lz(y4)) + 0.5 * Lx(-Ly(y3) + Lx(y4))
                           - 0.5j * (cy1 * Ly(y1) - y1 * Ly(cy1)) - y4 * y1 * cy1))

    def F2(self, y1, y2, y3, y4):
        return self.f2[None, None, :] * self.Lz_(self.Lx_(y3) + self.Ly_(y4))

    def Constraint(self, y1, y2, y3, y4):
        """``Constraint`` (:77): d_z(d_x A_x + d_y A_y - d_z A_t) - 2 Im(conj(Psi) d_z Psi)."""
        return (self.Lz_(self.Lx_(y3) + self.Ly_(y4) - self.Lz_(y2))
                - 2 * np.imag(np.conj(y1) * self.Lz_(y1)))

    def ff(self, X):
        return ff(X, self.mask2d)


# --------------------------------------------------------------------------------------
# One RK4 step (tevol_driven_cluster.m:120-142)
# --------------------------------------------------------------------------------------

@dataclass
class Params:
    """Run parameters of the loop body.  ``ax_b``/``ay_b`` are the z=0 values of A_x and
    A_y restored after every step (``:133`` uses A_x=0 and A_y=-vy, uniform)."""
    mu: float = 6.0
    h: float = 0.02
    ax_b: np.ndarray | float = 0.0
    ay_b: np.ndarray | float = 0.0
    winding: bool = True
    filter: bool = True


def rhs(ops: Operators, Z1, Z2, Z3, Z4):
    """The four closures at one state: (F1, F2, F3, F4) (``:72-76``)."""
    return ops.F1(Z1, Z2, Z3, Z4), ops.F2(Z1, Z2, Z3, Z4), ops.F3(Z1, Z2, Z3, Z4), ops.F4(Z1, Z2, Z3, Z4)


def boundary_order_parameter(ops: Operators, Z1):
    """O = (func_Lz(Z1))(:,:,end) (``:136``, ``scripts/analysis/movie_show.m:91``)."""
    return ops.Lz_(Z1)[:, :, -1]


def winding_number(O):
    """``round(mean(sum(angle(circshift(O,-1,2)./O),1))/(2*pi))`` (``:137``)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.roll(O, -1, axis=1) / O
    return matlab_round(np.mean(np.sum(np.angle(ratio), axis=0)) / (2 * math.pi))


def step(ops: Operators, prm: Params, Z1, Z2, Z3, Z4, rhoi):
    """One pass of the loop body ``tevol_driven_cluster.m:120-142``.

    Returns ``(Z1, Z2, Z3, Z4, rhoi, winding)``.  With ``prm.winding=False`` this is
    the ``tevol_lattice.m:154-173`` body (winding removal commented out there)."""
    h = prm.h
    nz = ops.nz
    F1, F2, F3, F4 = ops.F1, ops.F2, ops.F3, ops.F4

    def stage(y1, y3, y4):
        k1 = h * F1(y1, Z2, y3, y4); k3 = h * F3(y1, Z2, y3, y4); k4 = h * F4(y1, Z2, y3, y4)
        k1[:, :, nz - 1] = 0; k3[:, :, nz - 1] = 0; k4[:, :, nz - 1] = 0
        k2 = h * F2(y1, Z2, y3, y4)
        return k1, k2, k3, k4

    k11, k21, k31, k41 = stage(Z1, Z3, Z4)
    k12, k22, k32, k42 = stage(Z1 + k11 / 2, Z3 + k31 / 2, Z4 + k41 / 2)
    k13, k23, k33, k43 = stage(Z1 + k12 / 2, Z3 + k32 / 2, Z4 + k42 / 2)
    k14, k24, k34, k44 = stage(Z1 + k13, Z3 + k33, Z4 + k43)
    k1 = (1 / 6) * (k11 + 2 * k12 + 2 * k13 + k14)
    k3 = (1 / 6) * (k31 + 2 * k32 + 2 * k33 + k34)
    k4 = (1 / 6) * (k41 + 2 * k42 + 2 * k43 + k44)
    k2 = (1 / 6) * (k21 + 2 * k22 + 2 * k23 + k24)
    Z1, Z3, Z4 = Z1 + k1, Z3 + k3, Z4 + k4
    rhoi = rhoi - np.repeat(k2[:, :, nz - 1:nz], nz, axis=2)
    cons = np.real(ops.Lz_(ops.Lx_(Z3) + ops.Ly_(Z4))
                   + 1j * (np.conj(Z1) * ops.Lz_(Z1) - Z1 * ops.Lz_(np.conj(Z1))))
    Z2 = ops.Lzr_(ops.Lzr_(cons) - rhoi) + prm.mu
    Z1 = Z1.copy(); Z3 = Z3.copy(); Z4 = Z4.copy()
    Z1[:, :, nz - 1] = 0
    Z3[:, :, nz - 1] = prm.ax_b
    Z2[:, :, nz - 1] = prm.mu
    Z4[:, :, nz - 1] = prm.ay_b
    W = 0.0
    if prm.winding:
        O1 = boundary_order_parameter(ops, Z1)
        W = winding_number(O1)
        # Z1.*exp(-1i*2*pi*yy/L*(winding+0)), yy=(yycheb*L)/2/pi (:43,:139)
        phase = np.exp(-1j * 2 * math.pi * ops.yy1 / ops.Ly0 * (W + 0))
        Z1 = Z1 * phase[None, :, None]
    if prm.filter:
        Z1 = ops.ff(np.real(Z1)) + 1j * ops.ff(np.imag(Z1))
        Z2, Z3, Z4 = ops.ff(Z2), ops.ff(Z3), ops.ff(Z4)
    return Z1, Z2, Z3, Z4, rhoi, W


# --------------------------------------------------------------------------------------
# Diagnostics: countN (src/diagnostics/countN.m:7-21)
# --------------------------------------------------------------------------------------

def countN(x1):
    """``[n, np, nm] = countN(x1)``: periodic plaquette windings +1 / -1 (tolerance 1e-4)."""
    x1 = np.asarray(x1)
    x2 = np.roll(x1, (-1, 0), axis=(0, 1))
    x3 = np.roll(x1, (-1, -1), axis=(0, 1))
    x4 = np.roll(x1, (0, -1), axis=(0, 1))
    with np.errstate(divide="ignore", invalid="ignore"):
        w = 1 / (2 * math.pi) * (np.angle(x1 / x2) + np.angle(x2 / x3) + np.angle(x3 / x4) + np.angle(x4 / x1))
    Lp = (w >= 1 - 1e-4) & (w <= 1 + 1e-4)
    Lm = (w <= -1 + 1e-4) & (w >= -1 - 1e-4)
    return int(np.sum(Lp | Lm)), int(np.sum(Lp)), int(np.sum(Lm))


# --------------------------------------------------------------------------------------
# Stationary Newton solver (src/solvers/NewtonRaphson_vy.m:8-105), plots/prints dropped
# --------------------------------------------------------------------------------------

def NewtonRaphson_vy(nz, zh, small, D, Z1, Z2, vy1, vy, it_max=1000, tol=1e-6):
    """``[z1,z2,VY1,it,error] = NewtonRaphson_vy(nz,zh,small,D,Z1,Z2,vy1,vy)``."""
    Z1 = np.array(Z1, dtype=np.float64).reshape(nz)
    Z2 = np.array(Z2, dtype=np.float64).reshape(nz)
    vy1 = np.array(vy1, dtype=np.float64).reshape(nz)
    zzcheb, Dzm = chebdif(nz, 1)
    lz = Dzm[0]
    zi = D; zf = 1 - small
    dlz = zf - zi
    pz = 2 / dlz
    zz = (zzcheb - 1) * dlz / 2 + zf * np.ones(nz)
    z = zz
    Lz = pz * lz
    Lz2 = Lz @ Lz
    f = 1 - (zz / zh) ** 3
    df = -3 * zz ** 2 / zh ** 3
    pZ2 = Lz @ Z2
    Z2_f = np.zeros(nz)
    Z2_f_temp = Z2_f.copy()
    Z2_f[1:] = Z2[1:] / (1 - (zz[1:] / zh) ** 3)
    Z2_f[0] = -pZ2[0] / 3
    vy1 = vy1 - vy1[nz - 1]
    Vy1 = vy + vy1

    def fun(Z1, Z2, Vy1, Z2_f):
        return np.concatenate((f * (Lz2 @ Z1) + df * (Lz @ Z1) - Vy1 ** 2 * Z1 + Z2 * Z2_f * Z1 - z * Z1,
                               f * (Lz2 @ Z2) - 2 * Z1 ** 2 * Z2,
                               -f * (Lz2 @ Vy1) - df * (Lz @ Vy1) + 2 * Vy1 * Z1 ** 2))

    def dfun(Z1, Z2, Vy1, Z2_f):
        Zn = np.zeros((nz, nz))
        return np.block([
            [np.diag(f) @ Lz2 + np.diag(df) @ Lz + np.diag(Z2 * Z2_f - Vy1 ** 2 - z), np.diag(2 * Z2_f * Z1), np.diag(-2 * Vy1 * Z1)],
            [np.diag(-4 * Z1 * Z2), np.diag(f) @ Lz2 - np.diag(2 * Z1 ** 2), Zn],
            [np.diag(4 * Vy1 * Z1), Zn, -np.diag(f) @ Lz2 - np.diag(df) @ Lz + np.diag(2 * Z1 ** 2)]])

    BCz = np.zeros((4, 3 * nz))
    BCz[0, nz - 1] = 1; BCz[1, nz] = 1      # rows 1-2: eye(2) at columns nz, nz+1
    BCz[2, 2 * nz - 1] = 1                   # row 3: column 2nz
    BCz[3, 3 * nz - 1] = 1                   # row 4: column 3nz
    kidx = np.array([nz - 1, nz, 2 * nz - 1, 3 * nz - 1])  # 0-based rows nz:nz+1, 2nz, 3nz
    K = 0
    err = float("nan")
    Fun_next = None
    while K < it_max:
        if K == 0:
            Fun = fun(Z1, Z2, Vy1, Z2_f)
            Fun[kidx] = 0
        else:
            Fun = Fun_next
        dFun = dfun(Z1, Z2, Vy1, Z2_f)
        dFun[kidx, :] = BCz
        y = -np.linalg.solve(dFun, Fun)
        Z1_temp = Z1 + y[:nz]
        Z2_temp = Z2 + y[nz:2 * nz]
        Vy1_temp = Vy1 + y[2 * nz:3 * nz]
        pAt = Lz @ Z2_temp
        Z2_f_temp = Z2_f_temp.copy()
        Z2_f_temp[1:] = Z2_temp[1:] / (1 - (z[1:] / zh) ** 3)
        Z2_f_temp[0] = -pAt[0] / 3
        Fun_next = fun(Z1_temp, Z2_temp, Vy1_temp, Z2_f_temp)
        Fun_next[kidx] = 0
        err2 = np.fmax.reduce(np.abs(Fun_next))   # MATLAB max skips NaN
        Z1, Z2, Vy1, Z2_f = Z1_temp, Z2_temp, Vy1_temp, Z2_f_temp
        err = float(np.max(np.abs(y)))
        if math.isnan(err):
            break
        elif err2 < tol:
            break
        K += 1
    return Z1, Z2, Vy1, K, err


def newton_mu_continuation(nz, zh, small, D, Z1, Z2, vy1, vy, mu_seed, dmu, it_max=1000, tol=1e-6):
    """Continuation in mu by rescaled seeds, the procedure of
    ``scripts/legacy/Statinary_state_with_v.m:11-17,50-52`` (load a solved profile, ``Z2=Z2/6*mu``,
    solve again, step mu) written as a loop.  The target mu is the guess's own A_t(z=0) = ``Z2[-1]``.
    The guess rescaled to A_t(z=0) = ``mu_seed`` is solved first; then mu steps to the target by at
    most ``dmu``, each solve starting from the last solution with Z2 rescaled by m/prev and its
    boundary entry set to m.  A step that does not converge is retried at half its length until the
    length drops below ``dmu/64``.  Every solve is :func:`NewtonRaphson_vy` with
    ``min(it_max, 100)`` iterations; at most 1024 solves.  Returns ``(z1, z2, Vy1, it, err, conv,
    solves)``: the last solve's iterate, ``it`` and ``err``, whether the target was reached, and
    the number of solves."""
    CONT_IT, MAX_SOLVES = 100, 1024
    itc = min(int(it_max), CONT_IT)
    Z1 = np.array(Z1, dtype=np.float64).reshape(nz)
    Z2 = np.array(Z2, dtype=np.float64).reshape(nz)
    V = np.array(vy1, dtype=np.float64).reshape(nz)
    mt = float(Z2[nz - 1])

    def solve(a, b, v, r, m):
        b = b * r
        b[nz - 1] = m
        z1, z2, vv, K, err = NewtonRaphson_vy(nz, zh, small, D, a, b, v, vy, it_max=itc, tol=tol)
        return z1, z2, vv, K, err, (K < itc and not math.isnan(err))

    prev = float(mu_seed)
    z1, z2, vv, K, err, ok = solve(Z1, Z2, V, prev / mt, prev)
    solves = 1
    if not ok:
        return z1, z2, vv, K, err, False, solves
    h = math.copysign(dmu, mt - prev)
    hmin = dmu * (1.0 / 64)
    while prev != mt and solves < MAX_SOLVES:
        C1, C2, CV = z1, z2, vv
        ok = False
        while not ok and solves < MAX_SOLVES:
            m = prev + h
            if (m >= mt) if h > 0 else (m <= mt):
                m = mt
            z1, z2, vv, K, err, ok = solve(C1, C2, CV, m / prev, m)
            solves += 1
            if ok:
                prev = m
            else:
                h *= 0.5
                if abs(h) < hmin:
                    break
        if not ok:
            break
    return z1, z2, vv, K, err, prev == mt, solves


# --------------------------------------------------------------------------------------
# Initial fields (tevol_driven_cluster.m:81-102, tevol_lattice.m:101-135)
# --------------------------------------------------------------------------------------

def spline_interp(X, Y, Xi):
    """MATLAB ``interp1(X, Y, Xi, 'spline')``: not-a-knot cubic spline (any ordering of X)."""
    from scipy.interpolate import CubicSpline
    X = np.asarray(X, dtype=np.float64).ravel()
    Y = np.asarray(Y, dtype=np.float64).ravel()
    order = np.argsort(X)
    return CubicSpline(X[order], Y[order], bc_type="not-a-knot")(np.asarray(Xi, dtype=np.float64))


def stationary_fields(ops: Operators, prof: dict, vy_profile_sign: float = -1.0):
    """Broadcast a stationary radial profile and gauge-rotate it (``:82-98``).

    ``prof`` holds nz0-point vectors ``Z1, Z2, Vy1, rhoi`` on the chebdif(nz0) grid.
    Returns ``(z1, z2, z3, z4, rhoi)`` of shape (nx, ny, nz) before any perturbation."""
    nx, ny, nz = ops.nx, ops.ny, ops.nz
    nz0 = len(np.ravel(prof["Z1"]))
    z0, _ = chebdif(nz0, 1)
    Zg = (z0 - 1) * 1.0 / 2 + 1.0                          # :45
    zz0 = ops.zz1
    Z1 = spline_interp(Zg, prof["Z1"], zz0)
    Z2 = spline_interp(Zg, prof["Z2"], zz0)
    Vy1 = spline_interp(Zg, prof["Vy1"], zz0)
    rh = spline_interp(Zg, prof["rhoi"], zz0)
    shp = (nx, ny, nz)
    z1 = np.broadcast_to(Z1, shp).copy()
    z2 = np.broadcast_to(Z2, shp).copy()
    vy1 = np.broadcast_to(Vy1, shp).copy()
    rho = np.broadcast_to(rh, shp).copy()
    zz = np.broadcast_to(ops.zz1, shp)
    pZ2 = ops.Lz_(z2)
    Z2_f = np.zeros(shp)
    Z2_f[:, :, 1:] = z2[:, :, 1:] / (1 - (zz[:, :, 1:] / 1.0) ** 3)
    Z2_f[:, :, 0] = -pZ2[:, :, 0] / 3
    theta = -ops.Lzr_(Z2_f)
    z4 = vy_profile_sign * vy1
    z1 = z1 * np.exp(1j * theta)
    z3 = ops.Lx_(theta)
    rhoi = np.real(np.repeat(rho[:, :, nz - 1:nz], nz, axis=2))
    return z1, z2, z3, z4, rhoi


def imprint(ops: Operators, z1, phase_xy=None, amp_xy=None):
    """Multiply the scalar by a boundary-plane pattern: ``z1 .* amp .* exp(1i*phase)``
    (the form of ``tevol_driven_cluster.m:101`` and ``tevol_lattice.m:135``)."""
    out = z1
    if amp_xy is not None:
        out = out * amp_xy[:, :, None]
    if phase_xy is not None:
        out = out * np.exp(1j * phase_xy)[:, :, None]
    return out


# --------------------------------------------------------------------------------------
# Device initial state (tevol_driven_cluster.m:81-102 with the plane pattern of :101 and
# tevol_lattice.m:113-135): restatement of the product's documented construction
# --------------------------------------------------------------------------------------

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
_MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) on uint32 counter arrays."""
    c = [np.asarray(x, dtype=np.uint64) & _MASK32 for x in (c0, c1, c2, c3)]
    k0 = np.uint64(k0) & _MASK32
    k1 = np.uint64(k1) & _MASK32
    for _ in range(10):
        p0 = _M0 * c[0]
        p1 = _M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK32
        c = [(hi1 ^ c[1] ^ k0) & _MASK32, lo1, (hi0 ^ c[3] ^ k1) & _MASK32, lo0]
        k0 = (k0 + _W0) & _MASK32
        k1 = (k1 + _W1) & _MASK32
    return c


def column_normals(nx, ny, seed):
    """(g1, g2) per column i + nx*j: Philox keyed by the seed, two 53-bit uniforms, Box-Muller."""
    col = np.arange(nx * ny, dtype=np.uint64)
    seed = int(seed)
    u = philox4x32_10(col & _MASK32, col >> np.uint64(32), 0, 0, seed & 0xFFFFFFFF, seed >> 32)
    r1 = (((u[0] >> np.uint64(6)) << np.uint64(27)) | (u[1] >> np.uint64(5))).astype(np.float64)
    r2 = (((u[2] >> np.uint64(6)) << np.uint64(27)) | (u[3] >> np.uint64(5))).astype(np.float64)
    r1 = (r1 + 0.5) * 2.0 ** -53
    r2 = (r2 + 0.5) * 2.0 ** -53
    rad = np.sqrt(-2.0 * np.log(r1))
    g1 = (rad * np.cos(2 * math.pi * r2)).reshape((ny, nx)).T
    g2 = (rad * np.sin(2 * math.pi * r2)).reshape((ny, nx)).T
    return g1, g2


def device_initial_state(ops: Operators, prof: dict, amp_noise=0.0, phase_noise=0.0, seed=0, vortices=None,
                         sigma=0.0):
    """Fields of the device initial-state builder: stationary profile (:82-98) times
    amp = (1 + amp_noise g1)(1 - sum exp(-r^2/sigma^2)) and exp(i phase),
    phase = phase_noise g2 + sum q atan2(y - Y, x - X); A_x set to 0 (func_Lx of an
    x-independent theta)."""
    z1, z2, z3, z4, rhoi = stationary_fields(ops, prof)
    nx, ny = ops.nx, ops.ny
    X = ops.xx1[:, None]                                  # xx, yy (:42-43)
    Y = ops.yy1[None, :]
    amp = np.ones((nx, ny))
    phase = np.zeros((nx, ny))
    if amp_noise or phase_noise:
        g1, g2 = column_normals(nx, ny, seed)
        amp = amp + amp_noise * g1
        phase = phase + phase_noise * g2
    dip = np.zeros((nx, ny))
    if vortices is not None:
        for x0, y0, q in zip(*vortices):
            phase = phase + q * np.arctan2(Y - y0, X - x0)
            if sigma > 0:
                dip = dip + np.exp(-((X - x0) ** 2 + (Y - y0) ** 2) / sigma ** 2)
    amp = amp * (1 - dip)
    z1 = imprint(ops, z1, phase, amp)
    return z1, z2, np.zeros_like(z3), z4, rhoi

