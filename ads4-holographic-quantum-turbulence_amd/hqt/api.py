"""Python mirror of the reference's interface for the stepper path, over the C ABI.

Names, argument meaning and error behaviour follow the reference's MATLAB routines:

* :func:`fourdif`, :func:`chebdif`  -- ``third_party/dmsuite/fourdif.m:1``, ``chebdif.m:1``
* :func:`ff`                        -- ``src/numerics/ff.m:1`` (computed on the GPU)
* :func:`NewtonRaphson_vy`          -- ``src/solvers/NewtonRaphson_vy.m:1``
* :class:`Stepper`                  -- the closures ``F1..F4`` (``tevol_driven_cluster.m:72-76``)
  and the loop body ``:120-142`` with its NaN guard ``:182-185``

Fields are NumPy arrays of shape ``(nx, ny, nz)`` (MATLAB ``[nx,ny,nz]``); they are handed
to the library in column-major order (x fastest), the reference's memory layout.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (HQT_ERR_ARG, HQT_ERR_NAN, HQT_ERR_NOCONV, HQT_ERR_TIMEOUT, HQT_FP64, HQT_LINES16, HQT_LINES32, HQT_MIXED, HQT_NO_GRAPH, HQT_RADIAL_W2, HQT_RADIAL_W3, HQT_DENSE_FILTER, HQT_DENSE_LINES, HQT_FACTORED_STAGES, HQT_FOLDED_FILTER, HQT_FOLDED_STAGES, HQT_FP32_FILTER, HQT_ORDER_XCD, HQT_SERVER_NCOUNTERS, HQTError, Init, Params, check, dptr, lib)

__all__ = ["fourdif", "chebdif", "ff", "NewtonRaphson_vy", "NewtonRaphson_vy_batch", "Stepper", "Ensemble", "Server", "footprint", "make_params", "checkpoint_name", "write_checkpoint", "spline", "HQTError", "mask_factors"]


def fourdif(N: int, m: int):
    """``[x, DM] = fourdif(N, m)``: Fourier differentiation matrix on N equispaced points."""
    x = np.empty(N)
    DM = np.empty((N, N), order="F")
    check(lib().hqt_fourdif(int(N), int(m), dptr(x), dptr(DM)))
    return x, DM


def chebdif(N: int, M: int):
    """``[x, DM] = chebdif(N, M)``: Chebyshev nodes (descending) and DM of shape (N, N, M)."""
    x = np.empty(N)
    DM = np.empty((N, N, M), order="F")
    check(lib().hqt_chebdif(int(N), int(M), dptr(x), dptr(DM)))
    return x, DM


def mask_factors(mask: np.ndarray):
    """Split a boolean ``[nx, ny(, nz)]`` mask of the form ``fx | fy`` (every z slice equal,
    ``tevol_driven_cluster.m:27-31``) into the 1-D factors ``(fx, fy)``."""
    m = np.asarray(mask, dtype=bool)
    if m.ndim == 3:
        if not (m == m[:, :, :1]).all():
            raise HQTError(HQT_ERR_ARG, "ff: mask slices differ; only masks fx|fy repeated over z are supported")
        m = m[:, :, 0]
    if m.ndim != 2:
        raise HQTError(HQT_ERR_ARG, "ff: mask must be [nx, ny] or [nx, ny, nz]")
    fx = m.all(axis=1)
    fy = m.all(axis=0)
    if not np.array_equal(m, fx[:, None] | fy[None, :]):
        raise HQTError(HQT_ERR_ARG, "ff: mask is not of the form fx | fy")
    return fx, fy


def ff(x, y):
    """``w = ff(x, y)``: zero the masked 2-D Fourier modes of every slice and return the real
    part (``src/numerics/ff.m:7-11``).  ``x`` is ``[nx, ny]`` or ``[nx, ny, nslices]``, real or
    complex; ``y`` is a logical mask ``[nx, ny]`` (every slice) or ``[nx, ny, nslices]``.  A mask of
    the form ``fx | fy`` repeated over the slices (the stepper's, ``tevol_driven_cluster.m:27-31``)
    runs through the separable line passes (``hqt_ff``), any other through direct 2-D DFTs
    (``hqt_ff_mask``); both on the GPU."""
    x = np.asarray(x)
    shape = x.shape
    if x.ndim == 2:
        x = x[:, :, None]
    if x.ndim != 3:
        raise HQTError(HQT_ERR_ARG, "ff: x must be 2-D or 3-D")
    nx, ny, ns = x.shape
    m = np.asarray(y, dtype=bool)
    if m.ndim not in (2, 3) or m.shape[:2] != (nx, ny) or (m.ndim == 3 and m.shape[2] not in (1, ns)):
        raise HQTError(HQT_ERR_ARG, "ff: mask shape does not match x")
    xr = np.asfortranarray(np.real(x), dtype=np.float64)
    xi = np.asfortranarray(np.imag(x), dtype=np.float64) if np.iscomplexobj(x) else None
    w = np.empty((nx, ny, ns), order="F")
    u8 = C.POINTER(C.c_uint8)
    try:
        fx, fy = mask_factors(m)
    except HQTError:
        fx = fy = None
    if fx is not None and nx <= 512 and ny <= 512:
        fxb = np.ascontiguousarray(fx, dtype=np.uint8)
        fyb = np.ascontiguousarray(fy, dtype=np.uint8)
        check(lib().hqt_ff(nx, ny, ns, dptr(xr), dptr(xi), fxb.ctypes.data_as(u8), fyb.ctypes.data_as(u8), dptr(w)))
    else:
        mb = np.asfortranarray(m if m.ndim == 3 else m[:, :, None], dtype=np.uint8)
        check(lib().hqt_ff_mask(nx, ny, ns, dptr(xr), dptr(xi), mb.ctypes.data_as(u8), mb.shape[2], dptr(w)))
    return w.reshape(shape)


def NewtonRaphson_vy(nz, zh, small, D, Z1, Z2, vy1, vy, raise_on_failure=False):
    """``[z1,z2,VY1,it,error] = NewtonRaphson_vy(nz,zh,small,D,Z1,Z2,vy1,vy)``.

    Like the reference, a run that stops at ``it_max`` or on a NaN update returns normally
    (``it`` then reports the count); pass ``raise_on_failure=True`` to raise instead."""
    nz = int(nz)
    a = np.ascontiguousarray(np.ravel(Z1), dtype=np.float64)
    b = np.ascontiguousarray(np.ravel(Z2), dtype=np.float64)
    v = np.ascontiguousarray(np.ravel(vy1), dtype=np.float64)
    if a.size != nz or b.size != nz or v.size != nz:
        raise HQTError(HQT_ERR_ARG, "NewtonRaphson_vy: profiles must have nz entries")
    z1, z2, VY1 = np.empty(nz), np.empty(nz), np.empty(nz)
    it = C.c_int32(0)
    err = C.c_double(0.0)
    rc = lib().hqt_newton_vy(nz, float(zh), float(small), float(D), dptr(a), dptr(b), dptr(v), float(vy),
                             dptr(z1), dptr(z2), dptr(VY1), C.byref(it), C.byref(err))
    if rc != 0 and raise_on_failure:
        check(rc)
    return z1, z2, VY1, int(it.value), float(err.value)


def NewtonRaphson_vy_batch(nz, zh, small, D, Z1, Z2, vy1, vy, it_max=1000, tol=1e-6, raise_on_failure=False,
                           allocation_free=False, mu_seed=None, dmu=0.25):
    """NewtonRaphson_vy over a batch on the GPU (one work-group per problem).

    ``Z1, Z2, vy1`` are ``[nz, nb]`` (problem ``b`` in column ``b``; an nz-vector is broadcast
    over the batch), ``vy`` is ``[nb]``.  Returns ``z1, z2, VY1`` (``[nz, nb]``), ``it``,
    ``err`` and ``conv`` (``[nb]``), each problem iterated exactly as ``NewtonRaphson_vy``.
    ``allocation_free=True`` runs ``hqt_newton_vy_batch_noalloc`` (no device allocation; the
    workspace is a fixed per-device array of ``HQT_NEWTON_NOALLOC_BYTES``).
    ``mu_seed`` (a number or ``[nb]``; NaN entries solve once) runs
    ``hqt_newton_vy_continue_batch``: each guess is taken as a solution near ``mu_seed`` and
    continued in mu to its own A_t(z=0) = ``Z2[-1]`` in steps of at most ``dmu``; the number of
    solves per problem is then returned as a seventh value."""
    nz = int(nz)
    vy = np.atleast_1d(np.asarray(vy, dtype=np.float64))
    nb = vy.size

    def cols(a):
        a = np.asarray(a, dtype=np.float64)
        if a.ndim == 1:
            a = np.repeat(a[:, None], nb, axis=1)
        if a.shape != (nz, nb):
            raise HQTError(HQT_ERR_ARG, f"NewtonRaphson_vy_batch: profiles must be [{nz}, {nb}]")
        return np.asfortranarray(a)

    a, b, v = cols(Z1), cols(Z2), cols(vy1)
    vy = np.ascontiguousarray(vy)
    z1, z2, VY1 = (np.empty((nz, nb), order="F") for _ in range(3))
    it = np.zeros(nb, dtype=np.int32)
    err = np.zeros(nb)
    conv = np.zeros(nb, dtype=np.uint8)
    tail = (dptr(z1), dptr(z2), dptr(VY1), it.ctypes.data_as(C.POINTER(C.c_int32)), dptr(err),
            conv.ctypes.data_as(C.POINTER(C.c_uint8)))
    if mu_seed is not None:
        if allocation_free:
            raise HQTError(HQT_ERR_ARG, "NewtonRaphson_vy_batch: continuation has no allocation-free form")
        ms = np.ascontiguousarray(np.broadcast_to(np.asarray(mu_seed, dtype=np.float64), (nb,)))
        steps = np.zeros(nb, dtype=np.int32)
        rc = lib().hqt_newton_vy_continue_batch(nz, float(zh), float(small), float(D), nb, dptr(a), dptr(b), dptr(v),
                                                dptr(vy), dptr(ms), float(dmu), int(it_max), float(tol), *tail,
                                                steps.ctypes.data_as(C.POINTER(C.c_int32)))
    else:
        fn = lib().hqt_newton_vy_batch_noalloc if allocation_free else lib().hqt_newton_vy_batch
        rc = fn(nz, float(zh), float(small), float(D), nb, dptr(a), dptr(b), dptr(v), dptr(vy), int(it_max),
                float(tol), *tail)
    if rc != 0 and (raise_on_failure or rc != HQT_ERR_NOCONV):
        check(rc)
    if mu_seed is not None:
        return z1, z2, VY1, it, err, conv.astype(bool), steps
    return z1, z2, VY1, it, err, conv.astype(bool)


def _f64F(a, shape):
    a = np.asarray(a, dtype=np.float64)
    if a.shape != shape:
        a = np.broadcast_to(a, shape)
    return np.asfortranarray(a)


def _c128F(a, shape):
    a = np.asarray(a, dtype=np.complex128)
    if a.shape != shape:
        a = np.broadcast_to(a, shape)
    return np.asfortranarray(a)


def make_params(nx, ny, nz, Lx, Ly, mu=6.0, dt=0.02, winding=True, filter=True, precision="fp64", ax_b=0.0,
                ay_b=0.0, graph=True, lines=None, radial_waves=None, dense_filter=False, dense_lines=False,
                xcd_order=False, factored=None, folded_filter=False, fp32_filter=False):
    """``hqt_params`` for a run, plus the arrays its boundary pointers borrow (keep them alive
    while the params are in use).  ``lines`` = 16 or 32 fixes the lines per work-group of the
    x/y passes; None lets the library measure and pick per pass.  ``radial_waves`` = 2 or 3 fixes the
    occupancy the radial stage kernel is compiled for (None: 2).
    ``dense_filter`` applies the filter's imaginary part as a dense Toeplitz product instead of
    its rank-4 form.  ``dense_lines`` applies the x/y operators and the filter as dense Toeplitz
    products (func_Lx's own form) instead of their even/odd folded form.  ``xcd_order`` (with ``lines``)
    runs the folded line passes in the XCD-aware block order.  ``factored`` chooses the form of the RK
    stages' x/y derivatives where the line length has a factored-DFT form (one DFT per source line, the
    operators' eigenvalues, one inverse per output): None takes it in the mixed form at every supported line
    length (FP32 transforms round far less than FP32 matrix products) and for FP64 lines of 128 points or
    more, the folded matrix products otherwise; True / False force one form.  The filter follows the stages' form
    unless ``folded_filter`` keeps its folded matrix products.  In the mixed form the filter computes in FP64
    (FP32 state in and out) unless ``fp32_filter`` runs it in FP32."""
    p = Params()
    p.nx, p.ny, p.nz = int(nx), int(ny), int(nz)
    p.Lx, p.Ly, p.mu, p.dt = float(Lx), float(Ly), float(mu), float(dt)
    p.winding = 1 if winding else 0
    p.filter = 1 if filter else 0
    if precision not in ("fp64", "mixed"):
        raise HQTError(HQT_ERR_ARG, "precision must be 'fp64' or 'mixed'")
    p.precision = HQT_FP64 if precision == "fp64" else HQT_MIXED
    p.flags = ((0 if graph else HQT_NO_GRAPH) | {None: 0, 16: HQT_LINES16, 32: HQT_LINES32}[lines]
               | {None: 0, 2: HQT_RADIAL_W2, 3: HQT_RADIAL_W3}[radial_waves]
               | (HQT_DENSE_FILTER if dense_filter else 0) | (HQT_DENSE_LINES if dense_lines else 0)
               | (HQT_ORDER_XCD if xcd_order else 0) | (HQT_FOLDED_STAGES if factored is False else 0)
               | (HQT_FACTORED_STAGES if factored is True else 0) | (HQT_FOLDED_FILTER if folded_filter else 0)
               | (HQT_FP32_FILTER if fp32_filter else 0))
    keep = []
    for name, val in (("ax_b", ax_b), ("ay_b", ay_b)):
        if np.ndim(val) == 0:
            setattr(p, name + "_const", float(val))
            setattr(p, name, None)
        else:
            arr = np.asfortranarray(np.broadcast_to(np.asarray(val, dtype=np.float64), (p.nx, p.ny)))
            keep.append(arr)
            setattr(p, name, dptr(arr))
    return p, keep


def spline(X, Y, Xi):
    """``interp1(X, Y, Xi, 'spline')``: not-a-knot cubic spline (``hqt_spline``)."""
    X = np.ascontiguousarray(X, dtype=np.float64).ravel()
    Y = np.ascontiguousarray(Y, dtype=np.float64).ravel()
    Xi = np.ascontiguousarray(Xi, dtype=np.float64).ravel()
    out = np.empty_like(Xi)
    check(lib().hqt_spline(X.size, dptr(X), dptr(Y), Xi.size, dptr(Xi), dptr(out)))
    return out


def checkpoint_name(t):
    """``['check_point_t=', num2str(t), '.mat']`` (``tevol_driven_cluster.m:179``)."""
    buf = C.create_string_buffer(128)
    check(lib().hqt_checkpoint_name(float(t), buf, 128))
    return buf.value.decode()


def write_checkpoint(path, Z1, Z2, Z3, Z4, rhoi):
    """``save(path, 'Z1','Z2','Z3','Z4','rhoi')`` of host arrays [nx, ny, nz] (MAT level 5)."""
    s = np.shape(Z2)
    z1 = _c128F(Z1, s)
    arrs = [_f64F(a, s) for a in (Z2, Z3, Z4, rhoi)]
    check(lib().hqt_checkpoint_write_host(str(path).encode(), s[0], s[1], s[2], dptr(z1), *[dptr(a) for a in arrs]))


def footprint(nx, ny, nz, precision="fp64"):
    """Device bytes a context of this size allocates (``hqt_ctx_footprint``)."""
    p, _ = make_params(nx, ny, nz, 50.0, 50.0, precision=precision)
    b = C.c_int64(0)
    check(lib().hqt_ctx_footprint(C.byref(p), C.byref(b)))
    return int(b.value)


class Stepper:
    """A device-resident AdS4 evolution (one ``hqt_ctx``).

    Parameters mirror the workspace of ``tevol_driven_cluster.m``: grid ``nx, ny, nz``, box
    sides ``Lx`` (L0) and ``Ly`` (L), ``mu``, time step ``dt`` (h), winding removal on/off
    (on in the driven scripts, off in ``tevol_lattice.m``), the filter on/off, the z=0
    values ``ax_b`` / ``ay_b`` of A_x / A_y restored after every step (scalars or [nx, ny]
    arrays; the reference uses 0 and -vy), and ``precision`` ``"fp64"`` or ``"mixed"``.
    ``graph=False`` launches each kernel directly instead of replaying the captured step graph."""

    def __init__(self, nx, ny, nz, Lx, Ly, mu=6.0, dt=0.02, winding=True, filter=True,
                 precision="fp64", ax_b=0.0, ay_b=0.0, graph=True, lines=None, radial_waves=None,
                 dense_filter=False, dense_lines=False, xcd_order=False, factored=None, folded_filter=False,
                 fp32_filter=False):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.shape = (self.nx, self.ny, self.nz)
        self.precision = precision
        p, keep = make_params(nx, ny, nz, Lx, Ly, mu, dt, winding, filter, precision, ax_b, ay_b, graph, lines,
                              radial_waves, dense_filter, dense_lines, xcd_order, factored, folded_filter,
                              fp32_filter)
        self._h = C.c_void_p()
        check(lib().hqt_ctx_create(C.byref(p), C.byref(self._h)))

    # -- life cycle ---------------------------------------------------------------------
    def close(self):
        ens = getattr(self, "_ensemble", None)
        if ens is not None and ens() is not None:
            ens().close()                       # release the context from its ensemble first
        if getattr(self, "_h", None) and self._h.value:
            lib().hqt_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state --------------------------------------------------------------------------
    def set_state(self, Z1, Z2, Z3, Z4, rhoi):
        s = self.shape
        z1 = _c128F(Z1, s)
        check(lib().hqt_set_state(self._h, dptr(z1), dptr(_f64F(Z2, s)), dptr(_f64F(Z3, s)),
                                  dptr(_f64F(Z4, s)), dptr(_f64F(rhoi, s))))

    def get_state(self, check_nan=True):
        """Return ``(Z1, Z2, Z3, Z4, rhoi)``; raises HQTError(HQT_ERR_NAN) if the guard tripped
        (unless ``check_nan=False``)."""
        s = self.shape
        Z1 = np.empty(s, dtype=np.complex128, order="F")
        Z2, Z3, Z4, rh = (np.empty(s, order="F") for _ in range(4))
        rc = lib().hqt_get_state(self._h, dptr(Z1), dptr(Z2), dptr(Z3), dptr(Z4), dptr(rh))
        if rc != 0 and not (rc == HQT_ERR_NAN and not check_nan):
            check(rc)
        return Z1, Z2, Z3, Z4, rh

    # -- stepping -----------------------------------------------------------------------
    def step_async(self, nsteps=1):
        check(lib().hqt_step_n(self._h, int(nsteps)))

    def sync(self):
        """Wait for queued steps; raise HQTError(HQT_ERR_NAN) if the NaN guard tripped."""
        ns = C.c_int32(0)
        check(lib().hqt_sync(self._h, C.byref(ns)))

    def nan_step(self):
        ns = C.c_int32(0)
        lib().hqt_sync(self._h, C.byref(ns))
        return int(ns.value)

    def step(self, nsteps=1):
        self.step_async(nsteps)
        self.sync()

    @property
    def last_winding(self):
        w = C.c_double(0.0)
        check(lib().hqt_last_winding(self._h, C.byref(w)))
        return w.value

    # -- closures F1..F4 ------------------------------------------------------------------
    def rhs(self, Z1, Z2, Z3, Z4):
        """``(F1, F2, F3, F4)`` at the given fields (``tevol_driven_cluster.m:72-76``)."""
        s = self.shape
        z1 = _c128F(Z1, s)
        F1 = np.empty(s, dtype=np.complex128, order="F")
        F2, F3, F4 = (np.empty(s, order="F") for _ in range(3))
        check(lib().hqt_rhs(self._h, dptr(z1), dptr(_f64F(Z2, s)), dptr(_f64F(Z3, s)),
                            dptr(_f64F(Z4, s)), dptr(F1), dptr(F2), dptr(F3), dptr(F4)))
        return F1, F2, F3, F4

    def F1(self, y1, y2, y3, y4):
        return self.rhs(y1, y2, y3, y4)[0]

    def F2(self, y1, y2, y3, y4):
        return self.rhs(y1, y2, y3, y4)[1]

    def F3(self, y1, y2, y3, y4):
        return self.rhs(y1, y2, y3, y4)[2]

    def F4(self, y1, y2, y3, y4):
        return self.rhs(y1, y2, y3, y4)[3]

    def init_state(self, prof, amp_noise=0.0, phase_noise=0.0, seed=0, vortices=None, sigma=0.0):
        """Build the initial fields on the device (``hqt_init_state``): the stationary profile
        ``prof`` (dict of nz0-vectors ``Z1, Z2, Vy1, rhoi`` on chebdif(nz0) nodes) splined to
        this grid and gauge-rotated, times amp = (1 + amp_noise g1)(1 - sum of Gaussian dips of
        width sigma) and exp(i phase), phase = phase_noise g2 + sum q atan2(y - Y, x - X).
        ``vortices`` = (X, Y, q) arrays."""
        arr = [np.ascontiguousarray(prof[k], dtype=np.float64).ravel() for k in ("Z1", "Z2", "Vy1", "rhoi")]
        d = Init()
        d.nz0 = arr[0].size
        d.Z1, d.Z2, d.Vy1, d.rhoi = (dptr(a) for a in arr)
        d.amp_noise, d.phase_noise, d.seed = float(amp_noise), float(phase_noise), int(seed) & ((1 << 64) - 1)
        keep = []
        if vortices is not None and len(vortices[0]):
            keep = [np.ascontiguousarray(v, dtype=np.float64).ravel() for v in vortices]
            d.nv = keep[0].size
            d.vx, d.vy, d.q = (dptr(a) for a in keep)
        else:
            d.nv = 0
        d.sigma = float(sigma)
        check(lib().hqt_init_state(self._h, C.byref(d)))

    def save_checkpoint(self, path, wait=True):
        """Write the current state as a reference checkpoint file.  ``wait=False`` snapshots the
        state on the device and writes the file from a host thread while stepping goes on;
        :meth:`checkpoint_wait` joins it."""
        if wait:
            check(lib().hqt_checkpoint_save(self._h, str(path).encode()))
        else:
            check(lib().hqt_checkpoint_save_async(self._h, str(path).encode()))

    def checkpoint_wait(self):
        check(lib().hqt_checkpoint_wait(self._h))

    @property
    def device_bytes(self):
        b = C.c_int64(0)
        check(lib().hqt_ctx_device_bytes(self._h, C.byref(b)))
        return int(b.value)

    def order_parameter(self):
        """O = (func_Lz(Z1))(:,:,end) of the current state, complex ``[nx, ny]``."""
        O = np.empty((self.nx, self.ny), dtype=np.complex128, order="F")
        check(lib().hqt_order_parameter(self._h, dptr(O)))
        return O

    def vortex_census(self, charges=False):
        """``[n, np, nm] = countN(O)`` of the current state, computed on the GPU.

        Returns ``(n, np, nm, Omax, Omean)``, plus the ``[nx, ny]`` int8 plaquette charge map
        when ``charges`` is true."""
        counts = np.zeros(3, dtype=np.int32)
        oabs = np.zeros(2)
        cmap = np.zeros((self.nx, self.ny), dtype=np.int8, order="F") if charges else None
        check(lib().hqt_vortex_census(self._h, counts.ctypes.data_as(C.POINTER(C.c_int32)),
                                      cmap.ctypes.data_as(C.POINTER(C.c_int8)) if charges else None, dptr(oabs)))
        out = (int(counts[0]), int(counts[1]), int(counts[2]), float(oabs[0]), float(oabs[1]))
        return out + (cmap,) if charges else out

    def debug_field(self, name):
        """An internal device field of the last rhs/step (x/y/z derivative buffers)."""
        out = np.empty(self.shape, order="F")
        check(lib().hqt_debug_field(self._h, name.encode(), dptr(out)))
        return out

    # -- measurement --------------------------------------------------------------------
    def profile(self, enable=True):
        check(lib().hqt_profile(self._h, 1 if enable else 0))

    def kernel_stats(self):
        """``{name: (launches, total_ms)}`` for kernels launched since ``profile(True)``."""
        out = {}
        buf = C.create_string_buffer(64)
        cnt = C.c_int64(0)
        tot = C.c_double(0.0)
        i = 0
        while lib().hqt_kernel_stats(self._h, i, buf, 64, C.byref(cnt), C.byref(tot)) == 0:
            if cnt.value:
                out[buf.value.decode()] = (int(cnt.value), float(tot.value))
            i += 1
        return out


class Ensemble:
    """Independent evolutions stepped together (``hqt_ensemble``): the members split into four
    lanes (members of one shape together), each lane one launch per kernel class of the step on
    its own stream, the steps replayed from captured hipGraphs of eight steps and of one step.

    ``steppers`` are :class:`Stepper` objects with their state set; they may differ in grid,
    nz, mu, drive, switches and precision.  Each member's result equals stepping it alone.
    While the ensemble lives, the members' own calls wait for its queued steps."""

    def __init__(self, steppers, graph=True):
        import weakref
        self.members = list(steppers)
        arr = (C.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        self._h = C.c_void_p()
        check(lib().hqt_ensemble_create(arr, len(self.members), 0 if graph else HQT_NO_GRAPH, C.byref(self._h)))
        for m in self.members:
            m._ensemble = weakref.ref(self)

    @property
    def launches_per_step(self):
        n = C.c_int32(0)
        check(lib().hqt_ensemble_launches(self._h, C.byref(n)))
        return int(n.value)

    def step_async(self, nsteps=1):
        check(lib().hqt_ensemble_step_n(self._h, int(nsteps)))

    def nan_steps(self):
        """Wait; per-member NaN step (0 = none), without raising."""
        ns = np.zeros(len(self.members), dtype=np.int32)
        lib().hqt_ensemble_sync(self._h, ns.ctypes.data_as(C.POINTER(C.c_int32)))
        return ns

    def sync(self):
        """Wait for queued steps; raise HQTError(HQT_ERR_NAN) if any member's guard tripped."""
        ns = np.zeros(len(self.members), dtype=np.int32)
        check(lib().hqt_ensemble_sync(self._h, ns.ctypes.data_as(C.POINTER(C.c_int32))))

    def step(self, nsteps=1):
        self.step_async(nsteps)
        self.sync()

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().hqt_ensemble_destroy(self._h)
            self._h = C.c_void_p()
            for m in self.members:
                m._ensemble = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Server:
    """The serving loop (``hqt_server_*``): evolution segments and stationary Newton solves
    submitted at any time, admitted by HBM budget and batched as they arrive.

    ``submit_evolution`` takes the :class:`Stepper` arguments plus an initial state and a step
    count; ``submit_newton`` takes ``NewtonRaphson_vy``'s arguments.  ``result(id)`` waits and
    returns ``(status, latency_s, payload)``: the state tuple (evolution) or
    ``(z1, z2, VY1, it, err)`` (Newton)."""

    def __init__(self, hbm_budget, quantum=25):
        self._h = C.c_void_p()
        self._kind = {}
        self._shape = {}
        check(lib().hqt_server_create(int(hbm_budget), int(quantum), C.byref(self._h)))

    def submit_evolution(self, state, nsteps, nx, ny, nz, Lx, Ly, **kw):
        p, keep = make_params(nx, ny, nz, Lx, Ly, **kw)
        s = (int(nx), int(ny), int(nz))
        Z1, Z2, Z3, Z4, rh = state
        z1 = _c128F(Z1, s)
        args = [dptr(z1)] + [dptr(_f64F(a, s)) for a in (Z2, Z3, Z4, rh)]
        rid = C.c_int64(0)
        check(lib().hqt_server_submit_evolution(self._h, C.byref(p), *args, int(nsteps), C.byref(rid)))
        self._kind[rid.value] = "evo"
        self._shape[rid.value] = s
        return int(rid.value)

    def submit_newton(self, nz, zh, small, D, Z1, Z2, vy1, vy, mu_seed=None, dmu=0.25):
        """A stationary solve; with ``mu_seed`` the guess is continued in mu from that seed to its
        own A_t(z=0) (``hqt_server_submit_newton_continued``)."""
        a, b, v = (np.ascontiguousarray(np.ravel(x), dtype=np.float64) for x in (Z1, Z2, vy1))
        rid = C.c_int64(0)
        if mu_seed is None:
            check(lib().hqt_server_submit_newton(self._h, int(nz), float(zh), float(small), float(D), dptr(a),
                                                 dptr(b), dptr(v), float(vy), C.byref(rid)))
        else:
            check(lib().hqt_server_submit_newton_continued(self._h, int(nz), float(zh), float(small), float(D),
                                                           dptr(a), dptr(b), dptr(v), float(vy), float(mu_seed),
                                                           float(dmu), C.byref(rid)))
        self._kind[rid.value] = "newton"
        self._shape[rid.value] = int(nz)
        return int(rid.value)

    def result(self, rid, timeout=-1.0):
        lat = C.c_double(0.0)
        st = lib().hqt_server_wait(self._h, int(rid), float(timeout), C.byref(lat))
        if st == HQT_ERR_TIMEOUT or st == HQT_ERR_ARG:
            check(st)
        kind = self._kind.pop(rid)
        shp = self._shape.pop(rid)
        if kind == "evo":
            Z1 = np.empty(shp, dtype=np.complex128, order="F")
            Z2, Z3, Z4, rh = (np.empty(shp, order="F") for _ in range(4))
            ns = C.c_int32(0)
            lib().hqt_server_evolution_result(self._h, int(rid), dptr(Z1), dptr(Z2), dptr(Z3), dptr(Z4), dptr(rh),
                                              C.byref(ns))
            return st, lat.value, (Z1, Z2, Z3, Z4, rh)
        z1, z2, v = np.empty(shp), np.empty(shp), np.empty(shp)
        it = C.c_int32(0)
        err = C.c_double(0.0)
        lib().hqt_server_newton_result(self._h, int(rid), dptr(z1), dptr(z2), dptr(v), C.byref(it), C.byref(err))
        return st, lat.value, (z1, z2, v, int(it.value), float(err.value))

    def stats(self, kind=2):
        """``(p50, p99, completed)`` latency over completed requests: kind 0 evolutions, 1
        Newton solves, 2 both."""
        p50, p99, n = C.c_double(0), C.c_double(0), C.c_int64(0)
        check(lib().hqt_server_stats(self._h, int(kind), C.byref(p50), C.byref(p99), C.byref(n)))
        return p50.value, p99.value, int(n.value)

    def counters(self):
        out = (C.c_int64 * HQT_SERVER_NCOUNTERS)()
        check(lib().hqt_server_counters(self._h, out))
        keys = ("admitted", "evolutions_done", "newton_done", "ensemble_builds", "newton_batches", "peak_bytes",
                "peak_resident", "steps", "newton_solves", "newton_continued")
        return dict(zip(keys, [int(x) for x in out]))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().hqt_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

