"""hqt_newton_vy_batch_noalloc (the batched Newton solve with its device workspace in a fixed
per-device array, no allocation) against hqt_newton_vy_batch, through the C ABI: bitwise the
same outputs and status, over consecutive calls that reuse the array with a smaller batch, on
the LDS-resident (nz <= 41) and the global-workspace (nz > 41) paths; oversize batches are
rejected before any GPU work."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VY_REF = -2 * math.pi * 7 / 20


def _batch(nz, nb, seed):
    from hqt import loadgen
    rng = np.random.default_rng(seed)
    cols = [loadgen.newton_guess(nz, float(rng.uniform(4.5, 8.0))) for _ in range(nb)]
    Z1, Z2, V = (np.stack([c[k] for c in cols], axis=1) for k in range(3))
    vy = VY_REF * rng.uniform(0.0, 1.0, nb)
    return Z1, Z2, V, vy


def _both(nz, nb, seed):
    import hqt
    Z1, Z2, V, vy = _batch(nz, nb, seed)
    a = hqt.NewtonRaphson_vy_batch(nz, 1, 0, 0, Z1, Z2, V, vy)
    b = hqt.NewtonRaphson_vy_batch(nz, 1, 0, 0, Z1, Z2, V, vy, allocation_free=True)
    return a, b


@pytest.mark.parametrize("sizes", [((30, 40), (20, 5)), ((48, 20), (44, 3))], ids=["lds", "global"])
def test_noalloc_equals_allocating_entry(sizes):
    for k, (nz, nb) in enumerate(sizes):
        print(f"noalloc Newton call {k}: nz={nz} nb={nb}", flush=True)
        a, b = _both(nz, nb, 100 * nz + nb)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), (nz, nb)
        assert a[5].sum() >= nb // 2, "most problems of the batch converge"


def test_noalloc_status_matches_on_nonconvergence():
    """A problem that does not converge within it_max gives HQT_ERR_NOCONV from both entries."""
    import ctypes as C

    import hqt
    from hqt._lib import dptr
    nz, nb = 24, 3
    Z1, Z2, V, vy = _batch(nz, nb, 7)
    rcs = []
    outs = []
    for fn in (hqt.lib().hqt_newton_vy_batch, hqt.lib().hqt_newton_vy_batch_noalloc):
        z1, z2, v = (np.empty((nz, nb), order="F") for _ in range(3))
        it = np.zeros(nb, dtype=np.int32)
        err = np.zeros(nb)
        conv = np.zeros(nb, dtype=np.uint8)
        args = [np.asfortranarray(x) for x in (Z1, Z2, V)]
        rcs.append(fn(nz, 1.0, 0.0, 0.0, nb, *[dptr(x) for x in args], dptr(np.ascontiguousarray(vy)), 2, 1e-300,
                      dptr(z1), dptr(z2), dptr(v), it.ctypes.data_as(C.POINTER(C.c_int32)), dptr(err),
                      conv.ctypes.data_as(C.POINTER(C.c_uint8))))
        outs.append((z1, z2, v, it, err, conv))
    assert rcs == [hqt.HQT_ERR_NOCONV, hqt.HQT_ERR_NOCONV]
    for x, y in zip(*outs):
        assert np.array_equal(x, y)


def test_noalloc_rejects_oversize_batch():
    import hqt
    nz, nb = 64, 100          # 100 x (3 * 64)^2 * 8 bytes of Jacobians alone: past the fixed array
    Z1, Z2, V, vy = _batch(nz, 2, 3)
    Z1, Z2, V = (np.repeat(x[:, :1], nb, axis=1) for x in (Z1, Z2, V))
    with pytest.raises(hqt.HQTError) as ei:
        hqt.NewtonRaphson_vy_batch(nz, 1, 0, 0, Z1, Z2, V, np.zeros(nb), allocation_free=True)
    assert ei.value.code == hqt.HQT_ERR_ARG
    assert "allocation-free workspace" in str(ei.value)
    a, b = _both(30, 4, 5)                     # the array still serves the next call
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
