"""The radial contraction comparison (hqt_calibration_radial): the stage kernel's FP64 MFMA matvec on the
register slot layout and the vector-FMA matvec fed through LDS apply the same Chebyshev derivative twice
(func_Lz(func_Lz(.)), tevol_driven_cluster.m:62) to every column of a field: both forms agree with each other and
with a NumPy evaluation (the oracle's chebdif, Lz = 2 Dz on z in [0, 1]) of the same input."""

import ctypes as C

import numpy as np
import pytest

from oracle import hqt_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ncol,nz", [(304 * 301, 20), (64 * 64, 12), (1000, 20)])
def test_radial_forms_agree(ncol, nz):
    import hqt
    print("radial A/B", ncol, nz, flush=True)
    ss = {}
    for form in (0, 1):
        ms, s = C.c_double(0.0), C.c_double(0.0)
        hqt._lib.check(hqt.lib().hqt_calibration_radial(ncol, nz, form, 3, C.byref(ms), C.byref(s)))
        assert ms.value > 0
        ss[form] = s.value
    ps = (ncol + 31) // 32 * 32
    i = np.arange(ps * nz, dtype=np.float64)
    x = (np.sin(0.001 * i) + 0.5 * np.cos(0.013 * i)).reshape(nz, ps)[:, :ncol]
    _, DM = O.chebdif(nz, 1)
    Lz = 2.0 * DM[0]
    v = Lz @ (Lz @ x)
    ref = float(np.sum(v * v))
    assert abs(ss[0] - ss[1]) <= 1e-12 * ref
    assert abs(ss[0] - ref) <= 1e-10 * ref
