"""Pins of the oracle's continuation in mu (:func:`oracle.hqt_oracle.newton_mu_continuation`, the
procedure of scripts/legacy/Statinary_state_with_v.m:11-17,50-52 written as a loop).

The guesses are the serving workload's: the reference's mu = 6, kk = 7 profile splined to nz
nodes with A_t scaled to mu, at the profile's own flow vy = -2*pi*7/20.  What fixes the result
independently of the loop's own arithmetic:

* it is a solution at the target: NewtonRaphson_vy started from it stops at iteration 0, and its
  boundary value A_t(z=0) is the guess's own;
* it lies on the seed's (superfluid) branch whatever the path: other step lengths and another seed
  give the same profile to the Newton tolerance, and the condensate grows with mu along it;
* a seed equal to the target is one plain NewtonRaphson_vy solve, bit for bit;
* a single long step leaves the branch (lands on the normal phase Psi = 0), which is why the
  steps are short;
* below the end of the branch (a fold near mu = 5.79 at this flow) the walk lands on the normal
  phase, an exact solution (Psi = 0, A_t = mu(1 - z)).
"""

import math

import numpy as np
import pytest

from oracle import hqt_oracle as O

VY = -2 * math.pi * 7 / 20


def _guess(nz, mu):
    from hqt import loadgen
    return loadgen.newton_guess(nz, mu)


@pytest.mark.parametrize("nz,mu", [(32, 8.25), (48, 9.25), (16, 7.25), (24, 8.75)])
def test_continuation_reaches_a_solution_at_the_target(nz, mu):
    Z1, Z2, V = _guess(nz, mu)
    z1, z2, v, it, err, conv, solves = O.newton_mu_continuation(nz, 1, 0, 0, Z1, Z2, V, VY, 6.0, 0.25)
    assert conv and solves >= 2
    # the guess's own A_t(z=0) (the Dirichlet rows keep it up to the LU's rounding)
    assert abs(z2[-1] - Z2[-1]) < 1e-12 * mu
    assert abs(v[-1] - VY) < 1e-12 and abs(z1[-1]) < 1e-12 and abs(z2[0]) < 1e-12
    r1, r2, rv, rit, rerr = O.NewtonRaphson_vy(nz, 1, 0, 0, z1, z2, v, VY)
    assert rit == 0                                           # already a solution
    assert np.max(np.abs(z1)) > 1.0                           # the superfluid branch


def test_direct_solves_that_fail_are_reached():
    """At nz = 48, mu = 9.25 and nz = 24, mu = 8.75 NewtonRaphson_vy from the scaled guess wanders for
    its 1000 iterations; the continuation reaches the target."""
    for nz, mu in ((48, 9.25), (24, 8.75)):
        Z1, Z2, V = _guess(nz, mu)
        assert O.NewtonRaphson_vy(nz, 1, 0, 0, Z1, Z2, V, VY)[3] == 1000
        assert O.newton_mu_continuation(nz, 1, 0, 0, Z1, Z2, V, VY, 6.0, 0.25)[5]


@pytest.mark.parametrize("nz,mu", [(32, 8.25), (48, 9.25)])
def test_continuation_is_path_independent(nz, mu):
    Z1, Z2, V = _guess(nz, mu)
    base = O.newton_mu_continuation(nz, 1, 0, 0, Z1, Z2, V, VY, 6.0, 0.25)
    for seed, dmu in ((6.0, 0.1), (7.0, 0.3), (6.0, 1.0)):
        r = O.newton_mu_continuation(nz, 1, 0, 0, Z1, Z2, V, VY, seed, dmu)
        assert r[5]
        for a, b in zip(r[:3], base[:3]):
            assert np.max(np.abs(a - b)) < 1e-7, (seed, dmu)


def test_condensate_grows_along_the_branch():
    nz = 32
    mx = []
    for mu in np.arange(6.0, 10.01, 0.5):
        Z1, Z2, V = _guess(nz, float(mu))
        r = O.newton_mu_continuation(nz, 1, 0, 0, Z1, Z2, V, VY, 6.0, 0.25)
        assert r[5]
        mx.append(float(np.max(np.abs(r[0]))))
    assert all(b > a + 0.1 for a, b in zip(mx, mx[1:])), mx


def test_seed_at_the_target_is_one_plain_solve():
    nz = 30
    p_Z1, p_Z2, p_V = _guess(nz, 6.0)
    mu = float(p_Z2[-1])
    r = O.newton_mu_continuation(nz, 1, 0, 0, p_Z1, p_Z2, p_V, VY, mu, 0.25)
    d = O.NewtonRaphson_vy(nz, 1, 0, 0, p_Z1, p_Z2, p_V, VY, it_max=100)
    assert r[6] == 1 and r[5]
    for a, b in zip(r[:3], d[:3]):
        assert np.array_equal(a, b)
    assert r[3] == d[3]


def test_one_long_step_leaves_the_branch():
    nz = 32
    Z1, Z2, V = _guess(nz, 8.25)
    r = O.newton_mu_continuation(nz, 1, 0, 0, Z1, Z2, V, VY, 6.0, 4.0)
    assert r[5] and r[6] == 2
    assert np.max(np.abs(r[0])) < 1e-6                        # Psi = 0 (to the Newton tolerance): normal phase


def test_below_the_fold_lands_on_the_normal_phase():
    nz = 32
    zc, _ = O.chebdif(nz, 1)
    z = (zc - 1) / 2 + 1
    for mu in (5.5, 3.5):
        Z1, Z2, V = _guess(nz, mu)
        r = O.newton_mu_continuation(nz, 1, 0, 0, Z1, Z2, V, VY, 6.0, 0.25)
        assert r[5]
        assert np.max(np.abs(r[0])) < 1e-6
        assert np.max(np.abs(r[1] - Z2[-1] * (1 - z))) < 1e-6  # A_t = mu (1 - z)
        assert np.max(np.abs(r[2] - VY)) < 1e-6
