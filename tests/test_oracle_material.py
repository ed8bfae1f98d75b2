"""Pins of oracle/material.py (rock_prop, calculate_dt) by closed forms."""
import numpy as np
import pytest

from oracle import material as OM


@pytest.fixture(scope="module")
def nodes():
    rng = np.random.default_rng(0)
    X = rng.standard_normal((500, 3))
    X *= (0.546 + 0.454 * rng.random(500) ** 0.5)[:, None] / np.linalg.norm(X, axis=1)[:, None]
    T = 1300 * rng.random(500)
    return X, T


def test_constant_viscosity_and_linear_density(nodes):
    X, T = nodes
    v, d = OM.rock_prop(T, X, Visc0=2.0, Dens0=3300.0, T_ref=1300.0, alpha=3e-5)
    assert np.all(v == 2.0)
    assert np.allclose(d, 3300.0 * (1 - 3e-5 * (T - 1300.0)), rtol=1e-15)


def test_arrhenius_factor_is_one_at_the_reference_and_monotone(nodes):
    X, _ = nodes
    T = np.array([1300.0, 1000.0, 1500.0])
    v, _ = OM.rock_prop(T, X[:3], Visc0=1.0, Dens0=1.0, T_ref=1300.0, Ea=3e5)
    assert v[0] == 1.0 and v[1] > 1.0 > v[2]
    T0K, R = 1573.0, OM.R_GAS
    assert np.isclose(v[1], np.exp(3e5 / (R * T0K) * (T0K / 1273.0 - 1)), rtol=1e-14)
    # clamp to [1e-8, 1e8]
    v, _ = OM.rock_prop(np.array([0.0, 5000.0]), X[:2], 1.0, 1.0, 1300.0, Ea=1e6)
    assert v[0] == 1e8 and v[1] == 1e-8


def test_pressure_and_lower_mantle_factors_take_the_minimum(nodes):
    X, _ = nodes
    T = np.full(len(X), 1300.0)
    depth = np.linalg.norm(X, axis=1).max() - np.linalg.norm(X, axis=1)
    v, _ = OM.rock_prop(T, X, 1.0, 3300.0, 1300.0, Ea=3e5, Va=4e-6, g=-9.81, L0=1e6)
    fP = np.exp(np.abs(-9.81 * 3300.0 * 1e6 * depth) * 4e-6 / (OM.R_GAS * 1573.0))
    # without a lower-mantle factor the minimum with the T-only factor (1 here)
    # cancels the pressure effect (rock_prop.m:100-107 takes min([vALL2 vALL3]))
    assert np.all(v == 1.0) and fP.max() > 1.1
    v, _ = OM.rock_prop(T, X, 1.0, 3300.0, 1300.0, Ea=3e5, Va=4e-6, g=-9.81, L0=1e6, a_UM2LM=1.5)
    assert np.allclose(v, np.minimum(fP, 1.5), rtol=1e-14)


def test_cutoffs_and_kj_error(nodes):
    X, T = nodes
    v, d = OM.rock_prop(T, X, 1.0, 1.0, 650.0, alpha=1e-3, Ea=3e5, minVisc=0.5, maxVisc=2.0,
                        minDens=0.9, maxDens=1.1)
    assert v.min() >= 0.5 and v.max() <= 2.0 and d.min() >= 0.9 and d.max() <= 1.1
    with pytest.raises(ValueError):
        OM.rock_prop(T, X, 1.0, 1.0, 650.0, Ea=300.0)


def test_calculate_dt_closed_form(nodes):
    X, _ = nodes
    U = np.zeros_like(X)
    assert OM.calculate_dt(U, X, 0.5, 7.0) == 7.0
    U[3] = [0.0, -4.0, 1.0]
    L = X.max(axis=0) - X.min(axis=0)
    assert OM.calculate_dt(U, X, 0.5, 1e9) == pytest.approx(0.5 * min(L[1] / 4.0, L[2] / 1.0))
    assert OM.calculate_dt(U, X, 0.5, 1e-9) == 1e-9
