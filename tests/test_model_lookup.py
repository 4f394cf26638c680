"""Background-layer lookup and complex conductivity (P1): the host model (cartesian_model) against
the oracle, and both against hand-worked cases of the reference's rules.

Reference rules (depths are System.Decimal there, so sums and comparisons are exact):
  ModelUtils.FindCorrespondingBackgroundLayerIndex (Extreme.Cartesian/Model/ModelUtils.cs:35-52): the
  first layer whose closed interval [ss, es] holds [start, end]; none -> InvalidOperationException.
  ScalarMathUtils.DetermineCorrespondingBackgroundLayer (Green/Scalar/Impl/ScalarMathUtils.cs:15-31):
  half-open [d1, d2), so a depth on an interface belongs to the layer below it.
  zeta = sigma - i omega eps0 (Model/Omega/OmegaModelUtils.cs:41-42).
"""
import math

import numpy as np
import pytest

from extreme_amd import cartesian_model as cm


def _c5_section():
    # air | 100 ohm-m 0-2 km | 10 ohm-m 2-10 km | 1000 ohm-m slab 10 km - bottom | half-space
    return cm._three_layer_section(14503.969732826352)


@pytest.mark.parametrize("make", [cm.config_c1, cm.config_c2, cm.config_c5])
def test_zeta0_and_indices_match_oracle(oracle, make):
    m = make()
    omega = cm.omega_of(m.frequencies[-1])
    got = m.background_indices()
    ref = [oracle.background_layer(m.section, d, d + t) for d, t in zip(m.depth, m.thick)]
    assert got == ref
    z0 = m.zeta0(omega)
    for k, i in enumerate(ref):
        assert z0[k] == oracle.zeta(m.section.sigma[i], omega)
    # zeta of every background layer, air included (sigma = 0 gives -i omega eps0)
    zs = m.section.zeta(omega)
    for k, s in enumerate(m.section.sigma):
        assert zs[k] == oracle.zeta(s, omega)
    assert zs[0] == -1j * omega * (1.0 / (4e-7 * math.pi * 299792458.0 ** 2))


def test_half_open_interfaces(oracle):
    sec = _c5_section()
    cases = {0.0: 1, 1999.999: 1, 2000.0: 2, 9999.5: 2, 10000.0: 3}
    for z, k in cases.items():
        assert sec.layer_at_depth(z) == k, z
        assert oracle.layer_at_depth(sec, z) == k, z
    # above the zero level and below the last finite layer nothing holds the depth
    for z in (-1.0, sum(sec.thickness) + 1.0):
        assert oracle.layer_at_depth(sec, z) == -1
        with pytest.raises(ValueError):
            sec.layer_at_depth(z)


def test_nesting_rule(oracle):
    sec = _c5_section()
    # closed intervals: an anomaly layer may end on the interface below it and start on the one above
    assert sec.background_layer(1500.0, 2000.0) == 1 == oracle.background_layer(sec, 1500.0, 2000.0)
    assert sec.background_layer(2000.0, 2500.0) == 2 == oracle.background_layer(sec, 2000.0, 2500.0)
    # an anomaly layer across an interface has no background layer: the reference throws
    assert oracle.background_layer(sec, 1900.0, 2100.0) == -1
    with pytest.raises(ValueError):
        sec.background_layer(1900.0, 2100.0)
    # the lower half-space (thickness 0) holds nothing
    bottom = sum(sec.thickness)
    assert oracle.background_layer(sec, bottom, bottom + 10.0) == -1


def test_decimal_sums():
    """2.2 + 99.9 is 102.1 in decimal but 102.10000000000001 in binary: the reference accepts this
    layout, and so must the lookups."""
    assert 2.2 + 99.9 > 102.1
    sec = cm.Section1D(0.0, [0.0, 102.1, 0.0], [0.0, 1e-2, 1e-3])
    assert sec.background_layer(2.2, 2.2 + 99.9) == 1
    # and a real overlap by one micrometre is still refused
    with pytest.raises(ValueError):
        sec.background_layer(2.2, 102.100001)


def test_decimal_sums_oracle(oracle):
    sec = cm.Section1D(0.0, [0.0, 102.1, 0.0], [0.0, 1e-2, 1e-3])
    assert oracle.background_layer(sec, 2.2, 2.2 + 99.9) == 1
    assert oracle.background_layer(sec, 2.2, 102.100001) == -1
    # a section whose interface is itself a decimal sum: 0.1 + 0.2 is the interface, not above it
    sec2 = cm.Section1D(0.0, [0.0, 0.1, 0.2, 5.0, 0.0], [0.0, 1e-2, 1e-1, 1e-3, 1e-3])
    assert 0.1 + 0.2 != 0.3
    assert oracle.layer_at_depth(sec2, 0.3) == 3 == sec2.layer_at_depth(0.3)
    assert oracle.background_layer(sec2, 0.1, 0.3) == 2 == sec2.background_layer(0.1, 0.3)


def test_zeta_closed_form():
    omega = cm.omega_of(0.1)
    eps0 = 1.0 / (4e-7 * math.pi * 299792458.0 ** 2)
    np.testing.assert_array_equal(cm.zeta_of([0.0, 0.01], omega), np.array([-1j * omega * eps0, 0.01 - 1j * omega * eps0]))
