"""Host-side property naming (no GPU): the reference's public names map to
the tissue fields the kernels read (reference sequence/_parameters.py:234-252)."""

from __future__ import annotations

from torchsim_amd.model._simulator import PUBLIC_PROPERTIES
from torchsim_amd.sequence import _engine


def test_reference_public_names_map_to_tissue_fields():
    from torchsim_amd.simulators import MRFSimulator

    fields = MRFSimulator()._tissue_fields({
        "T1": 1.0, "T2": 2.0, "M0": 3.0, "B1": 4.0, "B1phase": 5.0, "B0": 6.0, "inv_efficiency": 7.0, "D": 8.0,
        "poolB_fraction": 9.0, "poolB_exchange": 10.0, "poolB_T1": 11.0, "poolB_T2": 12.0, "poolB_shift": 13.0,
    })
    assert fields == {
        "t1_ms": 1.0, "t2_ms": 2.0, "m0": 3.0, "b1": 4.0, "b1_phase_rad": 5.0, "b0_hz": 6.0,
        "inversion_efficiency": 7.0, "diffusion_um2_per_ms": 8.0, "pool_b_fraction": 9.0,
        "pool_b_exchange_rate": 10.0, "t1_pool_b_ms": 11.0, "t2_pool_b_ms": 12.0, "pool_b_shift_hz": 13.0,
    }


def test_every_public_name_is_a_tissue_field_the_kernels_read():
    """Every public name but the field spread (which reaches the samples,
    reference sequence/_parameters.py:164-173) is a kernel tissue field."""
    from torchsim_amd import _native
    from torchsim_amd.model._simulator import SAMPLE_FIELDS

    for name, field in PUBLIC_PROPERTIES.items():
        if name == "T2prime":
            assert field == "t2_prime_ms" and field in SAMPLE_FIELDS and field not in _native.TISSUE_INDEX
            continue
        assert field in _native.TISSUE_INDEX, (name, field)
    # a transmit phase and an off-resonance both route to the complex kernel
    assert _engine._FEATURE_OF["b1_phase_rad"] == _engine._FEATURE_OF["b0_hz"] == _native.FEATURE_OFF_AXIS
