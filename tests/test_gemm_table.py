"""The GEMM choice table of the autotuner (host-side only: no GPU needed): invalid choices
are refused, valid ones are counted and cleared."""
import pytest


def test_gemm_choice_rejects_bad_values():
    """variant 1..4 and split 1..64 only; clear reports how many choices it removed."""
    from dirac_caspt2_amd.api import gemm_clear_choices, gemm_set_choice
    with pytest.raises(RuntimeError):
        gemm_set_choice("N", "N", 10, 10, 10, 1, 5, 1)
    with pytest.raises(RuntimeError):
        gemm_set_choice("N", "N", 10, 10, 10, 1, 1, 0)
    gemm_set_choice("N", "N", 10, 10, 10, 1, 2, 4)
    assert gemm_clear_choices() == 1
