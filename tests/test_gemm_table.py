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


def test_autotune_table_round_trip(tmp_path):
    """A saved table (choices that beat the built-in rule) installs the same number of
    choices; records that kept the built-in choice are not saved."""
    from dirac_caspt2_amd import autotune
    from dirac_caspt2_amd.api import gemm_clear_choices
    table = [{"opA": "N", "opB": "C", "M": 792, "N": 792, "K": 1728, "batch": 1, "variant": 3, "split": 1,
              "seconds": 1e-4, "builtin_seconds": 2e-4},
             {"opA": "C", "opB": "N", "M": 144, "N": 20736, "K": 924, "batch": 1, "variant": 1, "split": 2,
              "seconds": 5e-4, "builtin_seconds": 6e-4},
             {"opA": "T", "opB": "N", "M": 40, "N": 24, "K": 30, "batch": 5, "variant": 0, "split": 0,
              "seconds": 1e-5, "builtin_seconds": 1e-5}]
    path = tmp_path / "gemm_table.json"
    autotune.save(table, str(path))
    assert autotune.load(str(path)) == 2
    assert gemm_clear_choices() == 2
