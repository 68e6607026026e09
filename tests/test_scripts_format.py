"""Output formatting of the script layer (no GPU): the DO_SAVE lines write
floats as Julia's string interpolation does."""
import pytest

from ptstokes.scripts import julia_float, _sig


@pytest.mark.parametrize("x,want", [
    (0.0, "0.0"), (1.0, "1.0"), (100.0, "100.0"), (123456.0, "123456.0"), (1234567.0, "1.234567e6"),
    (1e6, "1.0e6"), (1e-5, "1.0e-5"), (1e-4, "0.0001"), (0.00012345, "0.00012345"), (-2.5e7, "-2.5e7"),
    (1.5e-10, "1.5e-10"), (9.661317264, "9.661317264"), (0.1, "0.1"), (1e16, "1.0e16"), (-0.0, "-0.0"),
    (float("nan"), "NaN"), (float("inf"), "Inf"),
])
def test_julia_float(x, want):
    assert julia_float(x) == want


def test_round_sigdigits():
    assert _sig(3932.58, 2) == 3900.0 and _sig(3932.58, 3) == 3930.0 and _sig(0.012345, 2) == 0.012
