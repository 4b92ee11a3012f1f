"""The oracle's passes, one at a time, against hand-worked answers, closed
forms, invariants and brute force (tests/oracle_pins.py)."""
import pytest

from tests import oracle_pins


@pytest.mark.parametrize("name", sorted(oracle_pins.PINS))
def test_pin(oracle_lib, name):
    oracle_pins.PINS[name](oracle_lib)
