"""Pins the oracle by what the physics fixes (tests/physics_checks.py), run
on the CPU in the reference's own stream mode.  The oracle is also pinned
bit for bit to the reference build (test_oracle_vs_ref.py); these checks
hold independently of that build."""
import pytest

import physics_checks as P


@pytest.mark.parametrize("check", P.ALL, ids=[c.__name__ for c in P.ALL])
def test_oracle_physics(oracle_lib, check):
    check(oracle_lib)
