"""Argument checks of the serving loop that need no GPU."""

import pytest

from sovereign_solver_amd.lp.serve import FamilyServer


def test_model_and_features_go_together():
    with pytest.raises(ValueError, match="features"):
        FamilyServer([], model=object())
    with pytest.raises(ValueError, match="features"):
        FamilyServer([], features=lambda ps: None)
