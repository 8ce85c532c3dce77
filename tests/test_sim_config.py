"""Host-side validation of the simulation configuration (reference config.py:74-100 rejects
unknown solvers/presets and non-physical values with an error before anything runs). CPU only."""

from __future__ import annotations

import pytest

from galaxy_collision_amd.sim import SimConfig


def test_accepts_every_solver_and_preset():
    from galaxy_collision_amd.solver import SOLVER_NAMES

    for name in SOLVER_NAMES:
        for preset in ("plummer", "two_galaxy"):
            assert SimConfig(solver=name, ic_preset=preset).solver == name


@pytest.mark.parametrize("field,value", [("solver", "conjugate_gradient"), ("ic_preset", "disk"),
                                         ("dt", -0.01), ("dt", 0.0), ("steps", -1),
                                         ("n_particles", 0), ("grid_size", 2), ("dx", 0.0),
                                         ("output_cadence", -1), ("reorder_every", -4)])
def test_rejects_bad_values(field, value):
    with pytest.raises(ValueError):
        SimConfig(**{field: value})
