"""Pins of the oracle's nonlinear IT step and RT potential phase
(oracle/gpe_oracle.c, restating src/solver/cpu_solver/cpu_solver.cpp:50-131)
against the energy functional it restates from cpu_solver.cpp:137-194 and
against identities the physics fixes.  The checks live in _nonlinear_pins.py,
so that test_oracle_mutations.py can run them on copies of the oracle with
planted mistakes."""
import pytest

import _nonlinear_pins as pins


@pytest.mark.parametrize("pin", pins.ALL_PINS, ids=lambda f: f.__name__)
def test_nonlinear_pin(oracle, pin):
    pin(oracle.lib())


@pytest.mark.parametrize("shape", [(6, 5, 4), (8, 7, 6)])
def test_it_energy_gradient_other_shapes(oracle, shape):
    pins.pin_it_step_is_energy_gradient(oracle.lib(), shape=shape, seed=sum(shape))
