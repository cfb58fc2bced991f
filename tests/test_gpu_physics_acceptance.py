"""GPU: the orbit, matrix and acceptance drivers on the engine, against the
same known answers as tests/test_physics_acceptance.py (there with the
oracle tracking)."""

import pytest

from test_physics_acceptance import (  # noqa: F401  (collected here with the engine)
    hmba,
    test_1d_acceptance_known_answers,
    test_2d_recursive_acceptance_known_answer,
    test_find_elem_m66_known_answer,
    test_find_m44_symplectic,
    test_find_m66_known_answer,
    test_find_orbit4_batch_equals_one_by_one,
    test_find_orbit6_batch_equals_one_by_one,
    test_find_orbit4_zero_and_propagation,
    test_find_orbit6_known_answer,
    test_grid_modes_bracket_the_recursive_boundary,
    test_orbit_max_iterations_warns,
    test_recursive_with_no_survivor_warns,
)

pytestmark = pytest.mark.gpu


@pytest.fixture
def hooks():
    """No hooks: every tracking call goes to the engine."""
    return None
