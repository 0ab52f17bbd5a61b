"""The grid generator's security checks (grid_generator/src/coriolis.c:188-194,
393-427; rhombus_averaging.c:262-265) and their bounds: the reference's 1e-10
by default, a looser bound only where it is asked for (the res-8 scale grid),
and a clear error when a check fails."""
import pytest

from game_archived_amd.grid import Grid


def test_default_bounds_pass_at_res3():
    g = Grid(3, n_layers=5, n_oro_layers=4, n_lloyd=0, cache=False)
    assert g.dims.nsh == 642


def test_trsk_kite_area_check_reports_the_failing_ratio():
    # no generated grid sums its kite areas to the cell area within 1e-18
    with pytest.raises(RuntimeError) as e:
        Grid(3, n_layers=5, n_oro_layers=4, n_lloyd=0, cache=False, trsk_tolerance=1e-18)
    msg = str(e.value)
    assert "(30)" in msg and "kite areas / cell area" in msg and "bound 1e-18" in msg


def test_rhombus_weight_check_fails_at_a_tiny_bound():
    with pytest.raises(RuntimeError) as e:
        Grid(3, n_layers=5, n_oro_layers=4, n_lloyd=0, cache=False, rhombus_tolerance=1e-18)
    assert "rhombus weights failed" in str(e.value)
