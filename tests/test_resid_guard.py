"""The residual kernel's guard, checked on the host (no GPU): when the NaN
sentinel of b/dt is overwritten so that every slot looks live, the slots
the guard admits still read only inside the colour arrays and coefficient
tables.  The same check without the guard's interior test (the form that
once read one z-plane outside the array) must report reads outside, so the
check is known to see that failure."""
import pytest

from ib3d_amd import native

GRIDS = [(4, 4, 4), (5, 4, 7), (32, 24, 24), (48, 32, 32), (33, 17, 9), (128, 64, 64), (64, 48, 48)]


@pytest.mark.parametrize("n", GRIDS)
def test_guard_keeps_every_read_inside(n):
    assert native.resid_guard_selfcheck(*n) == 0


@pytest.mark.parametrize("n", GRIDS[:4])
def test_check_sees_a_guard_without_its_interior_test(n):
    assert native.resid_guard_selfcheck(*n, drop_interior=True) > 0
