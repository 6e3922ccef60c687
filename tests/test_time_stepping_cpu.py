"""Step selection of solve(): shortened steps land exactly on save times and
on the end of tspan, at any time scale (no absolute snapping tolerance)."""
import pytest


@pytest.fixture(scope="module")
def select_step(th):
    return th.time_integration.select_step


def test_plain_step_is_not_changed(select_step):
    assert select_step(0.0, 0.1, 1.0, [None]) == (0.1, 0.1)
    assert select_step(0.3, 0.1, 1.0, [0.9]) == (0.1, 0.3 + 0.1)


def test_step_shortened_to_end_of_tspan(select_step):
    h, t = select_step(0.95, 0.1, 1.0)
    assert t == 1.0 and h == pytest.approx(0.05, rel=1e-12)


def test_step_lands_on_save_time_exactly(select_step):
    h, t = select_step(0.1, 0.3, 1.0, [0.25])
    assert t == 0.25 and h == pytest.approx(0.15, rel=1e-12)
    # the earliest of several save times wins, and t1 if it comes first
    assert select_step(0.0, 1.0, 0.5, [0.7, 0.6])[1] == 0.5
    assert select_step(0.0, 1.0, 2.0, [0.7, 0.6])[1] == 0.6


@pytest.mark.parametrize("scale", [1e-15, 1e-9, 1.0, 1e6])
def test_small_and_large_time_scales(select_step, scale):
    # steps of 0.3 scale with a save time at 0.5 scale: the second step is
    # shortened to land on it exactly, never snapped from a neighbouring step
    t, t1, ts = 0.0, 1.0 * scale, 0.5 * scale
    seen = []
    while t < t1:
        h, t = select_step(t, 0.3 * scale, t1, [ts if t < ts else None])
        assert h > 0
        seen.append(t)
    assert ts in seen
    assert seen[-1] == t1
    assert all(a < b for a, b in zip(seen, seen[1:]))


def test_save_time_on_t1(select_step):
    h, t = select_step(0.8, 0.3, 1.0, [1.0])
    assert t == 1.0 and h == pytest.approx(0.2, rel=1e-12)
