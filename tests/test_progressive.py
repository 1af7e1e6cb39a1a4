"""progressive_batch against the reference's own test (path_tracer.rs:1045-1053) and the frame seed rule."""
import pytest

from toaster_amd.progressive import progressive_batch


def test_progressive_batch_clamps_to_the_target():
    assert progressive_batch(2, 0, 5) == 2
    assert progressive_batch(2, 2, 5) == 2
    assert progressive_batch(2, 4, 5) == 1
    assert progressive_batch(2, 5, 5) is None
    with pytest.raises(ValueError):
        progressive_batch(0, 0, 5)
    with pytest.raises(ValueError):
        progressive_batch(1, 6, 5)


def test_render_frame_seed():  # path_tracer.rs:891-897
    from toaster_amd.progressive import render_frame_seed
    assert [render_frame_seed(f, "animated") for f in range(3)] == [0, 1, 2]
    assert [render_frame_seed(f, "static_independent") for f in range(3)] == [0, 0, 0]
    with pytest.raises(ValueError):
        render_frame_seed(0, "looping")
