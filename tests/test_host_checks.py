"""Host-side argument checks of the Python front end (no GPU): caller-supplied output arrays must be writeable,
C-contiguous float64 arrays of exactly the size the library writes, or the call raises ValueError before any native
code runs."""
import numpy as np
import pytest

from classicaldft_amd import cdft as C


def test_output_array_accepts_the_right_array():
    a = np.zeros((4, 5, 6))
    assert C._out(a, 120, (4, 5, 6)) is a
    assert C._out(None, 120, (4, 5, 6)).shape == (4, 5, 6)


@pytest.mark.parametrize("bad", [
    np.zeros((4, 5, 6), dtype=np.float32),    # wrong dtype
    np.zeros((4, 5, 5)),                       # wrong size
    np.zeros((6, 5, 4)).transpose(),           # not C-contiguous
    np.zeros(240)[::2],                        # strided view
    [0.0] * 120,                               # not an ndarray
])
def test_output_array_rejects_unsafe_arrays(bad):
    with pytest.raises(ValueError):
        C._out(bad, 120, (4, 5, 6))


def test_output_array_rejects_read_only():
    a = np.zeros(120)
    a.flags.writeable = False
    with pytest.raises(ValueError):
        C._out(a, 120, (120,))
