"""Host-side validation of the arrays the Python mirror hands to the C-ABI.

The C entry points read raw float32 pointers of a fixed length, so the mirror
must convert or reject anything else before a pointer crosses: a float64
numpy array read as float32 is garbage, and a weight vector shorter than
n_vis would be read past its end.  No GPU is needed for these checks.
"""
import numpy as np
import pytest

from viss_recon_shade_amd.api import _f32_input


def test_float64_numpy_is_converted_not_reinterpreted():
    x = np.linspace(0.0, 1.0, 12)            # float64
    y = _f32_input(x, (12,), "w")
    assert y.dtype == np.float32 and y.flags["C_CONTIGUOUS"]
    assert np.array_equal(y, x.astype(np.float32))


def test_non_contiguous_numpy_is_made_contiguous():
    x = np.arange(24, dtype=np.float32).reshape(4, 6)[:, ::2]
    y = _f32_input(x, (4, 3), "I")
    assert y.flags["C_CONTIGUOUS"] and np.array_equal(y, x)


@pytest.mark.parametrize("shape,want", [((7,), (8,)), ((8, 2), (8,)), ((3, 8, 2), (None, 9, 2))])
def test_wrong_shape_rejected(shape, want):
    with pytest.raises(ValueError):
        _f32_input(np.zeros(shape, np.float32), want, "x")


def test_short_weight_vector_rejected():
    with pytest.raises(ValueError, match="w must have shape"):
        _f32_input(np.ones(99, np.float32), (100,), "w")


def test_complex_and_object_arrays_rejected():
    with pytest.raises(TypeError):
        _f32_input(np.zeros(4, np.complex64), (4,), "V")
    with pytest.raises(TypeError):
        _f32_input(np.array(["a", "b"]), (2,), "V")


def test_torch_tensors_must_be_float32_and_contiguous():
    torch = pytest.importorskip("torch")
    with pytest.raises(TypeError):
        _f32_input(torch.zeros(4, dtype=torch.float64), (4,), "I")
    with pytest.raises(ValueError):
        _f32_input(torch.zeros(4, 6)[:, ::2], (4, 3), "I")
    t = torch.zeros(4, 3)
    assert _f32_input(t, (4, None), "I") is t   # used in place, no copy
