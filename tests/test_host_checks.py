"""Host-side argument checks of the Python wrapper (holoflow_amd.hfx): a ring or output tensor
of the wrong dtype, shape, layout or size is rejected before any pointer reaches the C ABI,
which only receives a pointer and the ring's frame count.  CPU tensors only; no GPU."""
import pytest
import torch

from holoflow_amd.hfx import check_out, check_ring


def test_ring_checks():
    with pytest.raises(ValueError, match="uint16"):
        check_ring(torch.zeros(40, 64, 64, dtype=torch.float32), 64, 64, 32)
    with pytest.raises(ValueError, match="shape"):
        check_ring(torch.zeros(40, 64 * 64, dtype=torch.int16), 64, 64, 32)
    with pytest.raises(ValueError, match="shape"):
        check_ring(torch.zeros(40, 64, 32, dtype=torch.int16), 64, 64, 32)
    with pytest.raises(ValueError, match="fewer than the window"):
        check_ring(torch.zeros(16, 64, 64, dtype=torch.int16), 64, 64, 32)
    with pytest.raises(ValueError, match="contiguous"):
        check_ring(torch.zeros(40, 64, 64, dtype=torch.int16).transpose(1, 2), 64, 64, 32)
    with pytest.raises(ValueError, match="device"):
        check_ring(torch.zeros(40, 64, 64, dtype=torch.int16), 64, 64, 32)


def test_out_checks():
    with pytest.raises(ValueError, match="float32"):
        check_out(torch.zeros(3, 64, 64, dtype=torch.float64), 0, (3, 64, 64))
    with pytest.raises(ValueError, match="fewer than one window"):
        check_out(torch.zeros(3, 64, 32), 0, (3, 64, 64))
    with pytest.raises(ValueError, match="contiguous"):
        check_out(torch.zeros(3, 64, 64).transpose(1, 2), 0, (3, 64, 64))
    with pytest.raises(ValueError, match="shape"):
        check_out(torch.zeros(4, 3, 64, 64), 5, (3, 64, 64))            # one window short
    with pytest.raises(ValueError, match="shape"):
        check_out(torch.zeros(5, 3, 64, 32), 5, (3, 64, 64))            # wrong image size
    with pytest.raises(ValueError, match="device"):
        check_out(torch.zeros(5, 3, 64, 64), 5, (3, 64, 64))
