"""The streaming-copy entry point behind the measured HBM ceiling
(mg_stream_copy): byte-exact copies of every size class the kernel's grid
meets (a partial pass of 4 x 16-byte words per thread, a partial grid, a
grid-stride loop), and its argument checks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nbytes", [16, 4096 + 48, 3 * 2**20 + 16, 2**28 + 1024 * 16 + 32])
def test_stream_copy_is_exact(nbytes):
    import torch
    import cassia_amd
    g = torch.Generator(device="cpu").manual_seed(nbytes)
    src = torch.randint(0, 256, (nbytes,), dtype=torch.uint8, generator=g).cuda()
    dst = torch.zeros(nbytes + 64, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    ms = cassia_amd.stream_copy(dst.data_ptr(), src.data_ptr(), nbytes, 2)
    assert ms > 0
    out = dst.cpu().numpy()
    assert np.array_equal(out[:nbytes], src.cpu().numpy())
    assert not out[nbytes:].any()  # nothing written past the copy


def test_stream_copy_rejects_bad_sizes():
    import torch
    import cassia_amd
    a = torch.zeros(64, dtype=torch.uint8, device="cuda")
    for nbytes, repeats in ((0, 1), (24, 1), (32, 0)):
        with pytest.raises(cassia_amd.MoldError):
            cassia_amd.stream_copy(a.data_ptr(), a.data_ptr(), nbytes, repeats)
