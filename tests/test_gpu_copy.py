"""The streaming-copy entry point behind the measured HBM ceiling
(mg_stream_copy): exact copies of every size class the kernel's grid meets
(a partial pass of 4 x 16-byte words per thread, a partial grid, a
grid-stride loop), nothing written past the end, and its argument checks."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nbytes", [16, 4096 + 48, 3 * 2**20 + 16, 2**28 + 1024 * 16 + 32])
def test_stream_copy_is_exact(nbytes):
    import cassia_amd
    ms, bad = cassia_amd.stream_copy(nbytes, 2)
    assert ms > 0
    assert bad == 0


def test_stream_copy_rejects_bad_sizes():
    import cassia_amd
    for nbytes, repeats in ((0, 1), (24, 1), (32, 0)):
        with pytest.raises(cassia_amd.MoldError):
            cassia_amd.stream_copy(nbytes, repeats)
