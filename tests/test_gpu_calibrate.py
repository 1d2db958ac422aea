"""The streaming calibration kernels (the achievable-bandwidth ceiling of the bench) compute what
they claim: per-workgroup sums of the array (read) and an exact copy (GPU)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [2, 130, 1 << 20, (1 << 22) + 6])
@pytest.mark.parametrize("nt", [0, 1])
def test_stream_read_and_copy(require_gpu, n, nt):
    from cvfes_amd._lib import check, lib
    h = lib()
    g = torch.Generator(device="cuda").manual_seed(n)
    src = torch.rand(n, dtype=torch.float64, device="cuda", generator=g)
    nb = int(h.cvfes_stream_blocks(n))
    assert nb >= 8 and nb % 8 == 0
    part = torch.full((nb,), np.nan, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    check(h.cvfes_stream_read(st, src.data_ptr(), n, nt, part.data_ptr()))
    dst = torch.zeros_like(src)
    check(h.cvfes_stream_copy(st, src.data_ptr(), n, nt, dst.data_ptr()))
    torch.cuda.synchronize()
    assert torch.equal(dst, src)
    p = part.cpu().numpy()
    assert np.all(np.isfinite(p))
    total = float(src.sum())
    assert abs(p.sum() - total) <= 1e-12 * n


def test_stream_kernels_refuse_bad_arguments(require_gpu):
    from cvfes_amd._lib import CvfesError, check, lib
    h = lib()
    src = torch.zeros(8, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    with pytest.raises(CvfesError, match="even"):
        check(h.cvfes_stream_read(st, src.data_ptr(), 7, 1, src.data_ptr()))
    with pytest.raises(CvfesError, match="aligned"):
        check(h.cvfes_stream_copy(st, src.data_ptr() + 8, 4, 1, src.data_ptr() + 64))
