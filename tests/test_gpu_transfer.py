"""Device -> host copies of large outputs through the pinned staging chunks (_native.to_host).

The public calls return every field through ``to_host``; fields above the
32 MB chunk size take the staged path (two pinned chunks, copies overlapped
with the host-side copy-out).  The result must be bit-identical to torch's
own device -> host copy for any length (whole chunks, a partial last chunk,
one element past a chunk), dtype and layout.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("no HIP device")
    return torch


@pytest.mark.parametrize("n", [(32 << 20) // 16 * 3, (32 << 20) // 16 * 3 + 1, 5_000_003])
def test_staged_copy_matches_torch(torch, n):
    from green_laplace_wave_reconstruction_amd._native import to_host

    print(f"staged copy of {n} complex128 values", flush=True)
    g = torch.Generator(device="cuda").manual_seed(n)
    x = torch.complex(torch.randn(n, dtype=torch.float64, device="cuda", generator=g),
                      torch.randn(n, dtype=torch.float64, device="cuda", generator=g))
    got = to_host(x)
    ref = x.cpu().numpy()
    assert got.dtype == ref.dtype and got.shape == ref.shape
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_staged_copy_shapes_and_views(torch):
    from green_laplace_wave_reconstruction_amd._native import to_host

    x = torch.arange(3 * 1100 * 1100, dtype=torch.float64, device="cuda").reshape(3, 1100, 1100)  # 29 MB
    y = torch.cat([x, x + 0.5], dim=0)                                                              # 58 MB
    got = to_host(y)
    assert got.shape == (6, 1100, 1100) and np.array_equal(got, y.cpu().numpy())
    t = y.transpose(1, 2)  # non-contiguous view: copied in its logical order
    assert np.array_equal(to_host(t), t.cpu().numpy())
    z = torch.complex(y, -y)[1:4]  # complex slice of a time series (58 MB: staged)
    assert np.array_equal(to_host(z), z.cpu().numpy())
    small = torch.complex(y, -y)[1]  # one instant (19 MB: torch's own copy)
    assert np.array_equal(to_host(small), small.cpu().numpy())
