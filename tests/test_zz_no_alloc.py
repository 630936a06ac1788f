"""rg_postprocess_device_noalloc through the C ABI against
rg_postprocess_device (same kernels, so bit-identical images): two calls in a
row, the second on a smaller image, with and without bloom; sizes beyond the
fixed scratch fail with RG_E_BUDGET and leave the images untouched."""
import sys

import numpy as np
import pytest

import rgapi

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def field(w, h, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((h, w, 4), np.float32)
    x[..., :3] = rng.exponential(0.4, (h, w, 3)).astype(np.float32)
    x[h // 3, w // 2, :3] = 40.0           # a bright source the bloom spreads
    x[..., 3] = 1.0
    return x


@pytest.fixture(scope="module")
def ctx():
    return rgapi.Context(0)


def run(ctx, fn, img, bloom, want_ldr=True):
    import torch

    h, w = img.shape[:2]
    d_hdr = torch.from_numpy(img.copy()).to("cuda:0")
    d_ldr = torch.zeros_like(d_hdr) if want_ldr else None
    torch.cuda.synchronize()
    fn(ctx, w, h, d_hdr.data_ptr(), bloom, d_ldr.data_ptr() if want_ldr else None)
    ctx.synchronize()
    return d_hdr.cpu().numpy(), (d_ldr.cpu().numpy() if want_ldr else None)


@pytest.mark.parametrize("size", [(1920, 1080), (333, 217)])
def test_noalloc_equals_device_postprocess_twice(ctx, size):
    w, h = size
    for k, (ww, hh) in enumerate([(w, h), (w // 3 + 1, h // 2 + 1)]):    # second call: smaller image
        img = field(ww, hh, 7 + k)
        for bloom in (rgapi.default_bloom(), rgapi.default_bloom(threshold=0.5, radius=0.05, curve=2.0), None):
            say(f"[noalloc] {ww}x{hh} bloom={'on' if bloom is not None else 'off'}")
            a_hdr, a_ldr = run(ctx, rgapi.postprocess_device, img, bloom)
            b_hdr, b_ldr = run(ctx, rgapi.postprocess_device_noalloc, img, bloom)
            assert np.array_equal(a_hdr, b_hdr)
            assert np.array_equal(a_ldr, b_ldr)
            if bloom is not None:
                assert a_hdr[..., :3].sum() > img[..., :3].sum()


def test_noalloc_rejects_sizes_beyond_the_scratch(ctx):
    img = field(1100, 1000, 3)              # size_aspect 1: 1.1M glow pixels > 1,048,576
    with pytest.raises(rgapi.RGError) as e:
        run(ctx, rgapi.postprocess_device_noalloc, img, rgapi.default_bloom(size_aspect=1.0))
    assert e.value.code == -7 and "scratch" in str(e.value)
    small = field(64, 48, 4)                # a blur wider than 127 taps per side
    with pytest.raises(rgapi.RGError) as e:
        run(ctx, rgapi.postprocess_device_noalloc, small, rgapi.default_bloom(size_aspect=1.0, radius=1.0))
    assert e.value.code == -7
    # the same sizes still work on the allocating entry point, and the
    # no-allocation one keeps working after the refusals
    run(ctx, rgapi.postprocess_device, small, rgapi.default_bloom(size_aspect=1.0, radius=1.0))
    a, _ = run(ctx, rgapi.postprocess_device, small, rgapi.default_bloom())
    b, _ = run(ctx, rgapi.postprocess_device_noalloc, small, rgapi.default_bloom())
    assert np.array_equal(a, b)
