"""rvk_post_shade_noalloc (GPU): the allocation-free post pass against
rvk_post_shade, through the C-ABI, on two calls in a row (the second on a
smaller image, so it reads the scratch array the first one left behind), and
its size limit."""
import numpy as np
import pytest

from robartsvtk_amd import abi

pytestmark = pytest.mark.gpu


def _inputs(rng, W, H):
    rgba = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    depth = (rng.random((H, W)) * 60.0).astype(np.float32)
    mind = (rng.random((H, W)) * 5.0).astype(np.float32)
    maxd = (20.0 + rng.random((H, W)) * 80.0).astype(np.float32)
    maxd[rng.random((H, W)) < 0.05] = 0.0  # pixels without a ray
    return rgba, depth, mind, maxd


def _shade(ctx, fn, W, H, inputs):
    import torch
    rgba, depth, mind, maxd = (torch.from_numpy(a.copy()).cuda() for a in inputs)
    sp = abi.ShadeParams(0.7, 0.002, 1.0 / 0.003, 0.25, 0.45, 1.0 / 0.15)
    abi.check(fn(ctx.handle, W, H, sp, depth.data_ptr(), mind.data_ptr(), maxd.data_ptr(), rgba.data_ptr()))
    torch.cuda.synchronize()
    return rgba.cpu().numpy(), depth.cpu().numpy()


def test_post_shade_noalloc_matches_post_shade(gpu_ctx):
    rng = np.random.default_rng(12)
    lib = abi.lib()
    for W, H in ((1280, 720), (97, 53)):
        inputs = _inputs(rng, W, H)
        ref_rgba, ref_depth = _shade(gpu_ctx, lib.rvk_post_shade, W, H, inputs)
        got_rgba, got_depth = _shade(gpu_ctx, lib.rvk_post_shade_noalloc, W, H, inputs)
        assert np.array_equal(got_rgba, ref_rgba), (W, H)
        assert np.array_equal(got_depth, ref_depth), (W, H)
        assert not np.array_equal(got_rgba, inputs[0])  # the shading changed the image


def test_post_shade_noalloc_rejects_images_over_its_limit(gpu_ctx):
    import torch
    t = torch.zeros(16, dtype=torch.float32, device="cuda")
    sp = abi.ShadeParams(0.7, 0.002, 1.0 / 0.003, 0.25, 0.45, 1.0 / 0.15)
    W, H = 4096, 1025  # one row past 4096^2 / 4 pixels
    rc = abi.lib().rvk_post_shade_noalloc(gpu_ctx.handle, W, H, sp, t.data_ptr(), t.data_ptr(), t.data_ptr(),
                                         t.data_ptr())
    assert rc == 1 and b"exceed" in abi.lib().rvk_last_error_string()
