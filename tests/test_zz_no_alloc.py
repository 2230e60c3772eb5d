"""rgs_loss_grad_dev_noalloc (photometric_loss_grad on device buffers, its
partial sums in a static scratch array) against rgs_loss_grad, the
reference-API entry point (cpu/backward.rs:646-693): dL/drgb bit-identical,
the loss to float rounding (the partial sums are added in another order),
on two calls in a row, the second with a smaller image, so the second reads
nothing the first left in the shared scratch."""
import numpy as np
import pytest

from rgsplat_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _images(w, h, seed):
    r = S.Stream(seed)
    tgt = np.concatenate([np.clip(r.normal((h, w, 3)) * 0.2 + 0.5, 0, 1), np.ones((h, w, 1))], -1)
    tgt = tgt.astype(np.float32)
    rgb = np.clip(tgt[..., :3] + r.normal((h, w, 3)) * 0.05, 0, 1).astype(np.float32)
    rgb[h // 4:h // 2, w // 4:w // 2] = tgt[h // 4:h // 2, w // 4:w // 2, :3]  # zero residuals: signum(+0) = +1
    return rgb, tgt


def test_loss_grad_noalloc_matches_loss_grad():
    import torch
    from rgsplat_amd import runtime
    dev = torch.device("cuda", 0)
    ctx = runtime.Context(0)
    try:
        ctx.set_stream(torch.cuda.current_stream(dev).cuda_stream)
        for (w, h), seed in [((1920, 1080), 1), ((333, 77), 2)]:
            rgb, tgt = _images(w, h, seed)
            loss_ref, dl_ref = ctx.photometric_loss_grad(rgb, tgt, w, h, 0.2)
            rgb_d = torch.from_numpy(rgb).to(dev)
            tgt_d = torch.from_numpy(tgt).to(dev)
            dl_d = torch.full((w * h * 3,), float("nan"), device=dev)
            loss_d = torch.full((1,), float("nan"), device=dev)
            ctx.loss_grad_dev_noalloc(rgb_d.data_ptr(), tgt_d.data_ptr(), w, h, 0.2, loss_d.data_ptr(),
                                      dl_d.data_ptr())
            torch.cuda.synchronize(dev)
            dl = dl_d.cpu().numpy()
            loss = float(loss_d.cpu().numpy()[0])
            assert np.array_equal(dl.view(np.uint32), dl_ref.view(np.uint32)), \
                f"{w}x{h}: {np.count_nonzero(dl.view(np.uint32) != dl_ref.view(np.uint32))} dL/drgb values differ"
            assert abs(loss - loss_ref) <= 1e-6 * abs(loss_ref), (w, h, loss, loss_ref)
        # beyond the scratch: an error, not a launch
        with pytest.raises(runtime.RgsError):
            ctx.loss_grad_dev_noalloc(rgb_d.data_ptr(), tgt_d.data_ptr(), 32768, 32768, 0.2, loss_d.data_ptr(),
                                      dl_d.data_ptr())
    finally:
        ctx.set_stream(None)
        ctx.close()
