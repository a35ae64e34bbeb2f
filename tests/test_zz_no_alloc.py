"""The allocation-free score entry point (mch_score_gradients_noalloc, static
device scratch) against mch_score_gradients with a caller-provided partial
buffer, through the C ABI: bit-identical scores, gradients and template
adjoints on two calls in a row, the second with a smaller batch."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _inputs(eng, dev, B, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    N = eng.N
    vols = torch.randn((B, N, N, N), generator=g).to(dev)
    omega = (0.5 * torch.randn((B, 3), generator=g)).to(dev)
    shift = (2.0 * torch.rand((B, 3), generator=g) - 1.0).to(dev)
    aT = torch.complex(torch.randn(eng.ne, generator=g), torch.randn(eng.ne, generator=g)).to(dev)
    return omega, shift, aT, vols


def _call(eng, fn, B, angles, jr, coeffs, sigma, partial):
    from matcha_amd import _lib
    dev = eng.device
    out = torch.full((B, 7), float("nan"), dtype=torch.float32, device=dev)
    G = torch.full((B, eng.ne), float("nan"), dtype=torch.complex64, device=dev)
    args = [eng.idx.data_ptr(), eng.lmax, eng.kdim, sigma.data_ptr(), sigma.stride(0), B, coeffs.data_ptr(),
            coeffs.stride(0), eng.L, angles.data_ptr(), jr.data_ptr()]
    if partial is not None:
        args.append(partial.data_ptr())
    args += [out.data_ptr(), G.data_ptr(), G.stride(0), _lib.stream_handle(dev)]
    _lib.check(getattr(_lib.lib(), fn)(*args), fn)
    torch.cuda.synchronize(dev)
    return out.cpu().numpy(), G.cpu().numpy()


def test_noalloc_matches_allocating_entry_point(gpu_device):
    from matcha_amd.diff import ScoreEngine
    eng = ScoreEngine(24, 8, 11, 7, 6, gpu_device)
    for B, seed in ((6, 1), (3, 2)):  # the second call is smaller: stale partials must not leak
        B_, angles, jr, coeffs, sigma = eng._prepare(*_inputs(eng, gpu_device, B, seed))
        assert B_ == B
        partial = torch.empty((B, eng.L + 1, 8), dtype=torch.float32, device=gpu_device)
        ref_out, ref_G = _call(eng, "mch_score_gradients", B, angles, jr, coeffs, sigma, partial)
        out, G = _call(eng, "mch_score_gradients_noalloc", B, angles, jr, coeffs, sigma, None)
        assert np.all(np.isfinite(ref_out)) and np.all(np.isfinite(ref_G))
        assert np.array_equal(out, ref_out)
        assert np.array_equal(G, ref_G)


def test_noalloc_rejects_batches_beyond_the_scratch(gpu_device):
    from matcha_amd import _lib
    rc = _lib.lib().mch_score_gradients_noalloc(None, 10, 1, None, 0, 200000, None, 0, 0, None, None, None, None,
                                                0, None)
    assert rc != 0
    assert b"static scratch" in _lib.lib().mch_last_error()
