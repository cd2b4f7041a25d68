"""gl_ppr_update_normalized_f32_noalloc (the update with its partial sums in a static device
array instead of a workspace) against gl_ppr_update_normalized_f32 through the C ABI: the
same kernels, so X and the residuals must agree bit for bit, on two calls in a row (the
second smaller, so it runs over what the first left in the array), plus the size limit."""
import numpy as np
import pytest
import torch

from page_rank_amd import _native


def _case(n, K, seed, device):
    g = torch.Generator(device="cpu").manual_seed(seed)
    Y = torch.rand((n, K), generator=g).to(device)
    X = torch.rand((n, K), generator=g).to(device)
    src = torch.randint(0, n, (K,), generator=g, dtype=torch.int32).to(device)
    return Y, X, src


def test_size_limit_is_checked_before_any_device_work():
    L = _native.lib()
    rc = L.gl_ppr_update_normalized_f32_noalloc(8, 8, 8, 0.85, 1 << 20, 2048, 8, None)
    assert rc == _native.GL_ERR_INVALID and "static scratch" in _native.last_error()
    rc = L.gl_ppr_update_normalized_f32_noalloc(None, 8, 8, 0.85, 100, 2, 8, None)
    assert rc == _native.GL_ERR_INVALID


@pytest.mark.gpu
def test_noalloc_update_equals_workspace_update(device):
    L = _native.lib()
    st = _native.stream_handle(device)
    for n, K, seed in ((300_000, 24, 1), (5_000, 3, 2)):
        Y, X0, src = _case(n, K, seed, device)
        Xa, Xb = X0.clone(), X0.clone()
        ea = torch.full((K,), -1.0, device=device)
        eb = torch.full((K,), -2.0, device=device)
        wsb = L.gl_ppr_update_f32_workspace_bytes(n, K)
        ws = torch.empty(wsb, dtype=torch.uint8, device=device)
        with torch.cuda.device(device):
            _native.check(L.gl_ppr_update_normalized_f32(
                Y.data_ptr(), Xa.data_ptr(), src.data_ptr(), 0.85, n, K, ea.data_ptr(),
                ws.data_ptr(), wsb, st), "gl_ppr_update_normalized_f32")
            _native.check(L.gl_ppr_update_normalized_f32_noalloc(
                Y.data_ptr(), Xb.data_ptr(), src.data_ptr(), 0.85, n, K, eb.data_ptr(), st),
                "gl_ppr_update_normalized_f32_noalloc")
        torch.cuda.synchronize(device)
        assert torch.equal(Xa, Xb), (n, K)
        assert torch.equal(ea, eb), (n, K, ea, eb)
        want = np.abs(0.85 * Y.double().cpu().numpy() - X0.double().cpu().numpy())
        want[src.cpu().numpy(), np.arange(K)] = np.abs(
            0.85 * Y.double().cpu().numpy()[src.cpu().numpy(), np.arange(K)] + 0.15
            - X0.double().cpu().numpy()[src.cpu().numpy(), np.arange(K)])
        np.testing.assert_allclose(eb.cpu().numpy(), want.sum(0), rtol=1e-5)
