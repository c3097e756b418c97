"""GPU checks of the top-k inner-product backward (fv_topk_ip_backward behind
torch.autograd): against a dense FP64 torch reference on a 20,000-row
database (scores of every row, top-k, cross-entropy, autograd) and against
FP64 central finite differences (eps = 1e-3) of the loss with the selected
rows held fixed."""
import numpy as np
import pytest

from flashvector_amd import autograd as A
from flashvector_amd import lib as L
from flashvector_amd import synthetic as S

pytestmark = pytest.mark.gpu
TAU = 0.05


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a ROCm device"
    L.load()
    return t


def vmf_case(torch, n=20000, b=64, dim=768, seed=(41, 42, 43, 44)):
    mu = S.vmf_mean_directions(64, dim, seed[0], "cuda")
    x = S.vmf_points(mu, n, 50.0, seed[1]).bfloat16().float()   # BF16-valued rows
    q = S.vmf_points(mu, b, 50.0, seed[2])
    rng = np.random.default_rng(seed[3])
    return q, x, rng


def ce_loss64(q, x, ids, t, tau=TAU):
    """FP64 loss with the selected rows fixed: mean CE(softmax(s/tau), t)."""
    s = np.einsum("bd,bkd->bk", q, x[ids])
    z = s / tau
    m = z.max(1, keepdims=True)
    lse = m[:, 0] + np.log(np.exp(z - m).sum(1))
    return float(np.mean(lse - z[np.arange(len(t)), t]))


def test_fused_ce_matches_dense_fp64_reference(torch):
    k = 64
    q, x, rng = vmf_case(torch)
    t = torch.from_numpy(rng.integers(0, k, size=q.shape[0]).astype(np.int32)).cuda()
    qg = q.clone().requires_grad_(True)
    xg = x.clone().requires_grad_(True)
    loss = A.topk_softmax_ce(qg, xg, k, t, TAU)
    loss.backward()
    torch.cuda.synchronize()
    # Dense reference: scores of all rows in FP64, top-k, CE, autograd.
    q64 = q.double().cpu().requires_grad_(True)
    x64 = x.double().cpu().requires_grad_(True)
    s = q64 @ x64.T
    top = torch.topk(s.detach(), k, dim=1).indices
    z = s.gather(1, top) / TAU
    ref = (torch.logsumexp(z, 1) - z.gather(1, t.long().cpu()[:, None])[:, 0]).mean()
    ref.backward()
    assert abs(loss.item() - ref.item()) <= 1e-4 * abs(ref.item())
    np.testing.assert_allclose(qg.grad.cpu().numpy(), q64.grad.numpy(), rtol=2e-3, atol=2e-6)
    dx, dx_ref = xg.grad.cpu().numpy(), x64.grad.numpy()
    touched = np.unique(top.numpy())
    np.testing.assert_allclose(dx[touched], dx_ref[touched], rtol=2e-3, atol=2e-6)
    mask = np.ones(x.shape[0], bool)
    mask[touched] = False
    assert not dx[mask].any()


def test_fused_ce_matches_fp64_finite_differences(torch):
    k = 64
    q, x, rng = vmf_case(torch, n=8000, b=16, seed=(51, 52, 53, 54))
    t = rng.integers(0, k, size=q.shape[0]).astype(np.int32)
    qg = q.clone().requires_grad_(True)
    xg = x.clone().requires_grad_(True)
    scores, ids = A.topk_inner_product(qg, xg, k)
    torch.cuda.synchronize()
    ids_np = ids.cpu().numpy().astype(np.int64)
    loss = A.topk_softmax_ce(qg, xg, k, torch.from_numpy(t).cuda(), TAU)
    loss.backward()
    dq, dx = qg.grad.cpu().numpy(), xg.grad.cpu().numpy()
    q64, x64 = q.double().cpu().numpy(), x.double().cpu().numpy()
    eps = 1e-3
    for _ in range(12):
        i, d = rng.integers(0, q64.shape[0]), rng.integers(0, q64.shape[1])
        qp, qm = q64.copy(), q64.copy()
        qp[i, d] += eps
        qm[i, d] -= eps
        fd = (ce_loss64(qp, x64, ids_np, t) - ce_loss64(qm, x64, ids_np, t)) / (2 * eps)
        assert abs(dq[i, d] - fd) <= 1e-3 * abs(fd) + 1e-6, (i, d, dq[i, d], fd)
    rows = np.unique(ids_np)
    for _ in range(12):
        r, d = rows[rng.integers(0, len(rows))], rng.integers(0, x64.shape[1])
        xp, xm = x64.copy(), x64.copy()
        xp[r, d] += eps
        xm[r, d] -= eps
        fd = (ce_loss64(q64, xp, ids_np, t) - ce_loss64(q64, xm, ids_np, t)) / (2 * eps)
        assert abs(dx[r, d] - fd) <= 1e-3 * abs(fd) + 1e-6, (r, d, dx[r, d], fd)


def test_generic_backward_with_upstream_gradient(torch):
    k = 32
    q, x, rng = vmf_case(torch, n=5000, b=40, dim=256, seed=(61, 62, 63, 64))
    qg = q.clone().requires_grad_(True)
    xg = x.clone().requires_grad_(True)
    scores, ids = A.topk_inner_product(qg, xg, k)
    g = torch.from_numpy(rng.normal(size=(40, k)).astype(np.float32)).cuda()
    (scores * g).sum().backward()
    torch.cuda.synchronize()
    ids_np = ids.cpu().numpy().astype(np.int64)
    g64, q64, x64 = g.double().cpu().numpy(), q.double().cpu().numpy(), x.double().cpu().numpy()
    dq_ref = np.einsum("bk,bkd->bd", g64, x64[ids_np])
    dx_ref = np.zeros_like(x64)
    np.add.at(dx_ref, ids_np.ravel(), (g64[:, :, None] * q64[:, None, :]).reshape(-1, 256))
    np.testing.assert_allclose(qg.grad.cpu().numpy(), dq_ref, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(xg.grad.cpu().numpy(), dx_ref, rtol=1e-4, atol=1e-5)
    # The scores are the exact inner products of the returned rows.
    s_ref = np.einsum("bd,bkd->bk", q64, x64[ids_np])
    np.testing.assert_allclose(scores.detach().cpu().numpy(), s_ref, rtol=1e-5, atol=1e-6)


def test_fewer_rows_than_k_skips_empty_slots(torch):
    k = 64
    q, x, rng = vmf_case(torch, n=40, b=8, dim=128, seed=(71, 72, 73, 74))
    t = torch.from_numpy(rng.integers(0, 40, size=8).astype(np.int32)).cuda()
    qg = q.clone().requires_grad_(True)
    xg = x.clone().requires_grad_(True)
    scores, ids = A.topk_inner_product(qg, xg, k)
    assert (ids[:, 40:].cpu().numpy() == -1).all()
    loss = A.topk_softmax_ce(qg, xg, k, t, TAU)
    loss.backward()
    torch.cuda.synchronize()
    assert torch.isfinite(qg.grad).all() and torch.isfinite(xg.grad).all()
    q64 = q.double().cpu().requires_grad_(True)
    x64 = x.double().cpu().requires_grad_(True)
    s = q64 @ x64.T
    top = torch.topk(s.detach(), 40, dim=1).indices
    z = s.gather(1, top) / TAU
    ref = (torch.logsumexp(z, 1) - z.gather(1, t.long().cpu()[:, None])[:, 0]).mean()
    ref.backward()
    np.testing.assert_allclose(qg.grad.cpu().numpy(), q64.grad.numpy(), rtol=2e-3, atol=2e-6)
    np.testing.assert_allclose(xg.grad.cpu().numpy(), x64.grad.numpy(), rtol=2e-3, atol=2e-6)
