"""Differentiable top-k inner-product retrieval over the matrix-core flat scan.

The forward runs fv_flat_search (reduced-precision scan, exact FP32 rerank)
and returns the k best scores s_ij = q_i . x_{ids_ij} with their row ids; the
backward is the hand-written fv_topk_ip_backward kernel: dL/dq_i = sum_j
g_ij x_{ids_ij} and a scatter-add of g_ij q_i into the selected rows of
dL/dX. No Q x N score matrix is formed in either direction.

TopKInnerProduct: gradient flows from any downstream use of the scores.
TopKSoftmaxCE: mean softmax cross-entropy of s_i / tau against a target
position per query, fused: its backward recomputes the k scores of each query
from Q, X and the ids instead of keeping them.

The retrieval itself (which rows are selected) is piecewise constant in Q and
X and carries no gradient. No counterpart in the reference, which has no
training path.
"""
from __future__ import annotations

import ctypes

import torch

from . import lib as L
from . import search as F


def _split_device(out):
    """int64 [nq, k] GpuSearchResult -> (ids int32, distances float32) on the device."""
    pairs = out.view(torch.int32).view(out.shape[0], out.shape[1], 2)
    ids = pairs[..., 0].contiguous()
    dist = pairs[..., 1].contiguous().view(torch.float32)
    return ids, dist


def _search(index, queries, k, shortlist_factor, f64_boundary):
    out = index.search(queries.detach().contiguous(), k, "ip", shortlist_factor=shortlist_factor,
                       f64_boundary=f64_boundary)
    ids, dist = _split_device(out)
    return ids, -dist


def _backward(queries, vectors, ids, k, grad_scores=None, targets=None, inv_tau=0.0, scale=0.0,
              grad_out=None, want_dx=True):
    q = queries.detach().contiguous().float()
    x = vectors.detach()
    x32 = x if x.dtype == torch.float32 else x.float()
    x32 = x32.contiguous()
    nq, dim = q.shape
    dq = torch.empty_like(q)
    dx = torch.zeros_like(x32) if want_dx else None
    rc = L.load().fv_topk_ip_backward(
        L.ptr(q), nq, L.ptr(x32), x32.shape[0], dim, L.ptr(ids), k,
        L.ptr(grad_scores), L.ptr(targets), ctypes.c_float(inv_tau), ctypes.c_float(scale),
        L.ptr(grad_out), L.ptr(dq), L.ptr(dx), None, L.stream_handle())
    L.check(rc, "fv_topk_ip_backward")
    return dq.to(queries.dtype), (None if dx is None else dx.to(vectors.dtype))


def _index_for(vectors, index, dtype):
    if index is not None:
        return index
    x32 = vectors.detach()
    x32 = (x32 if x32.dtype == torch.float32 else x32.float()).contiguous()
    return F.FlatIndex(x32, dtype)


def _check_index(index, vectors):
    if index.n != vectors.shape[0] or index.dim != vectors.shape[1]:
        raise L.FvError("index does not cover the database tensor")


class TopKInnerProduct(torch.autograd.Function):
    """(scores [B, k], ids [B, k] int32) of the k rows with the largest
    q . x. `index` is a FlatIndex over the current rows (built from `vectors`
    when None)."""

    @staticmethod
    def forward(ctx, queries, vectors, k, index=None, shortlist_factor=8, f64_boundary=True):
        index = _index_for(vectors, index, "bf16")
        _check_index(index, vectors)
        ids, scores = _search(index, queries, k, shortlist_factor, f64_boundary)
        ctx.save_for_backward(queries, vectors, ids)
        ctx.k = k
        ctx.mark_non_differentiable(ids)
        return scores, ids

    @staticmethod
    def backward(ctx, grad_scores, grad_ids):
        queries, vectors, ids = ctx.saved_tensors
        dq, dx = _backward(queries, vectors, ids, ctx.k,
                           grad_scores=grad_scores.detach().float().contiguous(),
                           want_dx=ctx.needs_input_grad[1])
        return (dq if ctx.needs_input_grad[0] else None), dx, None, None, None, None


class TopKSoftmaxCE(torch.autograd.Function):
    """Mean over queries of CE(softmax(s_i / tau), targets[i]) over the top-k
    scores s_i; targets are positions in [0, k)."""

    @staticmethod
    def forward(ctx, queries, vectors, k, targets, tau, index=None, shortlist_factor=8,
                f64_boundary=True):
        index = _index_for(vectors, index, "bf16")
        _check_index(index, vectors)
        ids, scores = _search(index, queries, k, shortlist_factor, f64_boundary)
        z = scores / tau
        loss = (torch.logsumexp(z, dim=1) - z.gather(1, targets.long()[:, None])[:, 0]).mean()
        ctx.save_for_backward(queries, vectors, ids, targets.to(torch.int32).contiguous())
        ctx.k, ctx.tau = k, tau
        return loss

    @staticmethod
    def backward(ctx, grad_loss):
        queries, vectors, ids, targets = ctx.saved_tensors
        g = grad_loss.detach().float().reshape(1).contiguous()
        dq, dx = _backward(queries, vectors, ids, ctx.k, targets=targets, inv_tau=1.0 / ctx.tau,
                           scale=1.0 / queries.shape[0], grad_out=g,
                           want_dx=ctx.needs_input_grad[1])
        return (dq if ctx.needs_input_grad[0] else None), dx, None, None, None, None, None, None


def topk_inner_product(queries, vectors, k, index=None, shortlist_factor=8, f64_boundary=True):
    return TopKInnerProduct.apply(queries, vectors, k, index, shortlist_factor, f64_boundary)


def topk_softmax_ce(queries, vectors, k, targets, tau, index=None, shortlist_factor=8,
                    f64_boundary=True):
    return TopKSoftmaxCE.apply(queries, vectors, k, targets, tau, index, shortlist_factor,
                               f64_boundary)
