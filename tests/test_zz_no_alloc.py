"""fv_flat_scan_batch_noalloc (scratch in the library's static per-GPU array)
against cuda_flat_scan_batch through the C ABI: identical results, bit for
bit, on two calls in a row (the second smaller, so stale contents of the
shared array from the first call must not leak), for L2, inner product and
cosine, with a mask and remapped ids; and the size limit is an error that
launches nothing."""
import ctypes

import numpy as np
import pytest

from flashvector_amd import lib as L
from flashvector_amd import synthetic as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a ROCm device"
    L.load()
    return t


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def call(torch, fn, db, ids, qs, k, metric, mask):
    lib = L.load()
    width = max(1, min(k, 256))
    out = L.result_tensor(qs.shape[0], width, qs.device)
    out.fill_(0x5A5A5A5A)
    rc = getattr(lib, fn)(L.ptr(db), L.ptr(ids), db.shape[0], db.shape[1], L.ptr(qs),
                          qs.shape[0], k, L.METRICS[metric], L.ptr(mask), L.ptr(out),
                          L.stream_handle())
    assert rc == 0, (fn, rc)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("metric", ["l2", "ip", "cosine"])
def test_noalloc_equals_cuda_flat_scan_batch(torch, metric):
    db, qs = S.flat_workload(60000, 900, 128, 64, seeds=(201, 202, 203))
    mask = S.random_valid_mask(60000, 0.7, 204)
    ids = np.random.default_rng(205).permutation(60000).astype(np.uint32)
    d_db, d_q, d_m = dev(torch, db), dev(torch, qs), dev(torch, mask)
    d_ids = dev(torch, ids.view(np.int32))
    # First call: 900 queries, k = 100; second call: fewer rows and queries, k = 7.
    for n, nq, k in [(60000, 900, 100), (20000, 150, 7)]:
        args = (d_db[:n].contiguous(), d_ids[:n].contiguous(), d_q[:nq].contiguous(), k, metric,
                d_m[: (n + 31) // 32].contiguous())
        want = call(torch, "cuda_flat_scan_batch", *args)
        got = call(torch, "fv_flat_scan_batch_noalloc", *args)
        np.testing.assert_array_equal(got, want)


def test_noalloc_size_limit(torch):
    """A batch whose scratch exceeds the static array is refused with
    hipErrorOutOfMemory and the output is left untouched."""
    lib = L.load()
    db = torch.zeros((1000, 8), dtype=torch.float32, device="cuda")
    nq = 40000                      # 40000 * (512 + 256) * 8 bytes > 64 MiB
    qs = torch.zeros((nq, 8), dtype=torch.float32, device="cuda")
    out = L.result_tensor(nq, 256, qs.device)
    out.fill_(7)
    rc = lib.fv_flat_scan_batch_noalloc(L.ptr(db), None, 1000, 8, L.ptr(qs), nq, 256, 0, None,
                                        L.ptr(out), L.stream_handle())
    torch.cuda.synchronize()
    assert rc == 2                  # hipErrorOutOfMemory
    assert bool((out == 7).all())
    assert ctypes.c_int(rc).value != 0
