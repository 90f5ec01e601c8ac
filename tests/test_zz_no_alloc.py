"""mvmc_pfinv_f64_noalloc (no device allocation: the work buffer is the
library's per-device static array) against mvmc_pfinv_f64 through the C ABI:
Pfaffians, statuses and inverses bit-identical on consecutive calls, the
later ones smaller (the array keeps the earlier call's contents), with a
zero plane inside the batch and more plane groups than the array has work
slots at n = 224; sizes beyond the maximum are rejected."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import torch
    torch.zeros(1, device="cuda")
    torch.cuda.synchronize()
    import mvmc_hip
    return mvmc_hip


def _planes(n, count, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((count, n, n))
    a = a - a.transpose(0, 2, 1)
    a[count // 2] = 0.0
    return np.ascontiguousarray(a)


def _run(M, fn, planes, n, inplace=False):
    import torch
    count = planes.shape[0]
    dp = torch.tensor(planes, device="cuda")
    pf = torch.full((count,), 7.0, dtype=torch.float64, device="cuda")
    st = torch.full((count,), 99, dtype=torch.int32, device="cuda")
    inv = dp if inplace else torch.full_like(dp, 3.0)
    fn(dp, pf, inv, st, n, count)
    torch.cuda.synchronize()
    return pf.cpu().numpy(), inv.cpu().numpy(), st.cpu().numpy()


def test_noalloc_matches_allocating_entry_point(M):
    # (n, count): 250 plane groups of 4 at n = 224 against 83 work slots,
    # then smaller calls on the same array
    for k, (n, count) in enumerate([(224, 1000), (224, 200), (64, 300)]):
        log(f"no-alloc call {k}: n={n} count={count}")
        planes = _planes(n, count, 100 + k)
        ref = _run(M, M.pfaffian_inverse_batched, planes, n)
        got = _run(M, M.pfaffian_inverse_batched_noalloc, planes, n)
        assert ref[2][count // 2] > 0 and (ref[2] == 0).sum() == count - 1
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), (n, count)
    # in place (inv == planes)
    planes = _planes(64, 50, 7)
    ref = _run(M, M.pfaffian_inverse_batched, planes, 64)
    got = _run(M, M.pfaffian_inverse_batched_noalloc, planes, 64, inplace=True)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_noalloc_rejects_sizes_beyond_maximum(M):
    import torch
    n = 1026
    p = torch.zeros((1, n, n), dtype=torch.float64, device="cuda")
    pf = torch.zeros(1, dtype=torch.float64, device="cuda")
    st = torch.zeros(1, dtype=torch.int32, device="cuda")
    with pytest.raises(M.InvalidShape, match="no-allocation maximum"):
        M.pfaffian_inverse_batched_noalloc(p, pf, p, st, n, 1)
    with pytest.raises(M.InvalidShape):
        M.pfaffian_inverse_batched_noalloc(p, pf, p, st, 3, 1)
