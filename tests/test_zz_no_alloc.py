"""mm_rcvr_normalize_ssq_noalloc (static scratch, no allocation) against
mm_rcvr_normalize_ssq on two calls in a row, the second on a smaller
operator, both through the C ABI with device pointers."""
import numpy as np
import pytest

from conftest import load_pkg

pytestmark = pytest.mark.gpu


def test_noalloc_normalize_matches_allocating_entry():
    import torch
    mm = load_pkg()
    if mm.device_count() < 1:
        pytest.fail("no GPU visible to libmatmri_hip")
    L, check = mm._lib.lib(), mm._lib.check
    rng = np.random.default_rng(77)
    cases = [[(96, 96)] * 5 + [(64, 80)] * 3, [(40, 40), (32, 48)]]   # second call: smaller
    for dims in cases:
        print(f"[case] no-alloc coil normalisation, slices {dims}", flush=True)
        C = 6
        maps = [rng.standard_normal(d + (C,)) + 1j * rng.standard_normal(d + (C,)) for d in dims]
        R = mm.rcvrOp.batch(maps)
        n = int(sum(np.prod(d) for d in dims))
        x = torch.as_tensor((rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64),
                            device="cuda")
        want = torch.empty_like(x)
        got = torch.full_like(x, complex(np.nan, np.nan))
        torch.cuda.synchronize()
        check(L.mm_rcvr_normalize_ssq(R._plan.h, x.data_ptr(), want.data_ptr(), 0.05, 1))
        check(L.mm_rcvr_normalize_ssq_noalloc(R._plan.h, x.data_ptr(), got.data_ptr(), 0.05))
        check(L.mm_ctx_sync(mm._lib.context()))
        same = torch.equal(got, want)
        print(f"  {n} pixels: bitwise equal {same}", flush=True)
        assert same
    # an operator whose scratch exceeds the static maximum is rejected
    big = [np.ones((2048, 2048, 1), np.complex64)]
    Rb = mm.rcvrOp.batch(big)
    xb = torch.zeros(2048 * 2048, dtype=torch.complex64, device="cuda")
    rc = L.mm_rcvr_normalize_ssq_noalloc(Rb._plan.h, xb.data_ptr(), xb.data_ptr(), 0.0)
    assert rc == 1 and b"static maximum" in L.mm_last_error()
