"""cbct_fp_noalloc (the forward projection that allocates nothing) against
cbct_fp through the C ABI: the same kernels, so the same bits, on two calls
in a row (the second with a smaller volume and detector), for FP32 and BF16
working copies and with accumulation; a volume whose working copy exceeds
the fixed scratch is refused with an error and leaves the output alone; and
the call makes no device allocation (device free memory unchanged)."""
import numpy as np
import pytest

import oracle as O
from gtt import capi
from helpers import angles_full, geom, phantom_volume, rel_l2, to_oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def print_case(*a):
    print("CASE", *a, flush=True)


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_fp_noalloc_matches_fp_two_calls(storage):
    ctx = capi.Context(0)
    ctx.storage(storage)
    try:
        for (n, nu, nv, nviews) in [(96, 160, 120, 40), (48, 70, 54, 17)]:
            g = geom(n, n, n, nu, nv)
            angles = angles_full(nviews, 360.0, start=5.0)
            print_case("fp_noalloc", storage, g, nviews)
            ctx.set_views(angles)
            vol = torch.from_numpy(phantom_volume(g).astype(np.float32)).cuda()
            want = torch.zeros(g.proj_shape(nviews), device="cuda")
            ctx.fp(g, vol, want)
            got = torch.full_like(want, 7.0)           # overwritten, not accumulated
            ctx.fp_noalloc(g, vol, got)
            torch.cuda.synchronize()
            assert torch.equal(got, want)
            acc = torch.full_like(want, 0.5)
            ctx.fp_noalloc(g, vol, acc, accumulate=True)
            ref = torch.full_like(want, 0.5)
            ctx.fp(g, vol, ref, accumulate=True)
            torch.cuda.synchronize()
            assert torch.equal(acc, ref)
            if storage == "f32" and n == 48:   # and to the oracle's bar
                assert rel_l2(got.cpu().numpy(), O.fp(to_oracle(g), angles, None,
                                                      vol.cpu().numpy().astype(np.float64))) <= 1e-5
    finally:
        ctx.close()


def test_fp_noalloc_refuses_oversized_and_allocates_nothing():
    ctx = capi.Context(0)
    try:
        g = geom(64, 64, 64, 80, 60)
        angles = angles_full(12, 360.0)
        ctx.set_views(angles)
        vol = torch.from_numpy(phantom_volume(g).astype(np.float32)).cuda()
        out = torch.zeros(g.proj_shape(12), device="cuda")
        ctx.fp_noalloc(g, vol, out)                    # first call: address lookup
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        for _ in range(3):
            ctx.fp_noalloc(g, vol, out)
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free0
        # 300^3 FP32: the working copy (303 x 303 x 304 x 4 B) exceeds 64 MiB
        big = geom(300, 300, 300, 16, 16)
        print_case("fp_noalloc oversized", big)
        vb = torch.zeros((300, 300, 300), device="cuda")
        ob = torch.full((12, 16, 16), 3.0, device="cuda")
        with pytest.raises(capi.CbctError, match="scratch"):
            ctx.fp_noalloc(big, vb, ob)
        torch.cuda.synchronize()
        assert bool((ob == 3.0).all())
    finally:
        ctx.close()
