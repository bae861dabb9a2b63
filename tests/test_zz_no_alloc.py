"""ugx_gemm_bf16x3_static, the allocation-free variant of ugx_gemm with
UGX_PREC_BF16X3: through the C ABI, two calls in a row (the second smaller,
after the first left its staging in the shared static array), each bitwise
equal to the allocating entry point's result; sizes beyond the static
array's 64 MiB are rejected with UGX_BAD_INPUT and leave the output alone."""
import numpy as np
import pytest

from ugacc_amd import ugx

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).cuda()


def test_static_gemm_matches_allocating_entry_point():
    from tests.gpu_util import announce
    c = ugx.Context(0)
    try:
        for M, N, K, opa, opb, tile in [(300, 257, 190, "N", "T", -1), (70, 33, 64, "T", "N", 3)]:
            announce("static bf16x3 gemm M=%d N=%d K=%d %s%s tile=%d" % (M, N, K, opa, opb, tile))
            rng = np.random.default_rng(M + N + K)
            A = rng.standard_normal((M, K))
            B = rng.standard_normal((K, N))
            C0 = rng.standard_normal((M, N))
            As = A if opa == "N" else A.T.copy()
            Bs = B if opb == "N" else B.T.copy()
            dA, dB = _dev(As, torch.float32), _dev(Bs, torch.float32)
            d1, d2 = _dev(C0, torch.float64), _dev(C0, torch.float64)
            torch.cuda.synchronize()
            c.gemm(opa, opb, M, N, K, 0.7, dA.data_ptr(), As.shape[1], dB.data_ptr(), Bs.shape[1], -0.3,
                   d1.data_ptr(), N, precision=ugx.PREC_BF16X3, tile=tile)
            c.gemm_bf16x3_static(opa, opb, M, N, K, 0.7, dA.data_ptr(), As.shape[1], dB.data_ptr(), Bs.shape[1], -0.3,
                                 d2.data_ptr(), N, tile=tile)
            c.synchronize()
            g1, g2 = d1.cpu().numpy(), d2.cpu().numpy()
            assert np.array_equal(g1, g2)
            ref = 0.7 * (A.astype(np.float32).astype(np.float64) @ B.astype(np.float32).astype(np.float64)) - 0.3 * C0
            scale = 0.7 * np.abs(A) @ np.abs(B) + 1.0
            assert np.max(np.abs(g2 - ref) / scale) < 2e-5
        # (M + N) K beyond 16 Mi elements: rejected before any work
        announce("static bf16x3 gemm: oversize rejected")
        M = N = 4096
        K = 2049
        dummy = torch.zeros(1, dtype=torch.float32, device="cuda")
        dC = torch.full((2, 2), 5.0, dtype=torch.float64, device="cuda")
        with pytest.raises(ugx.UgxError) as e:
            c.gemm_bf16x3_static("N", "T", M, N, K, 1.0, dummy.data_ptr(), K, dummy.data_ptr(), K, 0.0,
                                 dC.data_ptr(), N)
        assert e.value.code == ugx.BAD_INPUT and "64 MiB" in str(e.value)
        assert torch.all(dC == 5.0).item()
    finally:
        c.close()
