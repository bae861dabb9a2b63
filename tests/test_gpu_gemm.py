"""MFMA GEMM kernel against a float64 reference of the same product, for the
three arithmetic forms, both operand layouts and every tile shape."""
import numpy as np
import pytest

from ugacc_amd import ugx

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).cuda()


@pytest.mark.parametrize("prec", [ugx.PREC_FP64, ugx.PREC_FP32, ugx.PREC_BF16X3])
@pytest.mark.parametrize("opa,opb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("M,N,K", [(1, 37, 50), (45, 70, 33), (130, 129, 67), (200, 40, 300)])
def test_gemm_forms(prec, opa, opb, M, N, K):
    from tests.gpu_util import context, announce
    announce("gemm prec=%d %s%s M=%d N=%d K=%d" % (prec, opa, opb, M, N, K))
    c = context()
    rng = np.random.default_rng(M * 1000 + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    C0 = rng.standard_normal((M, N))
    As = A if opa == "N" else A.T.copy()
    Bs = B if opb == "N" else B.T.copy()
    tdt = torch.float64 if prec == ugx.PREC_FP64 else torch.float32
    dA, dB, dC = _dev(As, tdt), _dev(Bs, tdt), _dev(C0, torch.float64)
    torch.cuda.synchronize()
    c.gemm(opa, opb, M, N, K, 0.7, dA.data_ptr(), As.shape[1], dB.data_ptr(), Bs.shape[1], -0.3, dC.data_ptr(), N,
           precision=prec)
    c.synchronize()
    ref = 0.7 * A @ B - 0.3 * C0
    got = dC.cpu().numpy()
    scale = np.abs(A) @ np.abs(B)
    tol = {ugx.PREC_FP64: 1e-13, ugx.PREC_FP32: 2e-6, ugx.PREC_BF16X3: 2e-5}[prec]
    err = np.max(np.abs(got - ref) / (scale + 1.0))
    assert err < tol, err


@pytest.mark.parametrize("prec", [ugx.PREC_FP64, ugx.PREC_FP32])
@pytest.mark.parametrize("layout", ["NT", "TN"])
@pytest.mark.parametrize("tile", range(12))
def test_gemm_tiles(tile, layout, prec):
    """Every tile shape (including the 80- and 160-row ones) with both the
    k-contiguous (NT) and the row-contiguous (TN) operand loaders."""
    from tests.gpu_util import context, announce
    announce("gemm tile=%d %s prec=%d" % (tile, layout, prec))
    c = context()
    rng = np.random.default_rng(tile)
    M, N, K = 300, 170, 90
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((N, K))
    tdt = torch.float64 if prec == ugx.PREC_FP64 else torch.float32
    dC = torch.zeros((M, N), dtype=torch.float64, device="cuda")
    if layout == "NT":
        dA, dB = _dev(A, tdt), _dev(B, tdt)
        torch.cuda.synchronize()
        c.gemm("N", "T", M, N, K, 1.0, dA.data_ptr(), K, dB.data_ptr(), K, 0.0, dC.data_ptr(), N, precision=prec,
               tile=tile)
    else:
        dA, dB = _dev(A.T.copy(), tdt), _dev(B.T.copy(), tdt)
        torch.cuda.synchronize()
        c.gemm("T", "N", M, N, K, 1.0, dA.data_ptr(), M, dB.data_ptr(), N, 0.0, dC.data_ptr(), N, precision=prec,
               tile=tile)
    c.synchronize()
    tol = 1e-12 if prec == ugx.PREC_FP64 else 1e-4
    np.testing.assert_allclose(dC.cpu().numpy(), A @ B.T, rtol=0, atol=tol)


@pytest.mark.parametrize("form", ["pair", "panel"])
@pytest.mark.parametrize("layout", ["NT", "TN"])
@pytest.mark.parametrize("tile", range(12))
@pytest.mark.parametrize("K", [152, 64])
def test_gemm_bf16x3_tiles(tile, layout, form, K):
    """Every tile of the two BF16x3 kernels against a float64 product:
    "pair" splits FP32 operands into BF16 hi/lo arrays (k_gemm_b3p, the
    CCSD contractions' register-staged kernel); "panel" packs FP64 operands
    into MFMA panels and runs the LDS-copy kernel of the (T) W3 build
    (k_gemm_b3d; tiles whose block copies do not split over the four waves
    take k_gemm_b3p on the same panels).  K = 152 ends in a partial 32-deep
    slice (zero padding), as the (T) W3's K = v does at (H2O)8; M and N are
    not multiples of the tiles.  BF16x3 drops only the lo*lo product
    (~2^-16 relative per product), well inside 2e-5 of sum |a||b|."""
    from tests.gpu_util import context, announce
    announce("gemm bf16x3 %s tile=%d %s K=%d" % (form, tile, layout, K))
    c = context()
    rng = np.random.default_rng(1000 + tile * 7 + K)
    M, N = 203, 170
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((N, K))
    C0 = rng.standard_normal((M, N))
    tdt = torch.float32 if form == "pair" else torch.float64
    prec = ugx.PREC_BF16X3 if form == "pair" else ugx.PREC_BF16X3_PANEL
    if form == "pair":
        A, B = A.astype(np.float32).astype(np.float64), B.astype(np.float32).astype(np.float64)
    dC = _dev(C0, torch.float64)
    if layout == "NT":
        dA, dB = _dev(A, tdt), _dev(B, tdt)
        torch.cuda.synchronize()
        c.gemm("N", "T", M, N, K, 0.5, dA.data_ptr(), K, dB.data_ptr(), K, 2.0, dC.data_ptr(), N, precision=prec,
               tile=tile)
    else:
        dA, dB = _dev(A.T.copy(), tdt), _dev(B.T.copy(), tdt)
        torch.cuda.synchronize()
        c.gemm("T", "N", M, N, K, 0.5, dA.data_ptr(), M, dB.data_ptr(), N, 2.0, dC.data_ptr(), N, precision=prec,
               tile=tile)
    c.synchronize()
    ref = 0.5 * A @ B.T + 2.0 * C0
    scale = 0.5 * np.abs(A) @ np.abs(B).T + 1.0
    err = np.max(np.abs(dC.cpu().numpy() - ref) / scale)
    assert err < 2e-5, err
