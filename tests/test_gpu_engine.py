"""GPU checks of the MFMA GEMM engine and the permutation kernel against numpy (FP64)."""
import numpy as np
import pytest

from gpu_helpers import gpu_contract, gpu_permute, hooks, requires_gpu

pytestmark = [pytest.mark.gpu, requires_gpu]

# index extents used by the contraction patterns of the CC residual
SIZES = {"small": dict(a=19, b=19, c=19, d=19, i=5, j=5, k=5, l=5, Q=72, p=24, q=24),
         "ragged": dict(a=131, b=131, c=131, d=131, i=7, j=7, k=7, l=7, Q=97, p=138, q=138)}

PATTERNS = [
    ("klcd", "cdij", "klij"),
    ("klij", "abkl", "abij"),
    ("klcd", "adkj", "lcaj"),
    ("lcaj", "cbil", "abij"),
    ("klcd", "cakl", "ad"),
    ("klcd", "cdil", "ki"),
    ("ad", "dbij", "abij"),
    ("ki", "abkj", "abij"),
    ("aiQ", "kcQ", "akic"),
    ("akic", "cbkj", "abij"),
    ("bkci", "ackj", "abij"),
    ("acQ", "kiQ", "akci"),
    ("kcQ", "jiQ", "kjci"),
    ("kjci", "cakj", "ai"),
    ("abij", "jb", "ai"),
    ("ibQ", "bj", "ijQ"),
    ("pqQ", "Q", "pq"),
    ("ai", "bj", "abij"),
]


def _rand(labels, ext, rng):
    return np.asfortranarray(rng.standard_normal(tuple(ext[c] for c in labels)))


@pytest.mark.parametrize("size", list(SIZES))
@pytest.mark.parametrize("la,lb,lc", PATTERNS)
def test_contract_matches_einsum(size, la, lb, lc):
    ext = SIZES[size]
    rng = np.random.default_rng(abs(hash((size, la, lb, lc))) % 2**32)
    A, B, C = _rand(la, ext, rng), _rand(lb, ext, rng), _rand(lc, ext, rng)
    ref = 0.7 * np.einsum(f"{la},{lb}->{lc}", A, B) - 0.3 * C
    got = gpu_contract(0.7, A, la, B, lb, -0.3, C, lc)
    scale = np.max(np.abs(ref)) + 1e-300
    assert np.max(np.abs(got - ref)) / scale < 1e-13


@pytest.mark.parametrize("shape,perm", [((19, 19, 5, 5), (1, 0, 2, 3)), ((19, 19, 5, 5), (1, 0, 3, 2)),
                                        ((33, 40, 7, 3), (2, 0, 3, 1)), ((130, 70), (1, 0)),
                                        ((5, 19, 19, 5), (0, 2, 1, 3))])
def test_permute(shape, perm):
    rng = np.random.default_rng(7)
    X = rng.standard_normal(shape)
    Y = rng.standard_normal(tuple(shape[p] for p in perm))
    got = gpu_permute(X, perm, 2.0, 0.5, Y)
    ref = 2.0 * np.transpose(X, perm) + 0.5 * Y
    assert np.array_equal(got, ref)


def test_gemm_large_split_k():
    """Small output, long K: exercises the split-K partial sums."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((40, 40, 150, 150))
    B = rng.standard_normal((150, 150, 40, 40))
    ref = np.einsum("klcd,cdij->klij", A, B)
    got = gpu_contract(1.0, A, "klcd", B, "cdij", 0.0, np.zeros((40, 40, 40, 40)), "klij")
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-13


@pytest.mark.parametrize("la,lb,lc", [("akic", "cbkj", "abij"), ("klcd", "cdij", "klij"), ("ad", "dbij", "abij")])
def test_contract_fp32_products(la, lb, lc):
    """FP32 MFMA products with FP32 accumulation: error bounded by K * 2^-23 relative."""
    ext = SIZES["ragged"]
    rng = np.random.default_rng(17)
    A, B, C = _rand(la, ext, rng), _rand(lb, ext, rng), _rand(lc, ext, rng)
    ref = np.einsum(f"{la},{lb}->{lc}", A, B) + C
    got = gpu_contract(1.0, A, la, B, lb, 1.0, C, lc, prec=1)
    K = np.prod([ext[c] for c in la if c not in lc])
    # one rounding of each operand to FP32 plus FP32 accumulation, against the FP64 result
    bound = 4.0 * np.sqrt(K) * 2.0**-23 * np.einsum(f"{la},{lb}->{lc}", np.abs(A), np.abs(B)).max()
    assert float(np.max(np.abs(got - ref))) < bound


# BF16 cases: M and N large enough for the BF16 path (M N / (M + N) >= 256), K from one
# 32-wide slab (k = 19) over two (Q = 60) and four (Q = 97) to many with split K (917)
BF16_CASES = [
    (dict(a=131, b=131, i=7, j=7, k=19), "aik", "kbj", "abij"),
    (dict(a=131, i=7, k=7, c=131, Q=60), "aiQ", "kcQ", "akic"),
    (dict(a=131, i=7, k=7, c=131, Q=97), "aiQ", "kcQ", "akic"),
    (dict(a=131, b=131, c=131, i=7, j=7, k=7), "akic", "cbkj", "abij"),
]


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("case", range(len(BF16_CASES)))
def test_contract_bf16x3_products(kernel, case):
    """BF16 hi/lo products (three BF16 MFMA terms, FP32 accumulation), for each of the
    engine's BF16 kernels (register-staged slabs; slabs loaded straight into LDS).

    Each operand is represented to 2^-17 relative and the dropped lo*lo term is
    below 2^-16 relative, so |error| <= (2^-14 + 4 sqrt(K) 2^-23) * (|A||B|);
    a lost hi*lo term (error ~2^-9 relative) would fail this bound."""
    ext, la, lb, lc = BF16_CASES[case]
    rng = np.random.default_rng(19)
    A, B, C = _rand(la, ext, rng), _rand(lb, ext, rng), _rand(lc, ext, rng)
    ref = np.einsum(f"{la},{lb}->{lc}", A, B) + C
    prev = hooks().ecgt_set_bf16_kernel(kernel)
    try:
        got = gpu_contract(1.0, A, la, B, lb, 1.0, C, lc, prec=2)
    finally:
        hooks().ecgt_set_bf16_kernel(prev)
    K = np.prod([ext[c] for c in la if c not in lc])
    scale = np.einsum(f"{la},{lb}->{lc}", np.abs(A), np.abs(B)).max()
    bound = (2.0**-14 + 4.0 * np.sqrt(K) * 2.0**-23) * scale
    err = float(np.max(np.abs(got - ref)))
    assert err < bound
    assert err > 0.0          # the path is not silently FP64


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("case", range(len(BF16_CASES)))
def test_contract_fp32_kernels(kernel, case):
    """Large FP32 products through both FP32 kernels (generic FP32 MFMA; operands packed into
    FP32 planes for the direct-to-LDS kernel, K from one slab to many with split K): error
    bounded by 4 sqrt(K) 2^-23 (|A||B|), and the packed kernel's result is not FP64's."""
    ext, la, lb, lc = BF16_CASES[case]
    rng = np.random.default_rng(23)
    A, B, C = _rand(la, ext, rng), _rand(lb, ext, rng), _rand(lc, ext, rng)
    ref = np.einsum(f"{la},{lb}->{lc}", A, B) + C
    prev = hooks().ecgt_set_f32_kernel(kernel)
    try:
        got = gpu_contract(1.0, A, la, B, lb, 1.0, C, lc, prec=1)
    finally:
        hooks().ecgt_set_f32_kernel(prev)
    K = np.prod([ext[c] for c in la if c not in lc])
    bound = 4.0 * np.sqrt(K) * 2.0**-23 * np.einsum(f"{la},{lb}->{lc}", np.abs(A), np.abs(B)).max()
    err = float(np.max(np.abs(got - ref)))
    assert err < bound
    assert err > 0.0


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("case", range(len(BF16_CASES)))
def test_contract_fp64_kernels(kernel, case):
    """Large FP64 products through both FP64 kernels (generic FP64 MFMA; operands packed into
    FP64 planes for the direct-to-LDS kernel, K from one slab to many with split K) against
    numpy: FP64 rounding only (1e-12 relative to |A||B|)."""
    ext, la, lb, lc = BF16_CASES[case]
    rng = np.random.default_rng(29)
    A, B, C = _rand(la, ext, rng), _rand(lb, ext, rng), _rand(lc, ext, rng)
    ref = np.einsum(f"{la},{lb}->{lc}", A, B) + C
    prev = hooks().ecgt_set_f64_kernel(kernel)
    try:
        got = gpu_contract(1.0, A, la, B, lb, 1.0, C, lc, prec=0)
    finally:
        hooks().ecgt_set_f64_kernel(prev)
    scale = np.einsum(f"{la},{lb}->{lc}", np.abs(A), np.abs(B)).max()
    assert float(np.max(np.abs(got - ref))) < 1e-12 * scale

