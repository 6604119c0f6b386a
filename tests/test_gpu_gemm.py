"""The BF16-MFMA GEMM (gpesolver_amd/csrc/gpe_gemm.hpp) against numpy on
BF16-rounded operands (so only FP32 accumulation order differs), for real and
complex operands, conjugation, odd sizes and K not a multiple of the tile."""
import ctypes as C

import numpy as np
import pytest

from gpesolver_amd import hip

pytestmark = pytest.mark.gpu


def bf16(a):
    """Round-to-nearest-even to bfloat16, returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def run(Ar, Ai, Br, Bi, cplx_out, conjA=0, conjB=0):
    L = hip.load()
    M, K = Ar.shape
    N = Br.shape[1]
    fp = C.POINTER(C.c_float)
    L.gpe_gemm_test.argtypes = [fp, fp, fp, fp, fp, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.gpe_gemm_test.restype = C.c_int

    def p(a):
        return None if a is None else np.ascontiguousarray(a, dtype=np.float32).ctypes.data_as(fp)

    keep = [np.ascontiguousarray(x, dtype=np.float32) if x is not None else None for x in (Ar, Ai, Br, Bi)]
    Cr = np.zeros((M, N), dtype=np.float32)
    Ci = np.zeros((M, N), dtype=np.float32) if cplx_out else None
    rc = L.gpe_gemm_test(*(p(x) for x in keep), p(Cr), p(Ci), M, N, K, conjA, conjB)
    assert rc == 0
    return Cr, Ci


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (16, 32, 16), (100, 70, 45), (9, 288, 16), (1, 5, 130)])
def test_real_gemm(M, N, K):
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    Cr, _ = run(A, None, B, None, False)
    ref = bf16(A).astype(np.float64) @ bf16(B).astype(np.float64)
    assert np.abs(Cr - ref).max() < 1e-4 * np.abs(ref).max() + 1e-6
    # and against the unrounded product: BF16 operand rounding (~2^-9 relative)
    exact = A.astype(np.float64) @ B.astype(np.float64)
    assert np.linalg.norm(Cr - exact) < 1e-2 * np.linalg.norm(exact)


@pytest.mark.parametrize("conjA,conjB", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(64, 64, 64), (33, 47, 20)])
def test_complex_gemm(M, N, K, conjA, conjB):
    rng = np.random.default_rng(7 * M + N)
    A = (rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K))).astype(np.complex64)
    B = (rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))).astype(np.complex64)
    Cr, Ci = run(A.real, A.imag, B.real, B.imag, True, conjA, conjB)
    Ab = bf16(A.real) + 1j * bf16(A.imag)
    Bb = bf16(B.real) + 1j * bf16(B.imag)
    ref = (np.conj(Ab) if conjA else Ab).astype(np.complex128) @ (np.conj(Bb) if conjB else Bb).astype(np.complex128)
    got = Cr + 1j * Ci
    assert np.abs(got - ref).max() < 1e-4 * np.abs(ref).max()


def test_mixed_real_complex_and_real_part_only():
    rng = np.random.default_rng(3)
    A = (rng.standard_normal((40, 24)) + 1j * rng.standard_normal((40, 24))).astype(np.complex64)
    B = rng.standard_normal((24, 36)).astype(np.float32)
    Cr, Ci = run(A.real, A.imag, B, None, True)
    ref = (bf16(A.real) + 1j * bf16(A.imag)).astype(np.complex128) @ bf16(B).astype(np.float64)
    assert np.abs(Cr + 1j * Ci - ref).max() < 1e-4 * np.abs(ref).max()
    Bc = (rng.standard_normal((24, 36)) + 1j * rng.standard_normal((24, 36))).astype(np.complex64)
    Cr, _ = run(A.real, A.imag, Bc.real, Bc.imag, False)
    ref = ((bf16(A.real) + 1j * bf16(A.imag)).astype(np.complex128)
           @ (bf16(Bc.real) + 1j * bf16(Bc.imag)).astype(np.complex128)).real
    assert np.abs(Cr - ref).max() < 1e-4 * np.abs(ref).max()


@pytest.mark.parametrize("transB", [0, 1])
@pytest.mark.parametrize("M,N,K,beta,with_bias", [(1000, 32, 32, 1.0, True), (777, 64, 32, 0.0, True),
                                                  (1024, 1, 64, 0.0, True), (513, 64, 1, 0.0, False),
                                                  (300, 32, 6, 0.0, True), (4096, 32, 64, 0.5, False)])
def test_row_gemm_equals_tile_gemm(M, N, K, beta, with_bias, transB):
    """The row GEMM of the network's pointwise layers (many rows, N, K <= 64)
    gives the tile kernel's results bit for bit: same BF16 operand rounding,
    k order and epilogue; and it is the BF16 product to FP32 accumulation."""
    L = hip.load()
    fp = C.POINTER(C.c_float)
    L.gpe_gemm_rows_test.argtypes = [fp, fp, fp, fp, C.c_long, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int]
    L.gpe_gemm_rows_test.restype = C.c_int
    rng = np.random.default_rng(M + 3 * N + 7 * K + transB)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if transB else (K, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32) if with_bias else None
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    out = []
    for rows in (0, 1):
        Cm = C0.copy()
        rc = L.gpe_gemm_rows_test(A.ctypes.data_as(fp), B.ctypes.data_as(fp),
                                  bias.ctypes.data_as(fp) if bias is not None else None, Cm.ctypes.data_as(fp), M, N,
                                  K, transB, beta, rows)
        assert rc == 0
        out.append(Cm)
    np.testing.assert_array_equal(out[1], out[0])
    Bk = B.T if transB else B
    ref = bf16(A).astype(np.float64) @ bf16(Bk).astype(np.float64) + beta * C0.astype(np.float64)
    if bias is not None:
        ref += bias.astype(np.float64)
    assert np.abs(out[1] - ref).max() < 1e-4 * np.abs(ref).max() + 1e-5


@pytest.mark.parametrize("K,N", [(32, 32), (6, 32), (32, 64), (64, 1), (64, 32), (1, 64)])
def test_wgrad_chunk_kernel_equals_tile_gemm(K, N):
    """The weight-gradient chunk kernel (coalesced 32-row slabs transposed in
    LDS) gives the tile kernel's per-chunk partials A_z^T D_z bit for bit, and
    they are the BF16 products with FP32 accumulation."""
    L = hip.load()
    fp = C.POINTER(C.c_float)
    L.gpe_gemm_wgrad_test.argtypes = [fp, fp, fp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.gpe_gemm_wgrad_test.restype = C.c_int
    chunks = 3
    rng = np.random.default_rng(K * 100 + N)
    A = rng.standard_normal((chunks * 512, K)).astype(np.float32)
    D = rng.standard_normal((chunks * 512, N)).astype(np.float32)
    out = []
    for wg in (0, 1):
        P = np.full((chunks, K, N), np.nan, dtype=np.float32)
        assert L.gpe_gemm_wgrad_test(A.ctypes.data_as(fp), D.ctypes.data_as(fp), P.ctypes.data_as(fp), chunks, K, N,
                                     wg) == 0
        out.append(P)
    np.testing.assert_array_equal(out[1], out[0])
    Ab, Db = bf16(A).astype(np.float64), bf16(D).astype(np.float64)
    ref = np.stack([Ab[512 * z:512 * (z + 1)].T @ Db[512 * z:512 * (z + 1)] for z in range(chunks)])
    assert np.abs(out[1] - ref).max() < 1e-4 * np.abs(ref).max()
