// gpe_gemm.hpp -- strided batched (complex) GEMM on gfx950 BF16 MFMA.
//
//   C[z] = alpha * op(A[z]) * op(B[z]) (+ bias[n]) + beta * C[z]
//
// A (m,k) lives at A[z*saz + m*sam + k*sak], B (k,n) at B[z*sbk...] (any
// strides, so transposes are free), C (m,n) at C[z*scz + m*scm + n*scn].
// Operands are FP32 in memory and rounded to BF16 (round to nearest even) as
// they are staged into LDS; products accumulate in FP32 in the MFMA unit.
// Complex operands are (re, im) pointer pairs with the same strides; an
// imaginary pointer of nullptr means a real operand, and Ci == nullptr writes
// only Re(C).  conjA / conjB conjugate the operand.
//
// Tile: 64x64 per block of 256 threads (4 waves, 32x32 each = 2x2 MFMA
// 16x16x32 tiles), K staged 32 at a time.  k_gemm_small (below) takes the
// batches of tiny products with a shared A in FP32.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gpe {
namespace gemm {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

struct Args {
    int M, N, K, batch;
    const float* Ar;
    const float* Ai;
    long sam, sak, saz;
    const float* Br;
    const float* Bi;
    long sbk, sbn, sbz;
    float* Cr;
    float* Ci;
    long scm, scn, scz;
    const float* bias;  // added to Re(C) per column n (may be null)
    float alpha, beta;
    int conjA, conjB;
};

constexpr int TM = 64, TN = 64, TK = 32, NT = 256;

__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));  // inf/nan
    const unsigned int rnd = 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)((u + rnd) >> 16);
}

// LDS tiles hold bf16 bit patterns: A as [m][k], B transposed as [n][k], so the
// 8 consecutive k values an MFMA lane needs are one 16-byte read.
struct Smem {
    unsigned short ar[TM][TK + 8];
    unsigned short ai[TM][TK + 8];
    unsigned short br[TN][TK + 8];
    unsigned short bi[TN][TK + 8];
};

__device__ __forceinline__ bf16x8 ld8(const unsigned short* p) {
    bf16x8 v;
    __builtin_memcpy(&v, p, 16);
    return v;
}

__device__ __forceinline__ bf16x8 neg8(bf16x8 v) {
    unsigned short t[8];
    __builtin_memcpy(t, &v, 16);
#pragma unroll
    for (int i = 0; i < 8; ++i) t[i] ^= 0x8000;
    __builtin_memcpy(&v, t, 16);
    return v;
}

__global__ void __launch_bounds__(NT) k_gemm(Args a) {
    __shared__ Smem sm;
    const int z = blockIdx.z;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const bool cA = a.Ai != nullptr, cB = a.Bi != nullptr;
    const bool cplx_out = a.Ci != nullptr;
    const float sgnA = a.conjA ? -1.f : 1.f, sgnB = a.conjB ? -1.f : 1.f;
    const float* Ar = a.Ar + (long)z * a.saz;
    const float* Ai = cA ? a.Ai + (long)z * a.saz : nullptr;
    const float* Br = a.Br + (long)z * a.sbz;
    const float* Bi = cB ? a.Bi + (long)z * a.sbz : nullptr;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 accr[2][2], acci[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            accr[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
            acci[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        }

    for (int k0 = 0; k0 < a.K; k0 += TK) {
        // stage A (64 x 32) and B (32 x 64): 8 elements per thread each
#pragma unroll
        for (int q = 0; q < (TM * TK) / NT; ++q) {
            const int e = threadIdx.x + q * NT;
            const int mm = e / TK, kk = e % TK;
            const int m = m0 + mm, k = k0 + kk;
            float vr = 0.f, vi = 0.f;
            if (m < a.M && k < a.K) {
                const long off = (long)m * a.sam + (long)k * a.sak;
                vr = Ar[off];
                if (cA) vi = sgnA * Ai[off];
            }
            sm.ar[mm][kk] = f2bf(vr);
            sm.ai[mm][kk] = f2bf(vi);
        }
#pragma unroll
        for (int q = 0; q < (TN * TK) / NT; ++q) {
            const int e = threadIdx.x + q * NT;
            const int nn = e / TK, kk = e % TK;
            const int n = n0 + nn, k = k0 + kk;
            float vr = 0.f, vi = 0.f;
            if (n < a.N && k < a.K) {
                const long off = (long)k * a.sbk + (long)n * a.sbn;
                vr = Br[off];
                if (cB) vi = sgnB * Bi[off];
            }
            sm.br[nn][kk] = f2bf(vr);
            sm.bi[nn][kk] = f2bf(vi);
        }
        __syncthreads();
        // MFMA 16x16x32: lane holds A[m = l%16][k = 8*(l/16) + 0..7] and
        // B[k = 8*(l/16) + 0..7][n = l%16].
        const int r16 = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const bf16x8 fa_r = ld8(&sm.ar[wm + 16 * i + r16][kq]);
            bf16x8 fa_i, fa_in;
            if (cA) {
                fa_i = ld8(&sm.ai[wm + 16 * i + r16][kq]);
                fa_in = neg8(fa_i);
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const bf16x8 fb_r = ld8(&sm.br[wn + 16 * j + r16][kq]);
                accr[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa_r, fb_r, accr[i][j], 0, 0, 0);
                if (cB) {
                    const bf16x8 fb_i = ld8(&sm.bi[wn + 16 * j + r16][kq]);
                    if (cA) accr[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa_in, fb_i, accr[i][j], 0, 0, 0);
                    if (cplx_out)
                        acci[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa_r, fb_i, acci[i][j], 0, 0, 0);
                }
                if (cA && cplx_out)
                    acci[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa_i, fb_r, acci[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // D[m = 4*(l/16) + r][n = l%16]
    float* Cr = a.Cr + (long)z * a.scz;
    float* Ci = cplx_out ? a.Ci + (long)z * a.scz : nullptr;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                const int n = n0 + wn + 16 * j + (lane & 15);
                if (m < a.M && n < a.N) {
                    const long off = (long)m * a.scm + (long)n * a.scn;
                    float vr = a.alpha * accr[i][j][r];
                    if (a.bias) vr += a.bias[n];
                    if (a.beta != 0.f) vr += a.beta * Cr[off];
                    Cr[off] = vr;
                    if (cplx_out) {
                        float vi = a.alpha * acci[i][j][r];
                        if (a.beta != 0.f) vi += a.beta * Ci[off];
                        Ci[off] = vi;
                    }
                }
            }
}

// Many small products with one shared A (saz == 0): M <= 16, K <= 32 and N
// dividing 256 (the z DFTs of the spectral layers, 9 x 16 by 16 x 32 per
// grid line).  They are memory-bound, so FP32 FMA on one thread per output
// column n of one product beats 64 x 64 MFMA tiles that such shapes fill to a
// few percent: a block runs 256 / N products, B and C are read and written
// along n (contiguous when scn = sbn = 1), and A sits in LDS.
constexpr int SM_M = 16, SM_K = 32;

inline bool small_ok(const Args& a) {
    return a.saz == 0 && a.M <= SM_M && a.K <= SM_K && a.N > 0 && 256 % a.N == 0 && a.bias == nullptr;
}

__global__ void __launch_bounds__(256) k_gemm_small(Args a) {
    __shared__ float far[SM_M * SM_K], fai[SM_M * SM_K];
    const bool cA = a.Ai != nullptr, cB = a.Bi != nullptr, cplx_out = a.Ci != nullptr;
    for (int e = threadIdx.x; e < a.M * a.K; e += 256) {
        const int m = e / a.K, k = e - m * a.K;
        const long off = (long)m * a.sam + (long)k * a.sak;
        far[e] = a.Ar[off];
        fai[e] = cA ? (a.conjA ? -a.Ai[off] : a.Ai[off]) : 0.f;
    }
    __syncthreads();
    const int per = 256 / a.N;
    const int n = threadIdx.x % a.N;
    const long z = (long)blockIdx.x * per + threadIdx.x / a.N;
    if (z >= a.batch) return;
    const float sgnB = a.conjB ? -1.f : 1.f;
    const float* Br = a.Br + z * a.sbz + (long)n * a.sbn;
    const float* Bi = cB ? a.Bi + z * a.sbz + (long)n * a.sbn : nullptr;
    float br[SM_K], bi[SM_K];
#pragma unroll
    for (int k = 0; k < SM_K; ++k) {
        br[k] = 0.f;
        bi[k] = 0.f;
        if (k < a.K) {
            br[k] = Br[(long)k * a.sbk];
            if (cB) bi[k] = sgnB * Bi[(long)k * a.sbk];
        }
    }
    float* Cr = a.Cr + z * a.scz + (long)n * a.scn;
    float* Ci = cplx_out ? a.Ci + z * a.scz + (long)n * a.scn : nullptr;
    for (int m = 0; m < a.M; ++m) {
        float sr = 0.f, si = 0.f;
#pragma unroll
        for (int k = 0; k < SM_K; ++k) {
            if (k < a.K) {
                const float ar = far[m * a.K + k], ai = fai[m * a.K + k];
                sr += ar * br[k] - ai * bi[k];
                si += ar * bi[k] + ai * br[k];
            }
        }
        const long off = (long)m * a.scm;
        float vr = a.alpha * sr;
        if (a.beta != 0.f) vr += a.beta * Cr[off];
        Cr[off] = vr;
        if (cplx_out) {
            float vi = a.alpha * si;
            if (a.beta != 0.f) vi += a.beta * Ci[off];
            Ci[off] = vi;
        }
    }
}

// Row GEMM: C[m][n] = alpha * sum_k A[m][k] B(k, n) (+ bias[n]) + beta * C[m][n]
// for many rows m (A row-major [M][K]) and N, K <= 64: the pointwise layers,
// lift and projection of the network, M = batch x grid points.  They are
// memory-bound (a row of A in, a row of C out), so each wave keeps the whole
// B as MFMA fragments in registers (B(k, n) = B[k N + n], or B[n K + k] with
// TRANSB), and streams groups of 16 rows: lane l loads row 16 g + l % 16,
// k = 32 s + 8 (l / 16) .. + 7 with two 16-byte loads when K % 8 == 0, and
// the group's 16 x N outputs take one MFMA 16x16x32 per 16 columns and 32 k.
// Operand rounding (f2bf), the k order and the epilogue are those of k_gemm,
// so the results equal k_gemm's bit for bit.
constexpr int kRowGroups = 8;  // row groups of 16 per wave

// A fragment of row m, k = k0 .. k0 + 7 (zeros past M or K)
__device__ __forceinline__ void row_frag_load(const float* __restrict__ A, long M, int K, long m, int k0, bool vec,
                                              float (&av)[8]) {
    if (vec) {
        float4 u = make_float4(0.f, 0.f, 0.f, 0.f), w = u;
        if (m < M && k0 < K) {
            const float4* p = reinterpret_cast<const float4*>(A + m * K + k0);
            u = p[0];
            w = p[1];
        }
        av[0] = u.x; av[1] = u.y; av[2] = u.z; av[3] = u.w;
        av[4] = w.x; av[5] = w.y; av[6] = w.z; av[7] = w.w;
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) av[j] = (m < M && k0 + j < K) ? A[m * K + k0 + j] : 0.f;
    }
}

template <bool TRANSB>
__global__ void __launch_bounds__(256) k_gemm_rows(const float* __restrict__ A, const float* __restrict__ B,
                                                   float* __restrict__ C, long M, int N, int K,
                                                   const float* __restrict__ bias, float alpha, float beta) {
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    const int ntn = (N + 15) / 16, nks = (K + 31) / 32;
    const long ngroups = (M + 15) / 16;
    const long g0 = wave * kRowGroups, g1 = g0 + kRowGroups < ngroups ? g0 + kRowGroups : ngroups;
    if (g0 >= g1) return;
    const bool vec = (K % 8) == 0;
    // the first group's rows are in flight while B becomes fragments
    float av[2][8];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
        if (ks < nks) row_frag_load(A, M, K, g0 * 16 + r16, ks * 32 + kq, vec, av[ks]);
    bf16x8 fb[4][2];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            if (nt >= ntn || ks >= nks) continue;
            unsigned short t[8];
            const int n = nt * 16 + r16;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = ks * 32 + kq + j;
                float v = 0.f;
                if (n < N && k < K) v = TRANSB ? B[(long)n * K + k] : B[(long)k * N + n];
                t[j] = f2bf(v);
            }
            __builtin_memcpy(&fb[nt][ks], t, 16);
        }
    for (long g = g0; g < g1; ++g) {
        const long m0 = g * 16;
        // this group's C (beta != 0) loads before its MFMAs, so the epilogue
        // does not wait a round trip per group
        float cold[4][4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long mm = m0 + 4 * (lane >> 4) + r;
                const int n = nt * 16 + r16;
                cold[nt][r] = (beta != 0.f && nt < ntn && mm < M && n < N) ? C[mm * N + n] : 0.f;
            }
        bf16x8 fa[2];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            unsigned short t[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) t[j] = f2bf(av[ks][j]);
            __builtin_memcpy(&fa[ks], t, 16);
        }
        // the next group's rows load while this group computes and stores
        if (g + 1 < g1) {
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
                if (ks < nks) row_frag_load(A, M, K, m0 + 16 + r16, ks * 32 + kq, vec, av[ks]);
        }
        f32x4 acc[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            if (ks >= nks) break;
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                if (nt < ntn) acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[ks], fb[nt][ks], acc[nt], 0, 0, 0);
        }
        // D[m = 4 (l / 16) + r][n = l % 16]
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            if (nt >= ntn) break;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long mm = m0 + 4 * (lane >> 4) + r;
                const int n = nt * 16 + r16;
                if (mm < M && n < N) {
                    const long off = mm * N + n;
                    float vr = alpha * acc[nt][r];
                    if (bias) vr += bias[n];
                    if (beta != 0.f) vr += beta * cold[nt][r];
                    C[off] = vr;
                }
            }
        }
    }
}

// Weight-gradient partials: for chunk z of `rows` rows, P_z[m][n] =
// sum_r A[r][m] D[r][n] over the chunk's rows (A [R][K], D [R][N] row-major,
// K, N <= 64), one block of 256 threads per chunk.  Each 32-row slab of A and
// D is a contiguous block of memory: it is read with 16-byte loads (the next
// slab's while this one computes), transposed into LDS as BF16 ([col][row],
// so an MFMA fragment is one 16-byte LDS read), and the 16 x 16 output tiles
// (m, n) are spread over the four waves.  Operand rounding, the k (row) order
// and the MFMA are those of k_gemm on the same products, so the partials
// equal k_gemm's bit for bit.
constexpr int kWgSlab = 32;
struct WgSmem {
    unsigned short a[64][kWgSlab + 8];
    unsigned short d[64][kWgSlab + 8];
};

__global__ void __launch_bounds__(256) k_gemm_wgrad(const float* __restrict__ A, const float* __restrict__ D,
                                                    float* __restrict__ P, int rows, int K, int N) {
    __shared__ WgSmem sm;
    const long z = blockIdx.x;
    const float* Az = A + z * (long)rows * K;
    const float* Dz = D + z * (long)rows * N;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    const int mtn = (K + 15) / 16, ntn = (N + 15) / 16, ntiles = mtn * ntn;
    const int na4 = kWgSlab * K / 4, nd4 = kWgSlab * N / 4;  // float4 per slab (32 K and 32 N are multiples of 4)
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    // slab loads: 32 K + 32 N floats <= 4096, at most 4 float4 per thread
    float4 pa[2], pd[2];
    auto load_slab = [&](int s0) {
        const float4* a4 = reinterpret_cast<const float4*>(Az + (long)s0 * K);
        const float4* d4 = reinterpret_cast<const float4*>(Dz + (long)s0 * N);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * 256;
            pa[q] = e < na4 ? a4[e] : make_float4(0.f, 0.f, 0.f, 0.f);
            pd[q] = e < nd4 ? d4[e] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    };
    load_slab(0);
    // zero padding past K / N (the slabs below write only columns < K, < N)
    for (int e = threadIdx.x; e < 64 * kWgSlab; e += 256) {
        const int c = e / kWgSlab, r = e % kWgSlab;
        if (c >= K) sm.a[c][r] = 0;
        if (c >= N) sm.d[c][r] = 0;
    }
    for (int s0 = 0; s0 < rows; s0 += kWgSlab) {
        // the slab transposed into LDS
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * 256;
            if (e < na4) {
                const float v[4] = {pa[q].x, pa[q].y, pa[q].z, pa[q].w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int f = 4 * e + j, r = f / K, c = f - r * K;
                    sm.a[c][r] = f2bf(v[j]);
                }
            }
            if (e < nd4) {
                const float v[4] = {pd[q].x, pd[q].y, pd[q].z, pd[q].w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int f = 4 * e + j, r = f / N, c = f - r * N;
                    sm.d[c][r] = f2bf(v[j]);
                }
            }
        }
        __syncthreads();
        if (s0 + kWgSlab < rows) load_slab(s0 + kWgSlab);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int tile = wave + 4 * t;
            if (tile < ntiles) {
                const int mt = tile / ntn, nt = tile - mt * ntn;
                const bf16x8 fa = ld8(&sm.a[mt * 16 + r16][kq]);
                const bf16x8 fb = ld8(&sm.d[nt * 16 + r16][kq]);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fb, acc[t], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // D[m = 4 (l / 16) + r][n = l % 16] of each tile
    float* Pz = P + z * (long)K * N;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int tile = wave + 4 * t;
        if (tile >= ntiles) break;
        const int mt = tile / ntn, nt = tile - mt * ntn;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = mt * 16 + 4 * (lane >> 4) + r, n = nt * 16 + r16;
            if (m < K && n < N) Pz[(long)m * N + n] = acc[t][r];
        }
    }
}

}  // namespace gemm
}  // namespace gpe
