// gpe_fno.hip -- Fourier-neural-operator warm start for imaginary time
// (no reference counterpart): training and inference on gfx950, every layer a
// BF16-MFMA GEMM (gpe_gemm.hpp) plus FP32 elementwise kernels.
//
// Network (grid nx x ny x nz, activations channels-last [B][x][y][z][c]):
//   h0 = x_in W_lift + b_lift                                  (cin -> W)
//   layer l = 0..L-1:  s = irdft_z idft_y idft_x ( R_l . dft_x dft_y rdft_z h_l )
//                      z_l = h_l W_l + b_l + s ;  h_{l+1} = GELU(z_l)  (no GELU after the last)
//   q = GELU(h_L W_q1 + b_q1) ;  y = q W_q2 + b_q2             (W -> hidden -> 1)
// The spectral transforms are truncated DFTs written as complex GEMMs with
// precomputed matrices: kx, ky in {0..m/2-1} U {n-m/2..n-1}, kz in
// {0..min(mz, nz/2+1)-1}; R_l holds one complex W x W matrix per retained mode.
// Loss: mean squared error against the target field.  Adam on FP32 master
// weights; GEMM operands are rounded to BF16, accumulation is FP32.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gpehip.h"
#include "gpe_gemm.hpp"

namespace {

using gpe::gemm::Args;

__device__ __forceinline__ float gelu(float x) {
    const float c = 0.7978845608028654f;  // sqrt(2/pi)
    return 0.5f * x * (1.f + tanhf(c * (x + 0.044715f * x * x * x)));
}
__device__ __forceinline__ float gelu_grad(float x) {
    const float c = 0.7978845608028654f;
    const float u = c * (x + 0.044715f * x * x * x);
    const float t = tanhf(u);
    return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * c * (1.f + 3.f * 0.044715f * x * x);
}

__global__ void k_gelu(const float* __restrict__ z, float* __restrict__ h, size_t n) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        h[e] = gelu(z[e]);
}
// dz = dh * gelu'(z)
__global__ void k_gelu_bwd(const float* __restrict__ dh, const float* __restrict__ z, float* __restrict__ dz,
                           size_t n) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        dz[e] = dh[e] * gelu_grad(z[e]);
}
// y [n], t [n]: per-block partial sums of (y - t)^2, dy = 2 (y - t) / n
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t, float* __restrict__ dy, size_t n,
                      double* __restrict__ part) {
    __shared__ double red[256];
    double acc = 0.0;
    const float sc = 2.f / (float)n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const float d = y[e] - t[e];
        acc += (double)d * d;
        if (dy) dy[e] = sc * d;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
// column sums of a [rows x cols] row-major matrix (bias gradients), fixed order
// Column sums, stage 1: block b adds rows [b rpb, (b+1) rpb) of x[rows][cols]
// (cols divides 256): thread t takes column t % cols and every (256/cols)-th
// row, so a wave reads whole rows; part[b][c] = the block's sum, added over
// the row groups in a fixed order.
__global__ void __launch_bounds__(256) k_colsum_part(const float* __restrict__ x, long rows, int cols, long rpb,
                                                     float* __restrict__ part) {
    __shared__ float red[256];
    const int c = threadIdx.x % cols, g = threadIdx.x / cols, ng = 256 / cols;
    const long r0 = (long)blockIdx.x * rpb, r1 = min(rows, r0 + rpb);
    float acc = 0.f;
    for (long r = r0 + g; r < r1; r += ng) acc += x[r * cols + c];
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < cols) {
        float s = 0.f;
        for (int q = 0; q < ng; ++q) s += red[q * cols + threadIdx.x];
        part[(long)blockIdx.x * cols + threadIdx.x] = s;
    }
}
// Split-K reduction, stage 1 of two: block (x, s) adds chunks
// [s cps, (s+1) cps) of part[chunk][n] for 256 consecutive elements;
// out2[s][e] = that sum (chunks in order).
__global__ void __launch_bounds__(256) k_sum_chunks_part(const float* __restrict__ part, int nchunk, int cps, long n,
                                                         float* __restrict__ out2) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const int c0 = blockIdx.y * cps, c1 = min(nchunk, c0 + cps);
    float acc = 0.f;
    for (int c = c0; c < c1; ++c) acc += part[(long)c * n + e];
    out2[(long)blockIdx.y * n + e] = acc;
}
// out[i] = sum_c part[c][i]  (split-K reduction)
__global__ void k_sum_chunks(const float* __restrict__ part, int nchunk, long n, float* __restrict__ out) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        float acc = 0.f;
        for (int c = 0; c < nchunk; ++c) acc += part[(long)c * n + e];
        out[e] = acc;
    }
}
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const float gg = g[e];
        const float mm = b1 * m[e] + (1.f - b1) * gg;
        const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
        m[e] = mm;
        v[e] = vv;
        w[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    }
}

struct Tensor {
    float* p = nullptr;
    size_t n = 0;
};

}  // namespace

// Parameter block layout (FP32, one contiguous array):
//   W_lift [cin][W], b_lift [W]
//   per layer: R_re [nm][W][W], R_im [nm][W][W], W_l [W][W], b_l [W]
//   W_q1 [W][H], b_q1 [H], W_q2 [H], b_q2 [1]
constexpr size_t kKpartFloats = (size_t)64 * 1024 * 1024;  // split-K / reduction scratch

struct gpe_fno {
    int nx, ny, nz, cin, width, layers, hidden, mx, my, mz, maxb;
    int nm;      // retained modes
    long P;      // points per sample
    size_t nparam = 0;
    size_t off_lift_w, off_lift_b, off_q1_w, off_q1_b, off_q2_w, off_q2_b;
    std::vector<size_t> off_rre, off_rim, off_w, off_b;
    float* w = nullptr;   // parameters
    float* g = nullptr;   // gradients
    float* am = nullptr;  // Adam moments
    float* av = nullptr;
    long adam_t = 0;
    // DFT matrices (re, im): Fz [mz][nz], Fy [my][ny], Fx [mx][nx], Gx [nx][mx], Gy [ny][my], Gz [nz][mz]
    float *fz[2], *fy[2], *fx[2], *gx[2], *gy[2], *gz[2];
    // activations for up to maxb samples
    float* xin = nullptr;
    std::vector<float*> h, zz, s3r, s3i;  // h[0..L], z[0..L-1], S3 per layer
    float *c1r, *c1i, *c2r, *c2i;          // complex scratch (largest: [B][nx][ny][mz][W])
    float *q_pre, *q, *y, *t;
    float *dh, *dz, *dq, *dy, *dtmp;       // backward scratch
    float* kpart = nullptr;                // split-K partials
    double* lpart = nullptr;
    // training set resident on the device (gpe_fno_set_data): inputs
    // [n][P][cin], targets [n][P]
    float* data_x = nullptr;
    float* data_t = nullptr;
    long data_n = 0;
    hipStream_t stream = nullptr;
    std::string err;
};

namespace {

int ffail(gpe_fno* f, const std::string& m) {
    if (f) f->err = m;
    return 1;
}

#define FCHECK(call)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return ffail(f, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int gemm(gpe_fno* f, const Args& a) {
    if (a.M <= 0 || a.N <= 0 || a.K <= 0 || a.batch <= 0) return 0;
    if (gpe::gemm::small_ok(a) && a.batch >= 256) {
        const long per = 256 / a.N;
        hipLaunchKernelGGL(gpe::gemm::k_gemm_small, dim3((unsigned)((a.batch + per - 1) / per)), dim3(256), 0,
                           f->stream, a);
        FCHECK(hipGetLastError());
        return 0;
    }
    dim3 grid((a.N + gpe::gemm::TN - 1) / gpe::gemm::TN, (a.M + gpe::gemm::TM - 1) / gpe::gemm::TM, a.batch);
    if (grid.z > 65535) return ffail(f, "gemm batch too large");
    hipLaunchKernelGGL(gpe::gemm::k_gemm, grid, dim3(gpe::gemm::NT), 0, f->stream, a);
    FCHECK(hipGetLastError());
    return 0;
}

Args mkargs(int M, int N, int K, int batch) {
    Args a;
    memset(&a, 0, sizeof(a));
    a.M = M;
    a.N = N;
    a.K = K;
    a.batch = batch;
    a.alpha = 1.f;
    return a;
}

// Row-major real GEMM (optionally batched by chunking the M rows for split-K
// is handled by callers): C[M x N] = A[M x K] . B[K x N] (+bias) (+beta C)
int gemm_rm(gpe_fno* f, const float* A, const float* B, float* C, long M, int N, int K, bool transA, bool transB,
            const float* bias, float beta) {
    if (!transA && N <= 64 && K <= 64 && ((uintptr_t)A % 16) == 0) {
        // many rows, small N and K: the row GEMM (same results as k_gemm)
        const long waves = ((M + 15) / 16 + gpe::gemm::kRowGroups - 1) / gpe::gemm::kRowGroups;
        const dim3 grid((unsigned)((waves + 3) / 4));
        if (transB)
            hipLaunchKernelGGL(gpe::gemm::k_gemm_rows<true>, grid, dim3(256), 0, f->stream, A, B, C, M, N, K, bias,
                               1.f, beta);
        else
            hipLaunchKernelGGL(gpe::gemm::k_gemm_rows<false>, grid, dim3(256), 0, f->stream, A, B, C, M, N, K, bias,
                               1.f, beta);
        FCHECK(hipGetLastError());
        return 0;
    }
    // one launch per 2^21 rows (grid.y = 32768 tiles of 64 rows at most)
    const long chunk = 1L << 21;
    for (long m0 = 0; m0 < M; m0 += chunk * 1) {
        const int mm = (int)std::min(chunk, M - m0);
        Args a = mkargs(mm, N, K, 1);
        if (!transA) {
            a.Ar = A + m0 * K;
            a.sam = K;
            a.sak = 1;
        } else {  // A stored [K][M]
            a.Ar = A + m0;
            a.sam = 1;
            a.sak = M;
        }
        if (!transB) {
            a.Br = B;
            a.sbk = N;
            a.sbn = 1;
        } else {  // B stored [N][K]
            a.Br = B;
            a.sbk = 1;
            a.sbn = K;
        }
        a.Cr = C + m0 * N;
        a.scm = N;
        a.scn = 1;
        a.bias = bias;
        a.beta = beta;
        if (gemm(f, a)) return 1;
    }
    return 0;
}

// out[e] = sum_c part[c][e] for nchunk chunks of n elements, in a fixed
// order: groups of 32 chunks in parallel, then the groups in order.  The
// group sums go to kpart after the nchunk * n partials.
int sum_chunks(gpe_fno* f, const float* part, int nchunk, long n, float* out) {
    const int cps = 32, ns = (nchunk + cps - 1) / cps;
    float* part2 = f->kpart + (size_t)nchunk * n;
    if ((size_t)(nchunk + ns) * n > kKpartFloats) return ffail(f, "split-K scratch too small");
    hipLaunchKernelGGL(k_sum_chunks_part, dim3((unsigned)((n + 255) / 256), ns), dim3(256), 0, f->stream, part, nchunk,
                       cps, n, part2);
    FCHECK(hipGetLastError());
    hipLaunchKernelGGL(k_sum_chunks, dim3(256), dim3(256), 0, f->stream, part2, ns, n, out);
    FCHECK(hipGetLastError());
    return 0;
}

// dW[K x N] = A^T[K x M] . D[M x N] over M = rows (large), split-K in chunks.
int gemm_wgrad(gpe_fno* f, const float* A, const float* D, float* dW, long M, int K, int N) {
    // 512-row chunks: M = 2^19 rows give 1024 blocks (one per chunk per tile)
    const long chunk = 512;
    const int nch = (int)((M + chunk - 1) / chunk);
    Args a = mkargs(K, N, (int)chunk, nch);
    a.Ar = A;
    a.sam = 1;
    a.sak = K;
    a.saz = chunk * K;
    a.Br = D;
    a.sbk = N;
    a.sbn = 1;
    a.sbz = chunk * N;
    a.Cr = f->kpart;
    a.scm = N;
    a.scn = 1;
    a.scz = (long)K * N;
    if (M % chunk != 0) return ffail(f, "rows must be a multiple of 512");
    if ((size_t)nch * K * N > kKpartFloats / 2) return ffail(f, "split-K scratch too small");
    if (K <= 64 && N <= 64 && ((uintptr_t)A % 16) == 0 && ((uintptr_t)D % 16) == 0) {
        // the chunk kernel: coalesced slab loads, same partials as k_gemm
        hipLaunchKernelGGL(gpe::gemm::k_gemm_wgrad, dim3(nch), dim3(256), 0, f->stream, A, D, f->kpart, (int)chunk, K,
                           N);
        FCHECK(hipGetLastError());
    } else if (gemm(f, a)) {
        return 1;
    }
    return sum_chunks(f, f->kpart, nch, (long)K * N, dW);
}

// out[c] = sum_r x[r][c] (cols divides 256), deterministic: per-block
// partials in kpart, then their fixed-order sum.
int colsum(gpe_fno* f, const float* x, long rows, int cols, float* out) {
    if (cols <= 0 || 256 % cols != 0) return ffail(f, "colsum: columns must divide 256");
    const int nb = 512;
    const long rpb = (rows + nb - 1) / nb;
    hipLaunchKernelGGL(k_colsum_part, dim3(nb), dim3(256), 0, f->stream, x, rows, cols, rpb, f->kpart);
    FCHECK(hipGetLastError());
    return sum_chunks(f, f->kpart, nb, (long)cols, out);
}

void fill_dft(std::vector<float>& re, std::vector<float>& im, int rows, int cols, const std::vector<int>& kl, int n,
              bool fwd_rows_are_k, double sign, const std::vector<double>& w) {
    re.assign((size_t)rows * cols, 0.f);
    im.assign((size_t)rows * cols, 0.f);
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) {
            const int k = fwd_rows_are_k ? kl[r] : kl[c];
            const int x = fwd_rows_are_k ? c : r;
            const double wt = w.empty() ? 1.0 : w[fwd_rows_are_k ? r : c];
            const double ang = sign * 2.0 * M_PI * (double)k * x / n;
            re[(size_t)r * cols + c] = (float)(wt * cos(ang));
            im[(size_t)r * cols + c] = (float)(wt * sin(ang));
        }
}

std::vector<int> mode_list(int n, int m) {
    std::vector<int> k;
    for (int t = 0; t < m / 2; ++t) k.push_back(t);
    for (int t = n - m / 2; t < n; ++t) k.push_back(t);
    return k;
}

// ---------------------------------------------------------------- forward
// complex activation layouts: C1 [B][nx][ny][mz][W], C2 [B][nx][my][mz][W],
// S3 [B][mx][my][mz][W]
int spectral_forward(gpe_fno* f, int l, int B, const float* hin, float* sout) {
    const int nx = f->nx, ny = f->ny, nz = f->nz, W = f->width, mx = f->mx, my = f->my, mz = f->mz;
    // rdft_z: per (b,x,y): [mz x nz] . [nz x W]
    Args a = mkargs(mz, W, nz, B * nx * ny);
    a.Ar = f->fz[0];
    a.Ai = f->fz[1];
    a.sam = nz;
    a.sak = 1;
    a.Br = hin;
    a.sbk = W;
    a.sbn = 1;
    a.sbz = (long)nz * W;
    a.Cr = f->c1r;
    a.Ci = f->c1i;
    a.scm = W;
    a.scn = 1;
    a.scz = (long)mz * W;
    if (gemm(f, a)) return 1;
    // dft_y: per (b,x): [my x ny] . [ny x (mz W)]
    a = mkargs(my, mz * W, ny, B * nx);
    a.Ar = f->fy[0];
    a.Ai = f->fy[1];
    a.sam = ny;
    a.sak = 1;
    a.Br = f->c1r;
    a.Bi = f->c1i;
    a.sbk = (long)mz * W;
    a.sbn = 1;
    a.sbz = (long)ny * mz * W;
    a.Cr = f->c2r;
    a.Ci = f->c2i;
    a.scm = (long)mz * W;
    a.scn = 1;
    a.scz = (long)my * mz * W;
    if (gemm(f, a)) return 1;
    // dft_x: per b: [mx x nx] . [nx x (my mz W)]
    a = mkargs(mx, my * mz * W, nx, B);
    a.Ar = f->fx[0];
    a.Ai = f->fx[1];
    a.sam = nx;
    a.sak = 1;
    a.Br = f->c2r;
    a.Bi = f->c2i;
    a.sbk = (long)my * mz * W;
    a.sbn = 1;
    a.sbz = (long)nx * my * mz * W;
    a.Cr = f->s3r[l];
    a.Ci = f->s3i[l];
    a.scm = (long)my * mz * W;
    a.scn = 1;
    a.scz = (long)mx * my * mz * W;
    if (gemm(f, a)) return 1;
    // mix: per mode m: [B x W] . R[m] [W x W]  -> C1 reused as [B][nm][W]
    const int nm = f->nm;
    a = mkargs(B, W, W, nm);
    a.Ar = f->s3r[l];
    a.Ai = f->s3i[l];
    a.sam = (long)nm * W;
    a.sak = 1;
    a.saz = W;
    a.Br = f->w + f->off_rre[l];
    a.Bi = f->w + f->off_rim[l];
    a.sbk = W;
    a.sbn = 1;
    a.sbz = (long)W * W;
    a.Cr = f->c1r;
    a.Ci = f->c1i;
    a.scm = (long)nm * W;
    a.scn = 1;
    a.scz = W;
    if (gemm(f, a)) return 1;
    // idft_x: per b: [nx x mx] . [mx x (my mz W)] -> C2 as [B][nx][my][mz][W]
    a = mkargs(nx, my * mz * W, mx, B);
    a.Ar = f->gx[0];
    a.Ai = f->gx[1];
    a.sam = mx;
    a.sak = 1;
    a.Br = f->c1r;
    a.Bi = f->c1i;
    a.sbk = (long)my * mz * W;
    a.sbn = 1;
    a.sbz = (long)mx * my * mz * W;
    a.Cr = f->c2r;
    a.Ci = f->c2i;
    a.scm = (long)my * mz * W;
    a.scn = 1;
    a.scz = (long)nx * my * mz * W;
    if (gemm(f, a)) return 1;
    // idft_y: per (b,x): [ny x my] . [my x (mz W)] -> C1 as [B][nx][ny][mz][W]
    a = mkargs(ny, mz * W, my, B * nx);
    a.Ar = f->gy[0];
    a.Ai = f->gy[1];
    a.sam = my;
    a.sak = 1;
    a.Br = f->c2r;
    a.Bi = f->c2i;
    a.sbk = (long)mz * W;
    a.sbn = 1;
    a.sbz = (long)my * mz * W;
    a.Cr = f->c1r;
    a.Ci = f->c1i;
    a.scm = (long)mz * W;
    a.scn = 1;
    a.scz = (long)ny * mz * W;
    if (gemm(f, a)) return 1;
    // irdft_z: per (b,x,y): Re([nz x mz] . [mz x W]) -> sout [B][nx][ny][nz][W]
    a = mkargs(nz, W, mz, B * nx * ny);
    a.Ar = f->gz[0];
    a.Ai = f->gz[1];
    a.sam = mz;
    a.sak = 1;
    a.Br = f->c1r;
    a.Bi = f->c1i;
    a.sbk = W;
    a.sbn = 1;
    a.sbz = (long)mz * W;
    a.Cr = sout;
    a.scm = W;
    a.scn = 1;
    a.scz = (long)nz * W;
    return gemm(f, a);
}

// Adjoint of spectral_forward w.r.t. its input: dh += Re(F^H ...) and the
// weight gradient dR.  ds: [B][nx][ny][nz][W] real.
int spectral_backward(gpe_fno* f, int l, int B, const float* ds, float* dh_accum) {
    const int nx = f->nx, ny = f->ny, nz = f->nz, W = f->width, mx = f->mx, my = f->my, mz = f->mz, nm = f->nm;
    // dC1 = Gz^H ds : per (b,x,y): [mz x nz] (conj transpose of Gz) . [nz x W]
    Args a = mkargs(mz, W, nz, B * nx * ny);
    a.Ar = f->gz[0];
    a.Ai = f->gz[1];
    a.conjA = 1;
    a.sam = 1;
    a.sak = mz;
    a.Br = ds;
    a.sbk = W;
    a.sbn = 1;
    a.sbz = (long)nz * W;
    a.Cr = f->c1r;
    a.Ci = f->c1i;
    a.scm = W;
    a.scn = 1;
    a.scz = (long)mz * W;
    if (gemm(f, a)) return 1;
    // dC2 = Gy^H dC1: per (b,x): [my x ny] . [ny x mz W]
    a = mkargs(my, mz * W, ny, B * nx);
    a.Ar = f->gy[0];
    a.Ai = f->gy[1];
    a.conjA = 1;
    a.sam = 1;
    a.sak = my;
    a.Br = f->c1r;
    a.Bi = f->c1i;
    a.sbk = (long)mz * W;
    a.sbn = 1;
    a.sbz = (long)ny * mz * W;
    a.Cr = f->c2r;
    a.Ci = f->c2i;
    a.scm = (long)mz * W;
    a.scn = 1;
    a.scz = (long)my * mz * W;
    if (gemm(f, a)) return 1;
    // dS4 = Gx^H dC2: per b: [mx x nx] . [nx x my mz W] -> dtmp (re) / q (im) as [B][nm][W]
    float* d4r = f->dtmp;
    float* d4i = f->dtmp + (size_t)f->maxb * nm * W;
    a = mkargs(mx, my * mz * W, nx, B);
    a.Ar = f->gx[0];
    a.Ai = f->gx[1];
    a.conjA = 1;
    a.sam = 1;
    a.sak = mx;
    a.Br = f->c2r;
    a.Bi = f->c2i;
    a.sbk = (long)my * mz * W;
    a.sbn = 1;
    a.sbz = (long)nx * my * mz * W;
    a.Cr = d4r;
    a.Ci = d4i;
    a.scm = (long)my * mz * W;
    a.scn = 1;
    a.scz = (long)nm * W;
    if (gemm(f, a)) return 1;
    // dR[m] = S3[:,m]^H . dS4[:,m] : [W x B] . [B x W]
    a = mkargs(W, W, B, nm);
    a.Ar = f->s3r[l];
    a.Ai = f->s3i[l];
    a.conjA = 1;
    a.sam = 1;
    a.sak = (long)nm * W;
    a.saz = W;
    a.Br = d4r;
    a.Bi = d4i;
    a.sbk = (long)nm * W;
    a.sbn = 1;
    a.sbz = W;
    a.Cr = f->g + f->off_rre[l];
    a.Ci = f->g + f->off_rim[l];
    a.scm = W;
    a.scn = 1;
    a.scz = (long)W * W;
    if (gemm(f, a)) return 1;
    // dS3 = dS4 . R^H : per mode [B x W] . [W x W]^H -> C1 as [B][nm][W]
    a = mkargs(B, W, W, nm);
    a.Ar = d4r;
    a.Ai = d4i;
    a.sam = (long)nm * W;
    a.sak = 1;
    a.saz = W;
    a.Br = f->w + f->off_rre[l];
    a.Bi = f->w + f->off_rim[l];
    a.conjB = 1;
    a.sbk = 1;
    a.sbn = W;
    a.sbz = (long)W * W;
    a.Cr = f->c1r;
    a.Ci = f->c1i;
    a.scm = (long)nm * W;
    a.scn = 1;
    a.scz = W;
    if (gemm(f, a)) return 1;
    // dC2 = Fx^H dS3: per b: [nx x mx] . [mx x my mz W]
    a = mkargs(nx, my * mz * W, mx, B);
    a.Ar = f->fx[0];
    a.Ai = f->fx[1];
    a.conjA = 1;
    a.sam = 1;
    a.sak = nx;
    a.Br = f->c1r;
    a.Bi = f->c1i;
    a.sbk = (long)my * mz * W;
    a.sbn = 1;
    a.sbz = (long)nm * W;
    a.Cr = f->c2r;
    a.Ci = f->c2i;
    a.scm = (long)my * mz * W;
    a.scn = 1;
    a.scz = (long)nx * my * mz * W;
    if (gemm(f, a)) return 1;
    // dC1 = Fy^H dC2: per (b,x): [ny x my] . [my x mz W]
    a = mkargs(ny, mz * W, my, B * nx);
    a.Ar = f->fy[0];
    a.Ai = f->fy[1];
    a.conjA = 1;
    a.sam = 1;
    a.sak = ny;
    a.Br = f->c2r;
    a.Bi = f->c2i;
    a.sbk = (long)mz * W;
    a.sbn = 1;
    a.sbz = (long)my * mz * W;
    a.Cr = f->c1r;
    a.Ci = f->c1i;
    a.scm = (long)mz * W;
    a.scn = 1;
    a.scz = (long)ny * mz * W;
    if (gemm(f, a)) return 1;
    // dh += Re(Fz^H dC1): per (b,x,y): [nz x mz] . [mz x W]
    a = mkargs(nz, W, mz, B * nx * ny);
    a.Ar = f->fz[0];
    a.Ai = f->fz[1];
    a.conjA = 1;
    a.sam = 1;
    a.sak = nz;
    a.Br = f->c1r;
    a.Bi = f->c1i;
    a.sbk = W;
    a.sbn = 1;
    a.sbz = (long)mz * W;
    a.Cr = dh_accum;
    a.scm = W;
    a.scn = 1;
    a.scz = (long)nz * W;
    a.beta = 1.f;
    return gemm(f, a);
}

int fno_forward(gpe_fno* f, int B) {
    const long R = (long)B * f->P;
    const int W = f->width;
    if (gemm_rm(f, f->xin, f->w + f->off_lift_w, f->h[0], R, W, f->cin, false, false, f->w + f->off_lift_b, 0.f))
        return 1;
    for (int l = 0; l < f->layers; ++l) {
        if (spectral_forward(f, l, B, f->h[l], f->zz[l])) return 1;
        if (gemm_rm(f, f->h[l], f->w + f->off_w[l], f->zz[l], R, W, W, false, false, f->w + f->off_b[l], 1.f))
            return 1;
        if (l < f->layers - 1) {
            hipLaunchKernelGGL(k_gelu, dim3(2048), dim3(256), 0, f->stream, f->zz[l], f->h[l + 1], (size_t)(R * W));
        } else {
            FCHECK(hipMemcpyAsync(f->h[l + 1], f->zz[l], (size_t)R * W * 4, hipMemcpyDeviceToDevice, f->stream));
        }
    }
    const int H = f->hidden;
    if (gemm_rm(f, f->h[f->layers], f->w + f->off_q1_w, f->q_pre, R, H, W, false, false, f->w + f->off_q1_b, 0.f))
        return 1;
    hipLaunchKernelGGL(k_gelu, dim3(2048), dim3(256), 0, f->stream, f->q_pre, f->q, (size_t)(R * H));
    if (gemm_rm(f, f->q, f->w + f->off_q2_w, f->y, R, 1, H, false, false, f->w + f->off_q2_b, 0.f)) return 1;
    FCHECK(hipGetLastError());
    return 0;
}

int fno_backward(gpe_fno* f, int B) {
    const long R = (long)B * f->P;
    const int W = f->width, H = f->hidden;
    FCHECK(hipMemsetAsync(f->g, 0, f->nparam * 4, f->stream));
    // y = q Wq2 + bq2
    if (gemm_wgrad(f, f->q, f->dy, f->g + f->off_q2_w, R, H, 1)) return 1;
    if (colsum(f, f->dy, R, 1, f->g + f->off_q2_b)) return 1;
    if (gemm_rm(f, f->dy, f->w + f->off_q2_w, f->dq, R, H, 1, false, true, nullptr, 0.f)) return 1;
    hipLaunchKernelGGL(k_gelu_bwd, dim3(2048), dim3(256), 0, f->stream, f->dq, f->q_pre, f->dq, (size_t)(R * H));
    // q_pre = h_L Wq1 + bq1
    if (gemm_wgrad(f, f->h[f->layers], f->dq, f->g + f->off_q1_w, R, W, H)) return 1;
    if (colsum(f, f->dq, R, H, f->g + f->off_q1_b)) return 1;
    if (gemm_rm(f, f->dq, f->w + f->off_q1_w, f->dh, R, W, H, false, true, nullptr, 0.f)) return 1;
    for (int l = f->layers - 1; l >= 0; --l) {
        if (l < f->layers - 1)
            hipLaunchKernelGGL(k_gelu_bwd, dim3(2048), dim3(256), 0, f->stream, f->dh, f->zz[l], f->dz,
                               (size_t)(R * W));
        else
            FCHECK(hipMemcpyAsync(f->dz, f->dh, (size_t)R * W * 4, hipMemcpyDeviceToDevice, f->stream));
        if (gemm_wgrad(f, f->h[l], f->dz, f->g + f->off_w[l], R, W, W)) return 1;
        if (colsum(f, f->dz, R, W, f->g + f->off_b[l])) return 1;
        if (gemm_rm(f, f->dz, f->w + f->off_w[l], f->dh, R, W, W, false, true, nullptr, 0.f)) return 1;
        if (spectral_backward(f, l, B, f->dz, f->dh)) return 1;
    }
    if (gemm_wgrad(f, f->xin, f->dh, f->g + f->off_lift_w, R, f->cin, W)) return 1;
    if (colsum(f, f->dh, R, W, f->g + f->off_lift_b)) return 1;
    FCHECK(hipGetLastError());
    return 0;
}

int fno_loss(gpe_fno* f, int B, bool grad, double* loss) {
    const size_t n = (size_t)B * f->P;
    hipLaunchKernelGGL(k_mse, dim3(256), dim3(256), 0, f->stream, f->y, f->t, grad ? f->dy : nullptr, n, f->lpart);
    FCHECK(hipGetLastError());
    std::vector<double> part(256);
    FCHECK(hipMemcpyAsync(part.data(), f->lpart, 256 * 8, hipMemcpyDeviceToHost, f->stream));
    FCHECK(hipStreamSynchronize(f->stream));
    double s = 0;
    for (double v : part) s += v;
    if (loss) *loss = s / (double)n;
    return 0;
}

void init_weights(gpe_fno* f, unsigned long long seed) {
    std::vector<float> w(f->nparam, 0.f);
    unsigned long long s = seed;
    auto uni = [&]() {
        s += 0x9E3779B97F4A7C15ULL;
        unsigned long long z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z = z ^ (z >> 31);
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](size_t off, size_t n, double bound) {
        for (size_t t = 0; t < n; ++t) w[off + t] = (float)((2 * uni() - 1) * bound);
    };
    const int W = f->width, H = f->hidden;
    fill(f->off_lift_w, (size_t)f->cin * W, 1.0 / sqrt((double)f->cin));
    fill(f->off_lift_b, W, 1.0 / sqrt((double)f->cin));
    for (int l = 0; l < f->layers; ++l) {
        const double sc = 1.0 / ((double)W * W);
        for (size_t t = 0; t < (size_t)f->nm * W * W; ++t) {
            w[f->off_rre[l] + t] = (float)(sc * uni());
            w[f->off_rim[l] + t] = (float)(sc * uni());
        }
        fill(f->off_w[l], (size_t)W * W, 1.0 / sqrt((double)W));
        fill(f->off_b[l], W, 1.0 / sqrt((double)W));
    }
    fill(f->off_q1_w, (size_t)W * H, 1.0 / sqrt((double)W));
    fill(f->off_q1_b, H, 1.0 / sqrt((double)W));
    fill(f->off_q2_w, H, 1.0 / sqrt((double)H));
    fill(f->off_q2_b, 1, 1.0 / sqrt((double)H));
    (void)hipMemcpy(f->w, w.data(), f->nparam * 4, hipMemcpyHostToDevice);
}

}  // namespace

extern "C" {

int gpe_fno_create(const gpe_fno_config* c, unsigned long long seed, gpe_fno** out) {
    if (!c || !out) return 1;
    *out = nullptr;
    gpe_fno* f = new gpe_fno();
    f->nx = c->nx;
    f->ny = c->ny;
    f->nz = c->nz;
    f->cin = c->cin;
    f->width = c->width;
    f->layers = c->layers;
    f->hidden = c->hidden;
    f->mx = std::min(c->modes, c->nx);
    f->my = std::min(c->modes, c->ny);
    f->mz = std::min(c->modes, c->nz / 2 + 1);
    f->maxb = c->max_batch;
    f->P = (long)c->nx * c->ny * c->nz;
    f->nm = f->mx * f->my * f->mz;
    auto bad = [&](const std::string& m) {
        f->err = m;
        *out = f;  // caller reads the message, then destroys
        return 1;
    };
    if (f->mx % 2 || f->my % 2 || f->width <= 0 || f->layers <= 0 || f->maxb <= 0 || (f->maxb * f->P) % 8192)
        return bad("bad FNO configuration (even modes, batch*points multiple of 8192)");
    if (hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess) return bad("stream");
    const int W = f->width, H = f->hidden;
    size_t o = 0;
    f->off_lift_w = o;
    o += (size_t)f->cin * W;
    f->off_lift_b = o;
    o += W;
    for (int l = 0; l < f->layers; ++l) {
        f->off_rre.push_back(o);
        o += (size_t)f->nm * W * W;
        f->off_rim.push_back(o);
        o += (size_t)f->nm * W * W;
        f->off_w.push_back(o);
        o += (size_t)W * W;
        f->off_b.push_back(o);
        o += W;
    }
    f->off_q1_w = o;
    o += (size_t)W * H;
    f->off_q1_b = o;
    o += H;
    f->off_q2_w = o;
    o += H;
    f->off_q2_b = o;
    o += 1;
    f->nparam = o;
    const long R = (long)f->maxb * f->P;
    const size_t cbig = (size_t)f->maxb * f->nx * f->ny * f->mz * W;
    const size_t cmid = (size_t)f->maxb * f->nx * f->my * f->mz * W;
    auto A = [&](float** p, size_t n) { return hipMalloc(p, std::max<size_t>(n, 1) * 4) == hipSuccess; };
    bool ok = A(&f->w, o) && A(&f->g, o) && A(&f->am, o) && A(&f->av, o) && A(&f->xin, (size_t)R * f->cin);
    f->h.assign(f->layers + 1, nullptr);
    f->zz.assign(f->layers, nullptr);
    f->s3r.assign(f->layers, nullptr);
    f->s3i.assign(f->layers, nullptr);
    for (int l = 0; l <= f->layers && ok; ++l) ok = A(&f->h[l], (size_t)R * W);
    for (int l = 0; l < f->layers && ok; ++l)
        ok = A(&f->zz[l], (size_t)R * W) && A(&f->s3r[l], (size_t)f->maxb * f->nm * W) &&
             A(&f->s3i[l], (size_t)f->maxb * f->nm * W);
    ok = ok && A(&f->c1r, cbig) && A(&f->c1i, cbig) && A(&f->c2r, std::max(cmid, cbig)) &&
         A(&f->c2i, std::max(cmid, cbig)) && A(&f->q_pre, (size_t)R * H) && A(&f->q, (size_t)R * H) &&
         A(&f->y, R) && A(&f->t, R) && A(&f->dh, (size_t)R * W) && A(&f->dz, (size_t)R * W) &&
         A(&f->dq, (size_t)R * H) && A(&f->dy, R) && A(&f->dtmp, 2 * (size_t)f->maxb * f->nm * W) &&
         A(&f->kpart, kKpartFloats) && hipMalloc(&f->lpart, 256 * 8) == hipSuccess;
    for (int t = 0; t < 2 && ok; ++t)
        ok = A(&f->fz[t], (size_t)f->mz * f->nz) && A(&f->fy[t], (size_t)f->my * f->ny) &&
             A(&f->fx[t], (size_t)f->mx * f->nx) && A(&f->gx[t], (size_t)f->nx * f->mx) &&
             A(&f->gy[t], (size_t)f->ny * f->my) && A(&f->gz[t], (size_t)f->nz * f->mz);
    if (!ok) return bad("FNO device allocation failed");
    // DFT matrices
    std::vector<int> kx = mode_list(f->nx, f->mx), ky = mode_list(f->ny, f->my), kz;
    for (int t = 0; t < f->mz; ++t) kz.push_back(t);
    std::vector<float> re, im;
    auto up = [&](float** d) {
        (void)hipMemcpy(d[0], re.data(), re.size() * 4, hipMemcpyHostToDevice);
        (void)hipMemcpy(d[1], im.data(), im.size() * 4, hipMemcpyHostToDevice);
    };
    fill_dft(re, im, f->mz, f->nz, kz, f->nz, true, -1.0, {});
    up(f->fz);
    fill_dft(re, im, f->my, f->ny, ky, f->ny, true, -1.0, {});
    up(f->fy);
    fill_dft(re, im, f->mx, f->nx, kx, f->nx, true, -1.0, {});
    up(f->fx);
    fill_dft(re, im, f->nx, f->mx, kx, f->nx, false, +1.0, {});
    up(f->gx);
    fill_dft(re, im, f->ny, f->my, ky, f->ny, false, +1.0, {});
    up(f->gy);
    // c2r along z: Hermitian weights (1 for kz = 0 and the Nyquist bin, else 2)
    // and the 1/(nx ny nz) normalisation of the inverse transform
    std::vector<double> wz(f->mz);
    for (int t = 0; t < f->mz; ++t)
        wz[t] = ((t == 0 || 2 * t == f->nz) ? 1.0 : 2.0) / ((double)f->nx * f->ny * f->nz);
    fill_dft(re, im, f->nz, f->mz, kz, f->nz, false, +1.0, wz);
    up(f->gz);
    init_weights(f, seed);
    if (hipMemset(f->am, 0, o * 4) != hipSuccess || hipMemset(f->av, 0, o * 4) != hipSuccess) return bad("memset");
    if (hipDeviceSynchronize() != hipSuccess) return bad("device synchronize after FNO set-up");
    *out = f;
    return 0;
}

void gpe_fno_destroy(gpe_fno* f) {
    if (!f) return;
    if (f->stream) (void)hipStreamSynchronize(f->stream);
    std::vector<void*> bufs = {f->w,  f->g,     f->am,    f->av, f->xin, f->c1r, f->c1i, f->c2r, f->c2i, f->q_pre,
                               f->q,  f->y,     f->t,     f->dh, f->dz,  f->dq,  f->dy,  f->dtmp, f->kpart, f->lpart,
                               f->data_x, f->data_t};
    for (auto* p : f->h) bufs.push_back(p);
    for (auto* p : f->zz) bufs.push_back(p);
    for (auto* p : f->s3r) bufs.push_back(p);
    for (auto* p : f->s3i) bufs.push_back(p);
    for (int t = 0; t < 2; ++t)
        for (float* p : {f->fz[t], f->fy[t], f->fx[t], f->gx[t], f->gy[t], f->gz[t]}) bufs.push_back(p);
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (f->stream) (void)hipStreamDestroy(f->stream);
    delete f;
}

const char* gpe_fno_last_error(const gpe_fno* f) { return f ? f->err.c_str() : ""; }
long gpe_fno_num_params(const gpe_fno* f) { return f ? (long)f->nparam : 0; }

// Host copies of the parameters and gradients go through the network's own
// stream: it is non-blocking, so a copy on the null stream would not wait for
// the launches still running on it (a backward pass enqueued by
// gpe_fno_loss_grad writes its last gradients after that call returns).
int gpe_fno_get_params(gpe_fno* f, float* host) {
    FCHECK(hipMemcpyAsync(host, f->w, f->nparam * 4, hipMemcpyDeviceToHost, f->stream));
    FCHECK(hipStreamSynchronize(f->stream));
    return 0;
}
int gpe_fno_set_params(gpe_fno* f, const float* host) {
    FCHECK(hipMemcpyAsync(f->w, host, f->nparam * 4, hipMemcpyHostToDevice, f->stream));
    FCHECK(hipStreamSynchronize(f->stream));
    return 0;
}
int gpe_fno_get_grads(gpe_fno* f, float* host) {
    FCHECK(hipMemcpyAsync(host, f->g, f->nparam * 4, hipMemcpyDeviceToHost, f->stream));
    FCHECK(hipStreamSynchronize(f->stream));
    return 0;
}

// x: [B][P][cin] on the host, y_out: [B][P]
int gpe_fno_forward(gpe_fno* f, const float* x, int B, float* y_out) {
    if (!f || B <= 0 || B > f->maxb) return ffail(f, "bad batch size");
    FCHECK(hipMemcpyAsync(f->xin, x, (size_t)B * f->P * f->cin * 4, hipMemcpyHostToDevice, f->stream));
    if (fno_forward(f, B)) return 1;
    if (y_out) FCHECK(hipMemcpyAsync(y_out, f->y, (size_t)B * f->P * 4, hipMemcpyDeviceToHost, f->stream));
    FCHECK(hipStreamSynchronize(f->stream));
    return 0;
}

// Forward + loss + backward (gradients left in the gradient buffer).
int gpe_fno_loss_grad(gpe_fno* f, const float* x, const float* target, int B, double* loss) {
    if (!f || B <= 0 || B > f->maxb) return ffail(f, "bad batch size");
    if (((long)B * f->P) % 8192) return ffail(f, "batch*points must be a multiple of 8192");
    FCHECK(hipMemcpyAsync(f->xin, x, (size_t)B * f->P * f->cin * 4, hipMemcpyHostToDevice, f->stream));
    FCHECK(hipMemcpyAsync(f->t, target, (size_t)B * f->P * 4, hipMemcpyHostToDevice, f->stream));
    if (fno_forward(f, B)) return 1;
    if (fno_loss(f, B, true, loss)) return 1;
    return fno_backward(f, B);
}

float* gpe_fno_input_dev(gpe_fno* f) { return f ? f->xin : nullptr; }
const float* gpe_fno_output_dev(gpe_fno* f) { return f ? f->y : nullptr; }

int gpe_fno_forward_dev(gpe_fno* f, int B) {
    if (!f || B <= 0 || B > f->maxb) return ffail(f, "bad batch size");
    if (fno_forward(f, B)) return 1;
    FCHECK(hipStreamSynchronize(f->stream));
    return 0;
}

int gpe_fno_set_data(gpe_fno* f, const float* x, const float* target, long n) {
    if (!f || n <= 0 || !x || !target) return ffail(f, "bad training set");
    FCHECK(hipStreamSynchronize(f->stream));  // no launch may still read the old set
    if (f->data_x) FCHECK(hipFree(f->data_x));
    if (f->data_t) FCHECK(hipFree(f->data_t));
    f->data_x = f->data_t = nullptr;
    f->data_n = 0;
    const size_t nx = (size_t)n * f->P * f->cin, nt = (size_t)n * f->P;
    FCHECK(hipMalloc(&f->data_x, nx * 4));
    FCHECK(hipMalloc(&f->data_t, nt * 4));
    FCHECK(hipMemcpyAsync(f->data_x, x, nx * 4, hipMemcpyHostToDevice, f->stream));
    FCHECK(hipMemcpyAsync(f->data_t, target, nt * 4, hipMemcpyHostToDevice, f->stream));
    FCHECK(hipStreamSynchronize(f->stream));
    f->data_n = n;
    return 0;
}

int gpe_fno_loss_grad_idx(gpe_fno* f, const int* idx, int B, double* loss) {
    if (!f || B <= 0 || B > f->maxb || !idx) return ffail(f, "bad batch size");
    if (((long)B * f->P) % 8192) return ffail(f, "batch*points must be a multiple of 8192");
    const size_t sx = (size_t)f->P * f->cin * 4, st = (size_t)f->P * 4;
    for (int q = 0; q < B; ++q) {
        if (idx[q] < 0 || idx[q] >= f->data_n) return ffail(f, "sample index outside the training set");
        FCHECK(hipMemcpyAsync((char*)f->xin + q * sx, (const char*)f->data_x + idx[q] * sx, sx,
                              hipMemcpyDeviceToDevice, f->stream));
        FCHECK(hipMemcpyAsync((char*)f->t + q * st, (const char*)f->data_t + idx[q] * st, st,
                              hipMemcpyDeviceToDevice, f->stream));
    }
    if (fno_forward(f, B)) return 1;
    if (fno_loss(f, B, true, loss)) return 1;
    return fno_backward(f, B);
}

int gpe_fno_adam_step(gpe_fno* f, float lr) {
    f->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)f->adam_t), c2 = 1.f - powf(b2, (float)f->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(1024), dim3(256), 0, f->stream, f->w, f->g, f->am, f->av, (long)f->nparam, lr,
                       b1, b2, 1e-8f, c1, c2);
    FCHECK(hipGetLastError());
    FCHECK(hipStreamSynchronize(f->stream));
    return 0;
}

// Row GEMM entry used by the tests: C = A op(B) (+ bias) + beta C on the row
// kernel (rows = 1) or the tile kernel (rows = 0), same operands; host
// row-major matrices, B [K][N] or (transB) [N][K], C in / out.
int gpe_gemm_rows_test(const float* A, const float* B, const float* bias, float* C, long M, int N, int K, int transB,
                       float beta, int rows) {
    if (M <= 0 || N <= 0 || N > 64 || K <= 0 || K > 64) return 1;
    float *da = nullptr, *db = nullptr, *dc = nullptr, *dbias = nullptr;
    int rc = 0;
    rc |= hipMalloc(&da, (size_t)M * K * 4) != hipSuccess;
    rc |= hipMalloc(&db, (size_t)K * N * 4) != hipSuccess;
    rc |= hipMalloc(&dc, (size_t)M * N * 4) != hipSuccess;
    if (bias) rc |= hipMalloc(&dbias, (size_t)N * 4) != hipSuccess;
    if (!rc) {
        (void)hipMemcpy(da, A, (size_t)M * K * 4, hipMemcpyHostToDevice);
        (void)hipMemcpy(db, B, (size_t)K * N * 4, hipMemcpyHostToDevice);
        (void)hipMemcpy(dc, C, (size_t)M * N * 4, hipMemcpyHostToDevice);
        if (bias) (void)hipMemcpy(dbias, bias, (size_t)N * 4, hipMemcpyHostToDevice);
        if (rows) {
            const long waves = ((M + 15) / 16 + gpe::gemm::kRowGroups - 1) / gpe::gemm::kRowGroups;
            const dim3 grid((unsigned)((waves + 3) / 4));
            if (transB)
                hipLaunchKernelGGL(gpe::gemm::k_gemm_rows<true>, grid, dim3(256), 0, 0, da, db, dc, M, N, K, dbias,
                                   1.f, beta);
            else
                hipLaunchKernelGGL(gpe::gemm::k_gemm_rows<false>, grid, dim3(256), 0, 0, da, db, dc, M, N, K, dbias,
                                   1.f, beta);
        } else {
            Args a = mkargs((int)M, N, K, 1);
            a.Ar = da;
            a.sam = K;
            a.sak = 1;
            a.Br = db;
            a.sbk = transB ? 1 : N;
            a.sbn = transB ? K : 1;
            a.Cr = dc;
            a.scm = N;
            a.scn = 1;
            a.bias = dbias;
            a.beta = beta;
            dim3 grid((N + 63) / 64, (unsigned)((M + 63) / 64), 1);
            hipLaunchKernelGGL(gpe::gemm::k_gemm, grid, dim3(256), 0, 0, a);
        }
        rc = hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess;
        if (!rc) (void)hipMemcpy(C, dc, (size_t)M * N * 4, hipMemcpyDeviceToHost);
    }
    for (float* p : {da, db, dc, dbias})
        if (p) (void)hipFree(p);
    return rc;
}

// Weight-gradient partials entry used by the tests: P[z] = A_z^T D_z for
// `chunks` chunks of 512 rows (A [chunks 512][K], D [chunks 512][N]) on the
// chunk kernel (wg = 1) or the tile kernel (wg = 0).
int gpe_gemm_wgrad_test(const float* A, const float* D, float* P, int chunks, int K, int N, int wg) {
    if (chunks <= 0 || K <= 0 || K > 64 || N <= 0 || N > 64) return 1;
    const long rows = 512;
    float *da = nullptr, *dd = nullptr, *dp = nullptr;
    int rc = 0;
    rc |= hipMalloc(&da, (size_t)chunks * rows * K * 4) != hipSuccess;
    rc |= hipMalloc(&dd, (size_t)chunks * rows * N * 4) != hipSuccess;
    rc |= hipMalloc(&dp, (size_t)chunks * K * N * 4) != hipSuccess;
    if (!rc) {
        (void)hipMemcpy(da, A, (size_t)chunks * rows * K * 4, hipMemcpyHostToDevice);
        (void)hipMemcpy(dd, D, (size_t)chunks * rows * N * 4, hipMemcpyHostToDevice);
        if (wg) {
            hipLaunchKernelGGL(gpe::gemm::k_gemm_wgrad, dim3(chunks), dim3(256), 0, 0, da, dd, dp, (int)rows, K, N);
        } else {
            Args a = mkargs(K, N, (int)rows, chunks);
            a.Ar = da;
            a.sam = 1;
            a.sak = K;
            a.saz = rows * K;
            a.Br = dd;
            a.sbk = N;
            a.sbn = 1;
            a.sbz = rows * N;
            a.Cr = dp;
            a.scm = N;
            a.scn = 1;
            a.scz = (long)K * N;
            dim3 grid((N + 63) / 64, (K + 63) / 64, chunks);
            hipLaunchKernelGGL(gpe::gemm::k_gemm, grid, dim3(256), 0, 0, a);
        }
        rc = hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess;
        if (!rc) (void)hipMemcpy(P, dp, (size_t)chunks * K * N * 4, hipMemcpyDeviceToHost);
    }
    for (float* p : {da, dd, dp})
        if (p) (void)hipFree(p);
    return rc;
}

// Dense GEMM entry used by the tests (C = A B, all row-major, real or complex).
int gpe_gemm_test(const float* Ar, const float* Ai, const float* Br, const float* Bi, float* Cr, float* Ci, int M,
                  int N, int K, int conjA, int conjB) {
    float *dar = nullptr, *dai = nullptr, *dbr = nullptr, *dbi = nullptr, *dcr = nullptr, *dci = nullptr;
    auto A = [&](float** p, size_t n) { return hipMalloc(p, n * 4) == hipSuccess; };
    if (!A(&dar, (size_t)M * K) || !A(&dbr, (size_t)K * N) || !A(&dcr, (size_t)M * N)) return 1;
    if (Ai && !A(&dai, (size_t)M * K)) return 1;
    if (Bi && !A(&dbi, (size_t)K * N)) return 1;
    if (Ci && !A(&dci, (size_t)M * N)) return 1;
    (void)hipMemcpy(dar, Ar, (size_t)M * K * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(dbr, Br, (size_t)K * N * 4, hipMemcpyHostToDevice);
    if (Ai) (void)hipMemcpy(dai, Ai, (size_t)M * K * 4, hipMemcpyHostToDevice);
    if (Bi) (void)hipMemcpy(dbi, Bi, (size_t)K * N * 4, hipMemcpyHostToDevice);
    Args a = mkargs(M, N, K, 1);
    a.Ar = dar;
    a.Ai = dai;
    a.sam = K;
    a.sak = 1;
    a.Br = dbr;
    a.Bi = dbi;
    a.sbk = N;
    a.sbn = 1;
    a.Cr = dcr;
    a.Ci = dci;
    a.scm = N;
    a.scn = 1;
    a.conjA = conjA;
    a.conjB = conjB;
    dim3 grid((N + 63) / 64, (M + 63) / 64, 1);
    hipLaunchKernelGGL(gpe::gemm::k_gemm, grid, dim3(256), 0, 0, a);
    int rc = hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess;
    if (!rc) {
        (void)hipMemcpy(Cr, dcr, (size_t)M * N * 4, hipMemcpyDeviceToHost);
        if (Ci) (void)hipMemcpy(Ci, dci, (size_t)M * N * 4, hipMemcpyDeviceToHost);
    }
    for (float* p : {dar, dai, dbr, dbi, dcr, dci})
        if (p) (void)hipFree(p);
    return rc;
}

}  // extern "C"
