// dtb_mlp.hip -- kernels of the learned surrogate: a bf16 MFMA GEMM with
// FP32 accumulation and fused bias + activation, and the small elementwise
// kernels of MLP training (ReLU backward, bias gradient, MSE gradient, Adam).
//
// GEMM form "NT": C[M,N] = act(A[M,K] * B[N,K]^T + bias[N]); both operands are
// K-contiguous row-major bf16, so a forward layer is (X, W^T) and the two
// backward products use transposed copies made by dtb_mlp_transpose_bf16.
// One workgroup (4 wavefronts) computes a 64x64 tile; each wavefront owns a
// 32x32 sub-tile and issues v_mfma_f32_32x32x16_bf16 over K in steps of 16.
//
// MFMA operand layout (32x32x16 bf16): lane l holds 8 consecutive K values
// k = 8*(l/32) .. +7 of row (l % 32) for A and of column (l % 32) for B; the
// 16 accumulator registers hold C[8*(r/4) + 4*(l/32) + r%4][l % 32].
#include <cmath>

#include "dtb_internal.hpp"

namespace dtb {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ inline uint16_t f2bf(float f)
{
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40); // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// act: 0 identity, 1 ReLU
// A workgroup of four wavefronts (2x2) computes a (64 TM) x (64 TN) tile of C;
// each wavefront TM x TN sub-tiles of 32x32, one v_mfma_f32_32x32x16_bf16 per
// sub-tile and 16-wide K step, so each fragment it loads feeds TN (or TM)
// products.  The fragments of the next K step are loaded before the products
// of the current one.  M and N are multiples of 32 only: a sub-tile past the
// edge loads clamped rows (inside A and B) and is not stored.
template <int TM, int TN>
__global__ __launch_bounds__(256) void gemm_nt_kernel(int32_t M, int32_t N, int32_t K, const uint16_t* __restrict__ A,
                                                      int64_t lda, const uint16_t* __restrict__ B, int64_t ldb,
                                                      const float* __restrict__ bias, int32_t act,
                                                      float* __restrict__ C, int64_t ldc,
                                                      uint16_t* __restrict__ Cb, int64_t ldcb)
{
    const int wid = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int m0 = blockIdx.y * (64 * TM) + (wid >> 1) * (32 * TM);
    const int n0 = blockIdx.x * (64 * TN) + (wid & 1) * (32 * TN);
    if (m0 >= M || n0 >= N) return; // wavefront-uniform
    const int r = lane & 31;
    const int kh = (lane >> 5) * 8;
    const uint16_t* pa[TM];
    const uint16_t* pb[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) pa[i] = A + (int64_t)(min(m0 + 32 * i, M - 32) + r) * lda + kh;
#pragma unroll
    for (int j = 0; j < TN; ++j) pb[j] = B + (int64_t)(min(n0 + 32 * j, N - 32) + r) * ldb + kh;
    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int x = 0; x < 16; ++x) acc[i][j][x] = 0.0f;
    bf16x8 a[TM], b[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) a[i] = *reinterpret_cast<const bf16x8*>(pa[i]);
#pragma unroll
    for (int j = 0; j < TN; ++j) b[j] = *reinterpret_cast<const bf16x8*>(pb[j]);
    for (int k0 = 16; k0 < K; k0 += 16) {
        bf16x8 an[TM], bn[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) an[i] = *reinterpret_cast<const bf16x8*>(pa[i] + k0);
#pragma unroll
        for (int j = 0; j < TN; ++j) bn[j] = *reinterpret_cast<const bf16x8*>(pb[j] + k0);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < TM; ++i) a[i] = an[i];
#pragma unroll
        for (int j = 0; j < TN; ++j) b[j] = bn[j];
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        if (n0 + 32 * j >= N) continue;
        const int col = n0 + 32 * j + r;
        const float bv = bias ? bias[col] : 0.0f;
#pragma unroll
        for (int i = 0; i < TM; ++i) {
            if (m0 + 32 * i >= M) continue;
#pragma unroll
            for (int x = 0; x < 16; ++x) {
                const int row = m0 + 32 * i + 8 * (x >> 2) + 4 * (lane >> 5) + (x & 3);
                float v = acc[i][j][x] + bv;
                if (act == 1) v = v > 0.0f ? v : 0.0f;
                if (C) C[(int64_t)row * ldc + col] = v;
                if (Cb) Cb[(int64_t)row * ldcb + col] = f2bf(v);
            }
        }
    }
}

// Large products: LDS-staged tiles.  A workgroup of eight wavefronts (4 x 2)
// computes a 256x256 tile of C; each wavefront a 64x128 sub-tile (2 x 4
// MFMA tiles of 32x32, eight v_mfma_f32_32x32x16_bf16 per 16-wide K step
// from six LDS fragments).  K advances 64 at a time: the next 256x64 slices
// of A and B are loaded from global memory into registers (16-byte loads, 8
// lanes per 128-byte row) while the current ones are multiplied out of LDS,
// then stored to the other half of a double buffer; one barrier per slice.
// LDS rows are 64 + 8 bf16 long, so the 32 rows a fragment read touches
// fall on distinct banks.  Workgroups are dealt to tiles XCD by XCD
// (workgroup w runs on XCD w mod 8): each XCD takes a contiguous run of
// tiles, whose A and B slices then repeat in its own L2.
// Requires M, N multiples of 256 and K of 64.
constexpr int kLdsBM = 256, kLdsBN = 256, kLdsKT = 64, kLdsRow = kLdsKT + 8;

__global__ __launch_bounds__(512, 1) void gemm_nt_lds_kernel(int32_t M, int32_t N, int32_t K,
                                                             const uint16_t* __restrict__ A, int64_t lda,
                                                             const uint16_t* __restrict__ B, int64_t ldb,
                                                             const float* __restrict__ bias, int32_t act,
                                                             float* __restrict__ C, int64_t ldc,
                                                             uint16_t* __restrict__ Cb, int64_t ldcb)
{
    extern __shared__ __align__(16) uint16_t lds[]; // [2][A 256 x 72 | B 256 x 72]
    constexpr int kStage = (kLdsBM + kLdsBN) * kLdsRow;
    const int t = threadIdx.x;
    const int wid = t >> 6;
    const int lane = t & 63;
    const int wm = wid >> 1, wn = wid & 1;
    // tile of this workgroup: XCD-contiguous runs of tiles, row-major within
    const int tn = N / kLdsBN;
    const int tiles = (M / kLdsBM) * tn;
    const int nx = 8;
    const int w = blockIdx.x;
    const int per = (tiles + nx - 1) / nx;
    const int xcd = w % nx, slot = w / nx;
    const int tile = xcd * per + slot;
    if (tile >= tiles) return; // workgroup-uniform
    const int m_blk = (tile / tn) * kLdsBM, n_blk = (tile % tn) * kLdsBN;
    // global -> register copies: chunk c = t + 512 u (u < 4) of each operand,
    // row c / 8, 8 bf16 at column 8 (c % 8) of the 64-wide slice
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const uint16_t* ga[4];
    const uint16_t* gb[4];
    int so[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int c = t + 512 * u;
        const int row = c >> 3, kc = (c & 7) * 8;
        ga[u] = A + (int64_t)(m_blk + row) * lda + kc;
        gb[u] = B + (int64_t)(n_blk + row) * ldb + kc;
        so[u] = row * kLdsRow + kc;
    }
    u32x4 ra[4], rb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        ra[u] = *reinterpret_cast<const u32x4*>(ga[u]);
        rb[u] = *reinterpret_cast<const u32x4*>(gb[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        *reinterpret_cast<u32x4*>(lds + so[u]) = ra[u];
        *reinterpret_cast<u32x4*>(lds + kLdsBM * kLdsRow + so[u]) = rb[u];
    }
    __syncthreads();
    f32x16 acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int x = 0; x < 16; ++x) acc[i][j][x] = 0.0f;
    const int r = lane & 31;
    const int kh = (lane >> 5) * 8;
    const int nk = K / kLdsKT;
    for (int kt = 0; kt < nk; ++kt) {
        const bool more = kt + 1 < nk;
        if (more) {
            const int64_t k1 = (int64_t)(kt + 1) * kLdsKT;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                ra[u] = *reinterpret_cast<const u32x4*>(ga[u] + k1);
                rb[u] = *reinterpret_cast<const u32x4*>(gb[u] + k1);
            }
        }
        const uint16_t* sa = lds + (kt & 1) * kStage;
        const uint16_t* sb = sa + kLdsBM * kLdsRow;
#pragma unroll
        for (int ks = 0; ks < kLdsKT / 16; ++ks) {
            bf16x8 a[2], b[4];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const bf16x8*>(sa + (wm * 64 + 32 * i + r) * kLdsRow + ks * 16 + kh);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                b[j] = *reinterpret_cast<const bf16x8*>(sb + (wn * 128 + 32 * j + r) * kLdsRow + ks * 16 + kh);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            uint16_t* da = lds + ((kt + 1) & 1) * kStage;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                *reinterpret_cast<u32x4*>(da + so[u]) = ra[u];
                *reinterpret_cast<u32x4*>(da + kLdsBM * kLdsRow + so[u]) = rb[u];
            }
        }
        __syncthreads();
    }
    const int m0 = m_blk + wm * 64, n0 = n_blk + wn * 128;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int col = n0 + 32 * j + r;
        const float bv = bias ? bias[col] : 0.0f;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
            for (int x = 0; x < 16; ++x) {
                const int row = m0 + 32 * i + 8 * (x >> 2) + 4 * (lane >> 5) + (x & 3);
                float v = acc[i][j][x] + bv;
                if (act == 1) v = v > 0.0f ? v : 0.0f;
                if (C) C[(int64_t)row * ldc + col] = v;
                if (Cb) Cb[(int64_t)row * ldcb + col] = f2bf(v);
            }
        }
    }
}

// Small products (too few 64x64 tiles to fill the chip): a workgroup computes
// one 32x32 tile of C and its four wavefronts split K in four consecutive
// ranges; wavefronts 1-3 hand their partial sums to wavefront 0 through LDS,
// which adds them in wavefront order (a fixed order: results do not depend on
// timing).  Each wavefront's serial chain of K steps is a quarter as long.
__global__ __launch_bounds__(256) void gemm_nt_splitk_kernel(int32_t M, int32_t N, int32_t K,
                                                             const uint16_t* __restrict__ A, int64_t lda,
                                                             const uint16_t* __restrict__ B, int64_t ldb,
                                                             const float* __restrict__ bias, int32_t act,
                                                             float* __restrict__ C, int64_t ldc,
                                                             uint16_t* __restrict__ Cb, int64_t ldcb)
{
    __shared__ float red[3][16][64];
    const int wid = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int m0 = blockIdx.y * 32;
    const int n0 = blockIdx.x * 32;
    const int r = lane & 31;
    const int kh = (lane >> 5) * 8;
    const int steps = K / 16;
    const int per = (steps + 3) / 4;
    const int kb = min(wid * per, steps) * 16;
    const int ke = min((wid + 1) * per, steps) * 16;
    const uint16_t* pa = A + (int64_t)(m0 + r) * lda + kh;
    const uint16_t* pb = B + (int64_t)(n0 + r) * ldb + kh;
    f32x16 acc;
#pragma unroll
    for (int x = 0; x < 16; ++x) acc[x] = 0.0f;
    for (int k0 = kb; k0 < ke; k0 += 16) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(pa + k0);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(pb + k0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    if (wid > 0) {
#pragma unroll
        for (int x = 0; x < 16; ++x) red[wid - 1][x][lane] = acc[x];
    }
    __syncthreads();
    if (wid != 0) return;
#pragma unroll
    for (int w = 0; w < 3; ++w)
#pragma unroll
        for (int x = 0; x < 16; ++x) acc[x] += red[w][x][lane];
    const int col = n0 + r;
    const float bv = bias ? bias[col] : 0.0f;
#pragma unroll
    for (int x = 0; x < 16; ++x) {
        const int row = m0 + 8 * (x >> 2) + 4 * (lane >> 5) + (x & 3);
        float v = acc[x] + bv;
        if (act == 1) v = v > 0.0f ? v : 0.0f;
        if (C) C[(int64_t)row * ldc + col] = v;
        if (Cb) Cb[(int64_t)row * ldcb + col] = f2bf(v);
    }
}

// dZ = dY * relu'(Z)  (Z > 0), written in FP32 and bf16
__global__ void relu_backward_kernel(int64_t n, const float* __restrict__ dY, const float* __restrict__ Z,
                                     float* __restrict__ dZ, uint16_t* __restrict__ dZb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = Z[i] > 0.0f ? dY[i] : 0.0f;
    if (dZ) dZ[i] = v;
    if (dZb) dZb[i] = f2bf(v);
}

// db[j] = sum_i dZ[i][j]
__global__ void bias_grad_kernel(int32_t M, int32_t N, const float* __restrict__ dZ, float* __restrict__ db)
{
    const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.0f;
    for (int32_t i = 0; i < M; ++i) s += dZ[(int64_t)i * N + j];
    db[j] = s;
}

// grad = scale * (out - target) masked; loss partial sums in lsum[blockIdx]
__global__ void mse_grad_kernel(int64_t n, const float* __restrict__ out, const float* __restrict__ target,
                                const float* __restrict__ mask, float scale, float* __restrict__ grad,
                                uint16_t* __restrict__ gradb, float* __restrict__ lsum)
{
    __shared__ float red[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float e = 0.0f;
    if (i < n) {
        const float mk = mask ? mask[i] : 1.0f;
        const float d = (out[i] - target[i]) * mk;
        e = d * d;
        const float g = scale * d;
        if (grad) grad[i] = g;
        if (gradb) gradb[i] = f2bf(g);
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && lsum) lsum[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba, 2015) on FP32 master weights, refreshing the bf16 copy
__global__ void adam_kernel(int64_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2,
                            uint16_t* __restrict__ wb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

__global__ void transpose_bf16_kernel(int32_t M, int32_t N, const uint16_t* __restrict__ in, uint16_t* __restrict__ out)
{
    __shared__ uint16_t tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5; // 256 threads: 32 x 8
    for (int y = ty; y < 32; y += 8) {
        const int r = by + y, c = bx + tx;
        if (r < M && c < N) tile[y][tx] = in[(int64_t)r * N + c];
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
        const int r = bx + y, c = by + tx; // out is [N][M]
        if (r < N && c < M) out[(int64_t)r * M + c] = tile[tx][y];
    }
}

__global__ void f32_to_bf16_kernel(int64_t n, const float* __restrict__ in, uint16_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

// out[r][c] = in[r][c] * scale[c] + shift[c] for c < cin, 0 for cin <= c < cout
__global__ void affine_kernel(int32_t rows, int32_t cin, int32_t cout, const float* __restrict__ in, int64_t ldi,
                              const float* __restrict__ scale, const float* __restrict__ shift,
                              float* __restrict__ out, uint16_t* __restrict__ outb, int64_t ldo)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rows * cout) return;
    const int32_t r = (int32_t)(i / cout), c = (int32_t)(i % cout);
    float v = 0.0f;
    if (c < cin) v = in[(int64_t)r * ldi + c] * (scale ? scale[c] : 1.0f) + (shift ? shift[c] : 0.0f);
    if (out) out[(int64_t)r * ldo + c] = v;
    if (outb) outb[(int64_t)r * ldo + c] = f2bf(v);
}

__device__ inline float sigm(float x) { return 1.0f / (1.0f + __expf(-x)); }

// GRU cell (gate order r, z, n as torch.nn.GRU):
//   r = s(gi_r + gh_r), z = s(gi_z + gh_z), n = tanh(gi_n + r * gh_n), h' = (1 - z) n + z h
// gi = x W_ih^T + b_ih and gh = h W_hh^T + b_hh come from the GEMM, [B][3H] each.
__global__ void gru_forward_kernel(int32_t B, int32_t H, const float* __restrict__ gi, const float* __restrict__ gh,
                                   const float* __restrict__ h, float* __restrict__ hn, uint16_t* __restrict__ hnb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)B * H) return;
    const int64_t b = i / H, j = i % H;
    const float* a = gi + b * 3 * H;
    const float* c = gh + b * 3 * H;
    const float r = sigm(a[j] + c[j]);
    const float z = sigm(a[H + j] + c[H + j]);
    const float n = tanhf(a[2 * H + j] + r * c[2 * H + j]);
    const float v = (1.0f - z) * n + z * h[i];
    hn[i] = v;
    if (hnb) hnb[i] = f2bf(v);
}

// Backward of one GRU step with the gates recomputed from gi, gh, h.
// dh = dh_a + dh_b + dh_c (readout gradient, direct and GEMM parts of the
// next step's recurrent gradient; dh_b/dh_c may be NULL).  Writes
// dgi = [dr', dz', dn'] and dgh = [dr', dz', dn' * r] (FP32 and bf16) and the
// direct recurrent part dh_z = dh * z.
__global__ void gru_backward_kernel(int32_t B, int32_t H, const float* __restrict__ gi, const float* __restrict__ gh,
                                    const float* __restrict__ h, const float* __restrict__ dh_a,
                                    const float* __restrict__ dh_b, const float* __restrict__ dh_c,
                                    float* __restrict__ dgi, uint16_t* __restrict__ dgib, float* __restrict__ dgh,
                                    uint16_t* __restrict__ dghb, float* __restrict__ dh_z)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)B * H) return;
    const int64_t b = i / H, j = i % H;
    const int64_t o = b * 3 * H;
    const float r = sigm(gi[o + j] + gh[o + j]);
    const float z = sigm(gi[o + H + j] + gh[o + H + j]);
    const float ghn = gh[o + 2 * H + j];
    const float n = tanhf(gi[o + 2 * H + j] + r * ghn);
    float d = dh_a[i];
    if (dh_b) d += dh_b[i];
    if (dh_c) d += dh_c[i];
    const float dn = d * (1.0f - z) * (1.0f - n * n);
    const float dz = d * (h[i] - n) * z * (1.0f - z);
    const float dr = dn * ghn * r * (1.0f - r);
    dh_z[i] = d * z;
    dgi[o + j] = dr;
    dgi[o + H + j] = dz;
    dgi[o + 2 * H + j] = dn;
    dgh[o + j] = dr;
    dgh[o + H + j] = dz;
    dgh[o + 2 * H + j] = dn * r;
    if (dgib) {
        dgib[o + j] = f2bf(dr);
        dgib[o + H + j] = f2bf(dz);
        dgib[o + 2 * H + j] = f2bf(dn);
    }
    if (dghb) {
        dghb[o + j] = f2bf(dr);
        dghb[o + H + j] = f2bf(dz);
        dghb[o + 2 * H + j] = f2bf(dn * r);
    }
}

static inline dim3 g1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

} // namespace dtb

using namespace dtb;

extern "C" {

int dtb_mlp_gemm_bf16_nt(int32_t M, int32_t N, int32_t K, const void* A, int64_t lda, const void* B, int64_t ldb,
                         const float* bias, int32_t act, float* C, int64_t ldc, void* Cb, int64_t ldcb, void* stream)
{
    DTB_REQUIRE(M > 0 && N > 0 && K > 0, "M, N, K must be positive");
    DTB_REQUIRE(M % 32 == 0 && N % 32 == 0 && K % 16 == 0, "M and N must be multiples of 32 and K of 16");
    DTB_REQUIRE(A && B && (C || Cb), "A, B and an output are required");
    DTB_REQUIRE(lda >= K && ldb >= K && lda % 8 == 0 && ldb % 8 == 0, "lda, ldb must be >= K and multiples of 8");
    DTB_REQUIRE(act == 0 || act == 1, "act must be 0 (identity) or 1 (ReLU)");
    DTB_REQUIRE((reinterpret_cast<uintptr_t>(A) & 15) == 0 && (reinterpret_cast<uintptr_t>(B) & 15) == 0,
                "A and B must be 16-byte aligned");
    // 128x128 workgroup tiles while they still give every compute unit a
    // workgroup, else 64x64, else 32x32 tiles with K split over the wavefronts
    // (the surrogate's own products are that small)
    const int64_t big = (int64_t)((N + 127) / 128) * ((M + 127) / 128);
    const int64_t lds_tiles = (int64_t)(N / kLdsBN) * (M / kLdsBM);
    if (M % kLdsBM == 0 && N % kLdsBN == 0 && K % kLdsKT == 0 && lds_tiles >= 64) {
        // LDS-staged 256x256 tiles, one workgroup per tile (the tile map deals them XCD by XCD)
        const size_t lds = sizeof(uint16_t) * 2 * (size_t)(kLdsBM + kLdsBN) * kLdsRow;
        DTB_TRY_HIP(hipFuncSetAttribute((const void*)gemm_nt_lds_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)lds));
        const unsigned nwg = (unsigned)((lds_tiles + 7) / 8 * 8);
        hipLaunchKernelGGL(gemm_nt_lds_kernel, dim3(nwg), dim3(512), lds, (hipStream_t)stream, M, N, K,
                           (const uint16_t*)A, lda, (const uint16_t*)B, ldb, bias, act, C, ldc, (uint16_t*)Cb, ldcb);
    } else if (big >= 256) {
        dim3 grid((unsigned)((N + 127) / 128), (unsigned)((M + 127) / 128));
        hipLaunchKernelGGL((gemm_nt_kernel<2, 2>), grid, dim3(256), 0, (hipStream_t)stream, M, N, K,
                           (const uint16_t*)A, lda, (const uint16_t*)B, ldb, bias, act, C, ldc, (uint16_t*)Cb, ldcb);
    } else if ((int64_t)((N + 63) / 64) * ((M + 63) / 64) >= 256) {
        dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
        hipLaunchKernelGGL((gemm_nt_kernel<1, 1>), grid, dim3(256), 0, (hipStream_t)stream, M, N, K,
                           (const uint16_t*)A, lda, (const uint16_t*)B, ldb, bias, act, C, ldc, (uint16_t*)Cb, ldcb);
    } else {
        dim3 grid((unsigned)(N / 32), (unsigned)(M / 32));
        hipLaunchKernelGGL(gemm_nt_splitk_kernel, grid, dim3(256), 0, (hipStream_t)stream, M, N, K,
                           (const uint16_t*)A, lda, (const uint16_t*)B, ldb, bias, act, C, ldc, (uint16_t*)Cb, ldcb);
    }
    DTB_TRY_HIP(hipGetLastError());
    return DTB_OK;
}

int dtb_mlp_relu_backward(int64_t n, const float* dY, const float* Z, float* dZ, void* dZb, void* stream)
{
    DTB_REQUIRE(n >= 0 && dY && Z && (dZ || dZb), "bad arguments");
    if (n == 0) return DTB_OK;
    hipLaunchKernelGGL(relu_backward_kernel, g1(n), dim3(256), 0, (hipStream_t)stream, n, dY, Z, dZ, (uint16_t*)dZb);
    DTB_TRY_HIP(hipGetLastError());
    return DTB_OK;
}

int dtb_mlp_bias_grad(int32_t M, int32_t N, const float* dZ, float* db, void* stream)
{
    DTB_REQUIRE(M > 0 && N > 0 && dZ && db, "bad arguments");
    hipLaunchKernelGGL(bias_grad_kernel, g1(N), dim3(256), 0, (hipStream_t)stream, M, N, dZ, db);
    DTB_TRY_HIP(hipGetLastError());
    return DTB_OK;
}

int dtb_mlp_mse_grad(int64_t n, const float* out, const float* target, const float* mask, float scale, float* grad,
                     void* gradb, float* lsum, void* stream)
{
    DTB_REQUIRE(n > 0 && out && target, "bad arguments");
    hipLaunchKernelGGL(mse_grad_kernel, g1(n), dim3(256), 0, (hipStream_t)stream, n, out, target, mask, scale, grad,
                       (uint16_t*)gradb, lsum);
    DTB_TRY_HIP(hipGetLastError());
    return DTB_OK;
}

int dtb_mlp_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                 int32_t t, void* wb, void* stream)
{
    DTB_REQUIRE(n > 0 && w && g && m && v && t >= 1, "bad arguments");
    const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(adam_kernel, g1(n), dim3(256), 0, (hipStream_t)stream, n, w, g, m, v, lr, b1, b2, eps, c1, c2,
                       (uint16_t*)wb);
    DTB_TRY_HIP(hipGetLastError());
    return DTB_OK;
}

int dtb_mlp_transpose_bf16(int32_t M, int32_t N, const void* in, void* out, void* stream)
{
    DTB_REQUIRE(M > 0 && N > 0 && in && out, "bad arguments");
    dim3 grid((unsigned)((N + 31) / 32), (unsigned)((M + 31) / 32));
    hipLaunchKernelGGL(transpose_bf16_kernel, grid, dim3(256), 0, (hipStream_t)stream, M, N, (const uint16_t*)in,
                       (uint16_t*)out);
    DTB_TRY_HIP(hipGetLastError());
    return DTB_OK;
}

int dtb_mlp_f32_to_bf16(int64_t n, const float* in, void* out, void* stream)
{
    DTB_REQUIRE(n >= 0 && in && out, "bad arguments");
    if (n == 0) return DTB_OK;
    hipLaunchKernelGGL(f32_to_bf16_kernel, g1(n), dim3(256), 0, (hipStream_t)stream, n, in, (uint16_t*)out);
    DTB_TRY_HIP(hipGetLastError());
    return DTB_OK;
}

int dtb_mlp_affine(int32_t rows, int32_t cin, int32_t cout, const float* in, int64_t ldi, const float* scale,
                   const float* shift, float* out, void* outb, int64_t ldo, void* stream)
{
    DTB_REQUIRE(rows >= 0 && cin >= 0 && cout >= cin && in && (out || outb), "bad arguments");
    DTB_REQUIRE(ldi >= cin && ldo >= cout, "ldi must be >= cin and ldo >= cout");
    if ((int64_t)rows * cout == 0) return DTB_OK;
    hipLaunchKernelGGL(affine_kernel, g1((int64_t)rows * cout), dim3(256), 0, (hipStream_t)stream, rows, cin, cout,
                       in, ldi, scale, shift, out, (uint16_t*)outb, ldo);
    DTB_TRY_HIP(hipGetLastError());
    return DTB_OK;
}

int dtb_gru_forward(int32_t B, int32_t H, const float* gi, const float* gh, const float* h, float* hn, void* hnb,
                    void* stream)
{
    DTB_REQUIRE(B > 0 && H > 0 && gi && gh && h && hn, "bad arguments");
    hipLaunchKernelGGL(gru_forward_kernel, g1((int64_t)B * H), dim3(256), 0, (hipStream_t)stream, B, H, gi, gh, h, hn,
                       (uint16_t*)hnb);
    DTB_TRY_HIP(hipGetLastError());
    return DTB_OK;
}

int dtb_gru_backward(int32_t B, int32_t H, const float* gi, const float* gh, const float* h, const float* dh_a,
                     const float* dh_b, const float* dh_c, float* dgi, void* dgib, float* dgh, void* dghb, float* dh_z,
                     void* stream)
{
    DTB_REQUIRE(B > 0 && H > 0 && gi && gh && h && dh_a && dgi && dgh && dh_z, "bad arguments");
    hipLaunchKernelGGL(gru_backward_kernel, g1((int64_t)B * H), dim3(256), 0, (hipStream_t)stream, B, H, gi, gh, h,
                       dh_a, dh_b, dh_c, dgi, (uint16_t*)dgib, dgh, (uint16_t*)dghb, dh_z);
    DTB_TRY_HIP(hipGetLastError());
    return DTB_OK;
}

} // extern "C"
