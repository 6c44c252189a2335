// unet.hip - the warm-start network's training and inference kernels (gfx950).
//
// Activations are BF16, NHWC (channels contiguous), with every channel count
// a multiple of 32; weights of a 3x3 convolution are [9][Cout][Cin] (tap =
// 3*(dy+1) + (dx+1)); accumulation is FP32 on the matrix cores
// (v_mfma_f32_16x16x32_bf16); activation gradients are FP32.  Rounding to
// BF16 is round-to-nearest-even, written out (f2bf), so the CPU oracle can
// reproduce every rounding point.
//
// 3x3 convolution = implicit GEMM: M = pixels, N = output channels, K = 9 x
// Cin.  A wave owns 64 pixels (4 MFMA row blocks) x 16*NJ channels, so each
// weight fragment is reused by 4 MFMAs from registers and each input
// fragment by NJ; the K loop walks the 9 taps x 32-channel chunks with 16-byte
// loads (zero outside the image).  The data gradient is the same kernel run
// on the flipped, transposed weights; the weight gradient is a split-K GEMM
// over pixels through LDS-transposed tiles with a fixed-order reduction
// (deterministic, no atomics).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>

#include "glide_unet.h"

namespace glide {
int set_error(int code, const char *msg);
}

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16_t;

int ufail(int code, const std::string &msg) { return glide::set_error(code, msg.c_str()); }
int ucheck(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ufail((int)e, std::string(what) + ": " + hipGetErrorString(e));
    return GLIDE_OK;
}

}  // namespace

__device__ __forceinline__ float bf2f(bf16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
// round to nearest even (NaN stays NaN)
__device__ __forceinline__ bf16_t f2bf(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}

union Frag {
    uint4 raw;
    bf16x8 v;
};

// ---------------------------------------------------------------------------
// 3x3 convolution, implicit GEMM on MFMA.
// MODE 0 (forward): out_bf16[p][co] = act(bias[co] + sum_{t,ci} W[t][co][ci] x[p+off_t][ci])
//        x = concat(xa (CA ch), xb (CB ch)); act = ReLU when relu != 0.
// MODE 1 (data gradient): the same sum without bias/act into FP32 outputs split
//        by channel: co < SA -> ga[p][co] (+)= ..., else gb[p][co - SA] (+)= ...
// Grid: x = pixel blocks of 256 (4 waves x 64), y = output-channel groups of 16*NJ.
// ---------------------------------------------------------------------------
template <int NJ, int MODE>
__global__ __launch_bounds__(256) void k_conv3x3(const bf16_t *__restrict__ xa, int CA, const bf16_t *__restrict__ xb,
                                                 int CB, int N, int H, int W, const bf16_t *__restrict__ Wt, int Cout,
                                                 const float *__restrict__ bias, int relu, bf16_t *__restrict__ out,
                                                 float *__restrict__ ga, int SA, int acc_a, float *__restrict__ gb,
                                                 int acc_b)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int Cin = CA + CB;
    const long P = (long)N * H * W;
    const long pbase = (long)blockIdx.x * 256 + wave * 64;
    const int co0 = blockIdx.y * 16 * NJ;
    const int r16 = lane & 15, kq = lane >> 4;   // fragment row / k-quarter
    // the 4 pixels this lane feeds as A rows
    int py[4], px[4], pn[4];
    bool pv[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
        long p = pbase + 16 * a + r16;
        pv[a] = p < P;
        long q = pv[a] ? p : 0;
        px[a] = (int)(q % W);
        long t = q / W;
        py[a] = (int)(t % H);
        pn[a] = (int)(t / H);
    }
    f32x4 acc[4][NJ];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int j = 0; j < NJ; j++) acc[a][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // K steps s = (tap t, 32-channel chunk c): fragments of step s + 1 are
    // loaded while the MFMAs of step s run (the accumulation order over s is
    // the plain tap-major, channel-minor order)
    const int nch = Cin / 32, nsteps = 9 * nch;
    auto load = [&](int s, Frag(&A)[4], Frag(&B)[NJ]) {
        const int t = s / nch, c0 = (s - t * nch) * 32;
        const int dy = t / 3 - 1, dx = t % 3 - 1;
        const bf16_t *xs = c0 < CA ? xa : xb;
        const int cs = c0 < CA ? CA : CB, cc = (c0 < CA ? c0 : c0 - CA) + 8 * kq;
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const int y = py[a] + dy, x = px[a] + dx;
            const bool ok = pv[a] && y >= 0 && y < H && x >= 0 && x < W;
            const long src = ok ? (((long)pn[a] * H + y) * W + x) : 0;
            A[a].raw = ok ? *(const uint4 *)(xs + src * cs + cc) : uint4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int j = 0; j < NJ; j++)
            B[j].raw = *(const uint4 *)(Wt + ((long)t * Cout + co0 + 16 * j + r16) * Cin + c0 + 8 * kq);
    };
    auto mma = [&](const Frag(&A)[4], const Frag(&B)[NJ]) {
#pragma unroll
        for (int j = 0; j < NJ; j++)
#pragma unroll
            for (int a = 0; a < 4; a++)
                acc[a][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[a].v, B[j].v, acc[a][j], 0, 0, 0);
    };
    Frag Ac[4], Bc[NJ], An[4], Bn[NJ];
    load(0, Ac, Bc);
#pragma unroll 1
    for (int s = 0; s < nsteps; s++) {
        if (s + 1 < nsteps) load(s + 1, An, Bn);
        mma(Ac, Bc);
#pragma unroll
        for (int a = 0; a < 4; a++) Ac[a] = An[a];
#pragma unroll
        for (int j = 0; j < NJ; j++) Bc[j] = Bn[j];
    }
    // D fragment: row (pixel) = 4*kq + i, column (channel) = r16
#pragma unroll
    for (int a = 0; a < 4; a++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const long p = pbase + 16 * a + 4 * kq + i;
            if (p >= P) continue;
#pragma unroll
            for (int j = 0; j < NJ; j++) {
                const int co = co0 + 16 * j + r16;
                float val = acc[a][j][i];
                if (MODE == 0) {
                    val += bias[co];
                    if (relu) val = fmaxf(val, 0.0f);
                    out[p * Cout + co] = f2bf(val);
                } else {
                    if (co < SA) {
                        float *d = ga + p * SA + co;
                        *d = acc_a ? *d + val : val;
                    } else {
                        float *d = gb + p * (Cout - SA) + (co - SA);
                        *d = acc_b ? *d + val : val;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: partial[z][t][co][ci] = sum over the z-th pixel chunk of
// dz[p][co] * x[p + off_t][ci]; one block = 64 co x 64 ci of one tap.
// ---------------------------------------------------------------------------
constexpr int WG_K = 32;          // pixels per K step
constexpr int WG_LD = WG_K + 8;   // LDS row pitch (bf16), breaks bank conflicts

__global__ __launch_bounds__(256) void k_conv3x3_wgrad(const bf16_t *__restrict__ dz, int Cout,
                                                       const bf16_t *__restrict__ xa, int CA,
                                                       const bf16_t *__restrict__ xb, int CB, int N, int H, int W,
                                                       long chunk, float *__restrict__ partial)
{
    __shared__ bf16_t sA[64][WG_LD];   // [co][pixel]
    __shared__ bf16_t sB[64][WG_LD];   // [ci][pixel]
    const int Cin = CA + CB;
    const int nco = (Cout + 63) / 64;
    const int t = blockIdx.y / nco, cog = blockIdx.y % nco;
    const int co0 = cog * 64, ci0 = blockIdx.x * 64;
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    const long P = (long)N * H * W;
    const long p_lo = (long)blockIdx.z * chunk, p_hi = p_lo + chunk < P ? p_lo + chunk : P;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, kq = lane >> 4;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;   // wave's 32 x 32 sub-tile
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    // staging: thread -> (pixel row kk = tid / 8, 8-channel group g = tid % 8)
    const int kk = tid >> 3, g = tid & 7;
    for (long p0 = p_lo; p0 < p_hi; p0 += WG_K) {
        const long p = p0 + kk;
        uint4 va = uint4{0u, 0u, 0u, 0u}, vb = uint4{0u, 0u, 0u, 0u};
        if (p < p_hi) {
            if (co0 + 8 * g < Cout) va = *(const uint4 *)(dz + p * Cout + co0 + 8 * g);
            const int x = (int)(p % W);
            const long q = p / W;
            const int y = (int)(q % H), n = (int)(q / H);
            const int ys = y + dy, xs = x + dx;
            const int c = ci0 + 8 * g;
            if (ys >= 0 && ys < H && xs >= 0 && xs < W && c < Cin) {
                const long s = ((long)n * H + ys) * W + xs;
                vb = c < CA ? *(const uint4 *)(xa + s * CA + c) : *(const uint4 *)(xb + s * CB + (c - CA));
            }
        }
        __syncthreads();
        const bf16_t *ea = (const bf16_t *)&va, *eb = (const bf16_t *)&vb;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            sA[8 * g + e][kk] = ea[e];
            sB[8 * g + e][kk] = eb[e];
        }
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 2; a++) {
            Frag A;
            A.raw = *(const uint4 *)&sA[wm + 16 * a + r16][8 * kq];
#pragma unroll
            for (int b = 0; b < 2; b++) {
                Frag B;
                B.raw = *(const uint4 *)&sB[wn + 16 * b + r16][8 * kq];
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.v, B.v, acc[a][b], 0, 0, 0);
            }
        }
    }
    float *dst = partial + ((long)blockIdx.z * 9 + t) * Cout * Cin;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int co = co0 + wm + 16 * a + 4 * kq + i, ci = ci0 + wn + 16 * b + r16;
                if (co < Cout && ci < Cin) dst[(long)co * Cin + ci] = acc[a][b][i];
            }
}

// out[k] = sum_z partial[z][k] in z order
__global__ void k_reduce_chunks(const float *__restrict__ partial, long n, int nz, float *__restrict__ out)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float s = 0.0f;
    for (int z = 0; z < nz; z++) s += partial[(long)z * n + k];
    out[k] = s;
}

// Plain GEMM C[m][n] = sum_k A[m][k] B[n][k] (both K-contiguous, BF16; M, N
// multiples of 16, K of 32): one wave per 16 x 16 tile.  Exposes the MFMA
// fragment layout for testing.
__global__ __launch_bounds__(64) void k_gemm_bf16(const bf16_t *__restrict__ A, const bf16_t *__restrict__ B,
                                                  float *__restrict__ C, int M, int N, int K)
{
    const int lane = threadIdx.x, r16 = lane & 15, kq = lane >> 4;
    const int m0 = blockIdx.y * 16, n0 = blockIdx.x * 16;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) {
        Frag a, b;
        a.raw = *(const uint4 *)(A + (long)(m0 + r16) * K + k0 + 8 * kq);
        b.raw = *(const uint4 *)(B + (long)(n0 + r16) * K + k0 + 8 * kq);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) C[(long)(m0 + 4 * kq + i) * N + n0 + r16] = acc[i];
}

// ---------------------------------------------------------------------------
// element-wise and small kernels
// ---------------------------------------------------------------------------
// [9][Cout][Cin] FP32 -> BF16 copy, and the data-gradient weights
// Wf[8 - t][ci][co] = W[t][co][ci]
__global__ void k_prep_weights(const float *__restrict__ Wm, int Cout, int Cin, bf16_t *__restrict__ Wb,
                               bf16_t *__restrict__ Wf)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = 9L * Cout * Cin;
    if (k >= n) return;
    const int ci = (int)(k % Cin);
    const long r = k / Cin;
    const int co = (int)(r % Cout), t = (int)(r / Cout);
    bf16_t h = f2bf(Wm[k]);
    Wb[k] = h;
    if (Wf) Wf[((long)(8 - t) * Cin + ci) * Cout + co] = h;
}

// dz = bf16(dy * (y > 0))  (ReLU backward from the post-activation output)
__global__ void k_relu_bwd(const float *__restrict__ dy, const bf16_t *__restrict__ y, long n, bf16_t *__restrict__ dz)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dz[k] = f2bf(bf2f(y[k]) > 0.0f ? dy[k] : 0.0f);
}

// 2x2 average pooling, NHWC
__global__ void k_pool2(const bf16_t *__restrict__ x, int N, int H, int W, int C, bf16_t *__restrict__ y)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int Ho = H / 2, Wo = W / 2;
    if (k >= (long)N * Ho * Wo * C) return;
    const int c = (int)(k % C);
    long r = k / C;
    const int xo = (int)(r % Wo);
    r /= Wo;
    const int yo = (int)(r % Ho), n = (int)(r / Ho);
    const long b = (((long)n * H + 2 * yo) * W + 2 * xo) * C + c;
    float s = ((bf2f(x[b]) + bf2f(x[b + C])) + bf2f(x[b + (long)W * C])) + bf2f(x[b + (long)W * C + C]);
    y[k] = f2bf(0.25f * s);
}
__global__ void k_pool2_bwd(const float *__restrict__ dy, int N, int H, int W, int C, float *__restrict__ dx, int acc)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)N * H * W * C) return;
    const int c = (int)(k % C);
    long r = k / C;
    const int x = (int)(r % W);
    r /= W;
    const int y = (int)(r % H), n = (int)(r / H);
    const float g = 0.25f * dy[(((long)n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
    dx[k] = acc ? dx[k] + g : g;
}
// 8 channels in one 16-byte access
union Bf8 {
    uint4 raw;
    bf16_t h[8];
};

// 2x2 average pooling over 8-channel groups (C % 8 == 0, 16-byte aligned
// tensors): per channel the same sums as k_pool2
__global__ void k_pool2_v8(const bf16_t *__restrict__ x, int N, int H, int W, int C, bf16_t *__restrict__ y)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    if (k >= (long)N * Ho * Wo * C8) return;
    const int c8 = (int)(k % C8);
    long r = k / C8;
    const int xo = (int)(r % Wo);
    r /= Wo;
    const int yo = (int)(r % Ho), n = (int)(r / Ho);
    const long b = (((long)n * H + 2 * yo) * W + 2 * xo) * C + 8 * c8;
    Bf8 q0, q1, q2, q3, o;
    q0.raw = *(const uint4 *)(x + b);
    q1.raw = *(const uint4 *)(x + b + C);
    q2.raw = *(const uint4 *)(x + b + (long)W * C);
    q3.raw = *(const uint4 *)(x + b + (long)W * C + C);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const float s = ((bf2f(q0.h[i]) + bf2f(q1.h[i])) + bf2f(q2.h[i])) + bf2f(q3.h[i]);
        o.h[i] = f2bf(0.25f * s);
    }
    *(uint4 *)(y + 8 * k) = o.raw;
}
// nearest x2 upsampling over 8-channel groups (C % 8 == 0, aligned)
__global__ void k_up2_v8(const bf16_t *__restrict__ x, int N, int H, int W, int C, bf16_t *__restrict__ y)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int C8 = C / 8;
    if (k >= (long)N * 4 * H * W * C8) return;
    const int c8 = (int)(k % C8);
    long r = k / C8;
    const int xo = (int)(r % (2 * W));
    r /= 2 * W;
    const int yo = (int)(r % (2 * H)), n = (int)(r / (2 * H));
    *(uint4 *)(y + 8 * k) = *(const uint4 *)(x + (((long)n * H + yo / 2) * W + xo / 2) * C + 8 * c8);
}
// nearest x2 upsampling, NHWC (input N x H x W)
__global__ void k_up2(const bf16_t *__restrict__ x, int N, int H, int W, int C, bf16_t *__restrict__ y)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)N * 4 * H * W * C) return;
    const int c = (int)(k % C);
    long r = k / C;
    const int xo = (int)(r % (2 * W));
    r /= 2 * W;
    const int yo = (int)(r % (2 * H)), n = (int)(r / (2 * H));
    y[k] = x[(((long)n * H + yo / 2) * W + xo / 2) * C + c];
}
__global__ void k_up2_bwd(const float *__restrict__ dy, int N, int H, int W, int C, float *__restrict__ dx)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)N * H * W * C) return;
    const int c = (int)(k % C);
    long r = k / C;
    const int x = (int)(r % W);
    r /= W;
    const int y = (int)(r % H), n = (int)(r / H);
    const long W2 = 2L * W;
    const long b = (((long)n * 2 * H + 2 * y) * W2 + 2 * x) * C + c;
    dx[k] = ((dy[b] + dy[b + C]) + dy[b + W2 * C]) + dy[b + W2 * C + C];
}

// bias gradient partials: part[z][c] = sum over pixel chunk z of dz[p][c]
__global__ void k_bias_grad(const bf16_t *__restrict__ dz, long P, int C, long chunk, float *__restrict__ part)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const long lo = (long)blockIdx.y * chunk, hi = lo + chunk < P ? lo + chunk : P;
    float s = 0.0f;
    for (long p = lo; p < hi; p++) s += bf2f(dz[p * C + c]);
    part[(long)blockIdx.y * C + c] = s;
}

// 1x1 output head with NO <= 4 outputs:
// out[p][o] = b[o] + sum_c Wh[o][c] x[p][c]  (FP32)
__global__ void k_head(const bf16_t *__restrict__ x, long P, int C, int NO, const float *__restrict__ Wh,
                       const float *__restrict__ bh, float *__restrict__ out)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    for (int o = 0; o < NO; o++) s[o] = bh[o];
    for (int c = 0; c < C; c++) {
        float v = bf2f(x[p * C + c]);
        for (int o = 0; o < NO; o++) s[o] += Wh[o * C + c] * v;
    }
    for (int o = 0; o < NO; o++) out[NO * p + o] = s[o];
}
// k_head reading 8 channels per 16-byte access (C % 8 == 0, aligned): the
// same sums in the same order
__global__ void k_head_v8(const bf16_t *__restrict__ x, long P, int C, int NO, const float *__restrict__ Wh,
                          const float *__restrict__ bh, float *__restrict__ out)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    for (int o = 0; o < NO; o++) s[o] = bh[o];
    for (int c = 0; c < C; c += 8) {
        Bf8 q;
        q.raw = *(const uint4 *)(x + p * C + c);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float v = bf2f(q.h[i]);
            for (int o = 0; o < NO; o++) s[o] += Wh[o * C + c + i] * v;
        }
    }
    for (int o = 0; o < NO; o++) out[NO * p + o] = s[o];
}
// dx[p][c] = sum_o dout[p][o] Wh[o][c]
__global__ void k_head_bwd_x(const float *__restrict__ dout, long P, int C, int NO, const float *__restrict__ Wh,
                             float *__restrict__ dx)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P * C) return;
    const long p = k / C;
    const int c = (int)(k % C);
    float s = dout[NO * p] * Wh[c];
    for (int o = 1; o < NO; o++) s += dout[NO * p + o] * Wh[o * C + c];
    dx[k] = s;
}
// partials over pixel chunks of dWh[o][c] = sum_p dout[p][o] x[p][c] and dbh[o] = sum_p dout[p][o]
__global__ void k_head_bwd_w(const float *__restrict__ dout, const bf16_t *__restrict__ x, long P, int C, int NO,
                             long chunk, float *__restrict__ part)
{
    const int k = threadIdx.x;   // 0 .. NO*C + NO - 1
    const int nk = NO * C + NO;
    if (k >= nk) return;
    const long lo = (long)blockIdx.x * chunk, hi = lo + chunk < P ? lo + chunk : P;
    float s = 0.0f;
    if (k < NO * C) {
        const int o = k / C, c = k % C;
        for (long p = lo; p < hi; p++) s += dout[NO * p + o] * bf2f(x[p * C + c]);
    } else {
        const int o = k - NO * C;
        for (long p = lo; p < hi; p++) s += dout[NO * p + o];
    }
    part[(long)blockIdx.x * nk + k] = s;
}

// masked mean-squared error: loss partials and d(loss)/d(pred) = 2 m (pred - target) / count
__global__ void k_mse(const float *__restrict__ pred, const float *__restrict__ target, const float *__restrict__ mask,
                      long n, float inv_count, float *__restrict__ grad, float *__restrict__ part)
{
    __shared__ float s[256];
    long k = (long)blockIdx.x * 256 + threadIdx.x;
    float e = 0.0f;
    if (k < n) {
        const float m = mask ? mask[k / 2] : 1.0f;
        const float d = (pred[k] - target[k]) * m;
        grad[k] = 2.0f * d * inv_count;
        e = d * d;
    }
    s[threadIdx.x] = e;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Adam (FP32 master weights); bias corrections c1 = 1 - b1^t, c2 = 1 - b2^t
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, long n, float lr, float b1, float b2, float eps, float c1, float c2,
                       float wd)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k] + wd * w[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    w[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// ---------------------------------------------------------------------------
// solver <-> network
// ---------------------------------------------------------------------------
// Network input of one member, NHWC with 32 channels (7 used):
//   0: H / 1000, 1: tanh((bed + 0.917 H) / 200) (flotation margin: the way
//   the bed enters the momentum balance besides the surface), 2: 100 |grad s|,
//   3: log(beta) + 2, 4: B / 24 - 1, 5, 6: 100 ds/dx, 100 ds/dy (central
//   differences, one-sided at the edges); channels 7..31 are 0.  Every
//   channel is invariant to a shift of the domain, so a network trained on
//   small domains applies to large ones.
// s = max(bed, -0.917 H) + H (the reference's surface, glide/physics.py:223-229).
__device__ __forceinline__ double surf(const double *H, const double *bed, long k)
{
    return fmax(bed[k], -0.917 * H[k]) + H[k];
}
__global__ void k_features(const double *__restrict__ H, const double *__restrict__ bed, const double *__restrict__ beta,
                           const double *__restrict__ B, int ny, int nx, double dx, bf16_t *__restrict__ x)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)ny * nx) return;
    const int i = (int)(k / nx), j = (int)(k % nx);
    const int jl = j > 0 ? j - 1 : j, jr = j < nx - 1 ? j + 1 : j;
    const int it = i > 0 ? i - 1 : i, ib = i < ny - 1 ? i + 1 : i;
    const double sx = (surf(H, bed, (long)i * nx + jr) - surf(H, bed, (long)i * nx + jl)) / ((jr - jl) * dx);
    const double sy = (surf(H, bed, (long)ib * nx + j) - surf(H, bed, (long)it * nx + j)) / ((ib - it) * dx);
    float f[7] = {(float)(H[k] / 1000.0), (float)tanh((bed[k] + 0.917 * H[k]) / 200.0),
                  (float)(100.0 * sqrt(sx * sx + sy * sy)), (float)(log(beta[k]) + 2.0), (float)(B[k] / 24.0 - 1.0),
                  (float)(100.0 * sx), (float)(100.0 * sy)};
    bf16_t *o = x + k * 32;
    for (int c = 0; c < 32; c++) o[c] = c < 7 ? f2bf(f[c]) : (bf16_t)0;
}
// targets (FP32, 3 per cell): cell-centred (u_l + u_r)/2 and (v_t + v_b)/2
// times scale, and the thickness change (H - H_prev) times hscale
__global__ void k_targets(const double *__restrict__ u, const double *__restrict__ v, const double *__restrict__ H,
                          const double *__restrict__ H_prev, int ny, int nx, double scale, double hscale,
                          float *__restrict__ tgt)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)ny * nx) return;
    const int i = (int)(k / nx), j = (int)(k % nx);
    tgt[3 * k] = (float)(0.5 * (u[(long)i * (nx + 1) + j] + u[(long)i * (nx + 1) + j + 1]) * scale);
    tgt[3 * k + 1] = (float)(0.5 * (v[(long)i * nx + j] + v[(long)(i + 1) * nx + j]) * scale);
    tgt[3 * k + 2] = (float)((H[k] - H_prev[k]) * hscale);
}
// the solver state from a prediction (3 per cell): faces from the cell-centred
// velocities / scale (interior faces average their two cells, boundary faces
// take their one cell); H = max(H_prev + dH / hscale, gamma)
__global__ void k_apply_prediction(const float *__restrict__ pred, int ny, int nx, double scale, double hscale,
                                   double *__restrict__ u, double *__restrict__ v, double *__restrict__ H,
                                   const double *__restrict__ H_prev, const double *__restrict__ gamma)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nu = (long)ny * (nx + 1), nv = (long)(ny + 1) * nx, nc = (long)ny * nx;
    if (k < nu) {
        const int i = (int)(k / (nx + 1)), j = (int)(k % (nx + 1));
        const int jl = j > 0 ? j - 1 : 0, jr = j < nx ? j : nx - 1;
        u[k] = 0.5 * ((double)pred[3 * ((long)i * nx + jl)] + (double)pred[3 * ((long)i * nx + jr)]) / scale;
    } else if (k < nu + nv) {
        const long kk = k - nu;
        const int i = (int)(kk / nx), j = (int)(kk % nx);
        const int it = i > 0 ? i - 1 : 0, ib = i < ny ? i : ny - 1;
        v[kk] = 0.5 * ((double)pred[3 * ((long)it * nx + j) + 1] + (double)pred[3 * ((long)ib * nx + j) + 1]) / scale;
    } else if (k < nu + nv + nc) {
        const long kk = k - nu - nv;
        H[kk] = fmax(H_prev[kk] + (double)pred[3 * kk + 2] / hscale, gamma[kk]);
    }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {
inline unsigned nblk(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }
inline hipStream_t SS(void *s) { return (hipStream_t)s; }
bool bad_ch(int c) { return c < 0 || c % 32 != 0 || c > 4096; }
// 8-channel (16-byte) accesses apply: C % 8 == 0 and both tensors 16-byte aligned
bool vec8(int C, const void *a, const void *b) { return C % 8 == 0 && (((uintptr_t)a | (uintptr_t)b) & 15) == 0; }
// Output channels per wave (16 * NJ) of a 3x3 convolution over P pixels: the
// widest of 64, 32 and 16 that still gives at least 2048 waves (two per
// SIMD), else 16.  Small levels are bound by each wave's serial K loop, so
// they take more, narrower waves; large ones keep the input fragments'
// reuse across output channels.  The sums are the same for every width.
int conv_nj(long P, int C)
{
    const long pw = (long)nblk(P) * 4;   // 64-pixel waves
    for (int nj = 4; nj > 1; nj /= 2)
        if (C % (16 * nj) == 0 && pw * (C / (16 * nj)) >= 2048) return nj;
    return 1;
}
}  // namespace

extern "C" {

int glide_unet_conv3x3(const uint16_t *xa, int CA, const uint16_t *xb, int CB, int N, int H, int W,
                       const uint16_t *weights, int Cout, const float *bias, int relu, uint16_t *out, void *stream)
{
    if (!xa || !weights || !bias || !out || (CB > 0 && !xb)) return ufail(GLIDE_E_NULL, "conv3x3: NULL argument");
    if (bad_ch(CA) || CA == 0 || bad_ch(CB) || bad_ch(Cout) || Cout == 0 || N < 1 || H < 1 || W < 1)
        return ufail(GLIDE_E_SIZE, "conv3x3: channels must be positive multiples of 32, sizes positive");
    const long P = (long)N * H * W;
    const int nj = conv_nj(P, Cout);
    if (nj == 4)
        hipLaunchKernelGGL((k_conv3x3<4, 0>), dim3(nblk(P), Cout / 64), dim3(256), 0, SS(stream), xa, CA, xb, CB, N, H,
                           W, weights, Cout, bias, relu, out, nullptr, 0, 0, nullptr, 0);
    else if (nj == 2)
        hipLaunchKernelGGL((k_conv3x3<2, 0>), dim3(nblk(P), Cout / 32), dim3(256), 0, SS(stream), xa, CA, xb, CB, N, H,
                           W, weights, Cout, bias, relu, out, nullptr, 0, 0, nullptr, 0);
    else
        hipLaunchKernelGGL((k_conv3x3<1, 0>), dim3(nblk(P), Cout / 16), dim3(256), 0, SS(stream), xa, CA, xb, CB, N, H,
                           W, weights, Cout, bias, relu, out, nullptr, 0, 0, nullptr, 0);
    return ucheck("conv3x3");
}

int glide_unet_conv3x3_dgrad(const uint16_t *dz, int Cout, int N, int H, int W, const uint16_t *wflip, int CA,
                             float *ga, int acc_a, int CB, float *gb, int acc_b, void *stream)
{
    if (!dz || !wflip || !ga || (CB > 0 && !gb)) return ufail(GLIDE_E_NULL, "conv3x3_dgrad: NULL argument");
    if (bad_ch(Cout) || Cout == 0 || bad_ch(CA) || CA == 0 || bad_ch(CB) || N < 1 || H < 1 || W < 1)
        return ufail(GLIDE_E_SIZE, "conv3x3_dgrad: channels must be positive multiples of 32, sizes positive");
    const long P = (long)N * H * W;
    const int Cin = CA + CB;   // output channels of the transposed convolution
    const int nj = conv_nj(P, Cin);
    if (nj == 4)
        hipLaunchKernelGGL((k_conv3x3<4, 1>), dim3(nblk(P), Cin / 64), dim3(256), 0, SS(stream), dz, Cout, nullptr, 0,
                           N, H, W, wflip, Cin, nullptr, 0, nullptr, ga, CA, acc_a, gb, acc_b);
    else if (nj == 2)
        hipLaunchKernelGGL((k_conv3x3<2, 1>), dim3(nblk(P), Cin / 32), dim3(256), 0, SS(stream), dz, Cout, nullptr, 0,
                           N, H, W, wflip, Cin, nullptr, 0, nullptr, ga, CA, acc_a, gb, acc_b);
    else
        hipLaunchKernelGGL((k_conv3x3<1, 1>), dim3(nblk(P), Cin / 16), dim3(256), 0, SS(stream), dz, Cout, nullptr, 0,
                           N, H, W, wflip, Cin, nullptr, 0, nullptr, ga, CA, acc_a, gb, acc_b);
    return ucheck("conv3x3_dgrad");
}

long glide_unet_wgrad_chunks(int N, int H, int W) { return ((long)N * H * W + 4095) / 4096; }

int glide_unet_conv3x3_wgrad(const uint16_t *dz, int Cout, const uint16_t *xa, int CA, const uint16_t *xb, int CB,
                             int N, int H, int W, float *partial, float *dw, void *stream)
{
    if (!dz || !xa || !partial || !dw || (CB > 0 && !xb)) return ufail(GLIDE_E_NULL, "conv3x3_wgrad: NULL argument");
    if (bad_ch(Cout) || Cout == 0 || bad_ch(CA) || CA == 0 || bad_ch(CB) || N < 1 || H < 1 || W < 1)
        return ufail(GLIDE_E_SIZE, "conv3x3_wgrad: channels must be positive multiples of 32, sizes positive");
    const long P = (long)N * H * W, chunk = 4096;
    const long nz = glide_unet_wgrad_chunks(N, H, W);
    const int Cin = CA + CB;
    dim3 grid((Cin + 63) / 64, 9 * ((Cout + 63) / 64), (unsigned)nz);
    hipLaunchKernelGGL(k_conv3x3_wgrad, grid, dim3(256), 0, SS(stream), dz, Cout, xa, CA, xb, CB, N, H, W, chunk,
                       partial);
    const long n = 9L * Cout * Cin;
    hipLaunchKernelGGL(k_reduce_chunks, dim3(nblk(n)), dim3(256), 0, SS(stream), partial, n, (int)nz, dw);
    (void)P;
    return ucheck("conv3x3_wgrad");
}

int glide_unet_gemm_bf16(const uint16_t *A, const uint16_t *B, float *C, int M, int N, int K, void *stream)
{
    if (!A || !B || !C) return ufail(GLIDE_E_NULL, "gemm: NULL argument");
    if (M < 16 || N < 16 || K < 32 || M % 16 || N % 16 || K % 32)
        return ufail(GLIDE_E_SIZE, "gemm: M, N multiples of 16 and K a multiple of 32");
    hipLaunchKernelGGL(k_gemm_bf16, dim3(N / 16, M / 16), dim3(64), 0, SS(stream), A, B, C, M, N, K);
    return ucheck("gemm");
}

int glide_unet_prep_weights(const float *wm, int Cout, int Cin, uint16_t *wb, uint16_t *wflip, void *stream)
{
    if (!wm || !wb) return ufail(GLIDE_E_NULL, "prep_weights: NULL argument");
    if (Cout < 1 || Cin < 1) return ufail(GLIDE_E_SIZE, "prep_weights: bad shape");
    const long n = 9L * Cout * Cin;
    hipLaunchKernelGGL(k_prep_weights, dim3(nblk(n)), dim3(256), 0, SS(stream), wm, Cout, Cin, wb, wflip);
    return ucheck("prep_weights");
}

int glide_unet_relu_bwd(const float *dy, const uint16_t *y, long n, uint16_t *dz, void *stream)
{
    if (!dy || !y || !dz) return ufail(GLIDE_E_NULL, "relu_bwd: NULL argument");
    if (n < 0) return ufail(GLIDE_E_SIZE, "relu_bwd: negative size");
    if (n) hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(n)), dim3(256), 0, SS(stream), dy, y, n, dz);
    return ucheck("relu_bwd");
}

int glide_unet_pool2(const uint16_t *x, int N, int H, int W, int C, uint16_t *y, void *stream)
{
    if (!x || !y) return ufail(GLIDE_E_NULL, "pool2: NULL argument");
    if (N < 1 || H < 2 || W < 2 || H % 2 || W % 2 || C < 1) return ufail(GLIDE_E_SIZE, "pool2: even sizes required");
    const long n = (long)N * (H / 2) * (W / 2) * C;
    if (vec8(C, x, y))
        hipLaunchKernelGGL(k_pool2_v8, dim3(nblk(n / 8)), dim3(256), 0, SS(stream), x, N, H, W, C, y);
    else
        hipLaunchKernelGGL(k_pool2, dim3(nblk(n)), dim3(256), 0, SS(stream), x, N, H, W, C, y);
    return ucheck("pool2");
}

int glide_unet_pool2_bwd(const float *dy, int N, int H, int W, int C, float *dx, int accumulate, void *stream)
{
    if (!dy || !dx) return ufail(GLIDE_E_NULL, "pool2_bwd: NULL argument");
    if (N < 1 || H < 2 || W < 2 || H % 2 || W % 2 || C < 1) return ufail(GLIDE_E_SIZE, "pool2_bwd: even sizes required");
    const long n = (long)N * H * W * C;
    hipLaunchKernelGGL(k_pool2_bwd, dim3(nblk(n)), dim3(256), 0, SS(stream), dy, N, H, W, C, dx, accumulate);
    return ucheck("pool2_bwd");
}

int glide_unet_up2(const uint16_t *x, int N, int H, int W, int C, uint16_t *y, void *stream)
{
    if (!x || !y) return ufail(GLIDE_E_NULL, "up2: NULL argument");
    if (N < 1 || H < 1 || W < 1 || C < 1) return ufail(GLIDE_E_SIZE, "up2: bad size");
    const long n = (long)N * 4 * H * W * C;
    if (vec8(C, x, y))
        hipLaunchKernelGGL(k_up2_v8, dim3(nblk(n / 8)), dim3(256), 0, SS(stream), x, N, H, W, C, y);
    else
        hipLaunchKernelGGL(k_up2, dim3(nblk(n)), dim3(256), 0, SS(stream), x, N, H, W, C, y);
    return ucheck("up2");
}

int glide_unet_up2_bwd(const float *dy, int N, int H, int W, int C, float *dx, void *stream)
{
    if (!dy || !dx) return ufail(GLIDE_E_NULL, "up2_bwd: NULL argument");
    if (N < 1 || H < 1 || W < 1 || C < 1) return ufail(GLIDE_E_SIZE, "up2_bwd: bad size");
    const long n = (long)N * H * W * C;
    hipLaunchKernelGGL(k_up2_bwd, dim3(nblk(n)), dim3(256), 0, SS(stream), dy, N, H, W, C, dx);
    return ucheck("up2_bwd");
}

long glide_unet_bias_chunks(long P) { return (P + 4095) / 4096; }

int glide_unet_bias_grad(const uint16_t *dz, long P, int C, float *partial, float *db, void *stream)
{
    if (!dz || !partial || !db) return ufail(GLIDE_E_NULL, "bias_grad: NULL argument");
    if (P < 1 || C < 1) return ufail(GLIDE_E_SIZE, "bias_grad: bad size");
    const long nz = glide_unet_bias_chunks(P);
    hipLaunchKernelGGL(k_bias_grad, dim3((C + 63) / 64, (unsigned)nz), dim3(64), 0, SS(stream), dz, P, C, 4096L,
                       partial);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(nblk(C)), dim3(256), 0, SS(stream), partial, (long)C, (int)nz, db);
    return ucheck("bias_grad");
}

int glide_unet_head(const uint16_t *x, long P, int C, int NO, const float *wh, const float *bh, float *out,
                    void *stream)
{
    if (!x || !wh || !bh || !out) return ufail(GLIDE_E_NULL, "head: NULL argument");
    if (P < 1 || C < 1 || NO < 1 || NO > 4) return ufail(GLIDE_E_SIZE, "head: bad size (1 <= outputs <= 4)");
    if (vec8(C, x, x))
        hipLaunchKernelGGL(k_head_v8, dim3(nblk(P)), dim3(256), 0, SS(stream), x, P, C, NO, wh, bh, out);
    else
        hipLaunchKernelGGL(k_head, dim3(nblk(P)), dim3(256), 0, SS(stream), x, P, C, NO, wh, bh, out);
    return ucheck("head");
}

int glide_unet_head_bwd(const float *dout, const uint16_t *x, long P, int C, int NO, const float *wh, float *dx,
                        float *partial, float *dwh_dbh, void *stream)
{
    if (!dout || !x || !wh || !dx || !partial || !dwh_dbh) return ufail(GLIDE_E_NULL, "head_bwd: NULL argument");
    const int nk = NO * C + NO;
    if (P < 1 || C < 1 || NO < 1 || NO > 4 || nk > 1024) return ufail(GLIDE_E_SIZE, "head_bwd: bad size");
    hipLaunchKernelGGL(k_head_bwd_x, dim3(nblk(P * C)), dim3(256), 0, SS(stream), dout, P, C, NO, wh, dx);
    const long nz = glide_unet_bias_chunks(P);
    hipLaunchKernelGGL(k_head_bwd_w, dim3((unsigned)nz), dim3(((nk + 63) / 64) * 64), 0, SS(stream), dout, x, P, C,
                       NO, 4096L, partial);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(nblk(nk)), dim3(256), 0, SS(stream), partial, (long)nk, (int)nz, dwh_dbh);
    return ucheck("head_bwd");
}

long glide_unet_mse_blocks(long n) { return (n + 255) / 256; }

int glide_unet_mse(const float *pred, const float *target, const float *mask, long n, float inv_count, float *grad,
                   float *partial, void *stream)
{
    if (!pred || !target || !grad || !partial) return ufail(GLIDE_E_NULL, "mse: NULL argument");
    if (n < 1) return ufail(GLIDE_E_SIZE, "mse: empty");
    hipLaunchKernelGGL(k_mse, dim3((unsigned)glide_unet_mse_blocks(n)), dim3(256), 0, SS(stream), pred, target, mask, n,
                       inv_count, grad, partial);
    return ucheck("mse");
}

int glide_unet_adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1, float b2, float eps,
                    int step, float weight_decay, void *stream)
{
    if (!w || !g || !m || !v) return ufail(GLIDE_E_NULL, "adam: NULL argument");
    if (n < 1 || step < 1) return ufail(GLIDE_E_SIZE, "adam: bad size or step");
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(256), 0, SS(stream), w, g, m, v, n, lr, b1, b2, eps, c1, c2,
                       weight_decay);
    return ucheck("adam");
}

int glide_unet_features(const double *H, const double *bed, const double *beta, const double *B, int ny, int nx,
                        double dx, uint16_t *x, void *stream)
{
    if (!H || !bed || !beta || !B || !x) return ufail(GLIDE_E_NULL, "features: NULL argument");
    if (ny < 2 || nx < 2 || !(dx > 0.0)) return ufail(GLIDE_E_SIZE, "features: bad grid");
    const long n = (long)ny * nx;
    hipLaunchKernelGGL(k_features, dim3(nblk(n)), dim3(256), 0, SS(stream), H, bed, beta, B, ny, nx, dx, x);
    return ucheck("features");
}

int glide_unet_targets(const double *u, const double *v, const double *H, const double *H_prev, int ny, int nx,
                       double scale, double hscale, float *tgt, void *stream)
{
    if (!u || !v || !H || !H_prev || !tgt) return ufail(GLIDE_E_NULL, "targets: NULL argument");
    if (ny < 1 || nx < 1) return ufail(GLIDE_E_SIZE, "targets: bad grid");
    hipLaunchKernelGGL(k_targets, dim3(nblk((long)ny * nx)), dim3(256), 0, SS(stream), u, v, H, H_prev, ny, nx, scale,
                       hscale, tgt);
    return ucheck("targets");
}

int glide_unet_apply_prediction(const float *pred, int ny, int nx, double scale, double hscale, double *u, double *v,
                                double *H, const double *H_prev, const double *gamma, void *stream)
{
    if (!pred || !u || !v || !H || !H_prev || !gamma) return ufail(GLIDE_E_NULL, "apply_prediction: NULL argument");
    if (ny < 1 || nx < 1 || !(scale > 0.0) || !(hscale > 0.0))
        return ufail(GLIDE_E_SIZE, "apply_prediction: bad grid or scale");
    const long n = (long)ny * (nx + 1) + (long)(ny + 1) * nx + (long)ny * nx;
    hipLaunchKernelGGL(k_apply_prediction, dim3(nblk(n)), dim3(256), 0, SS(stream), pred, ny, nx, scale, hscale, u, v,
                       H, H_prev, gamma);
    return ucheck("apply_prediction");
}

}  // extern "C"
