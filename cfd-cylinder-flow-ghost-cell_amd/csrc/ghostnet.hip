// Warm-start network for the capacitance (Schur) system of the ghost-cell pressure solve.
//
// A periodic 1D convolution network over the boundary angle: `layers` convolutions of kernel 5
// with per-layer dilation, channels cin -> ch -> ... -> ch -> 1 and ReLU between layers.  Given a
// right-hand side c of S y = c (angle order, scaled to unit norm) and per-row geometry features,
// it predicts y / |c|.  The reference solves S directly (pressure_schur.py:92-111); this network
// has no counterpart there.
//
// Every convolution (forward, backward-data, backward-weights) runs on the matrix cores as
// v_mfma_f32_16x16x16_bf16 with BF16 operands and FP32 accumulation.  Lane layout (checked on
// gfx950): lane l holds A[l%16][4*(l/16)+e], B[4*(l/16)+e][l%16] and D[4*(l/16)+e][l%16], e = 0..3.
// Master weights, gradients and the Adam moments are FP32.  Activations are stored channel-last,
// position p = sample * L + t, with channel counts padded to a multiple of 4 (padding is zero).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "ghostflow.h"

namespace {

typedef unsigned short bf16_t;
typedef __attribute__((__vector_size__(4 * sizeof(short)))) short s4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4;

constexpr int KS = 5;          // kernel taps
constexpr int NET_MAXL = 8;
constexpr int PB = 64;         // positions per block

// round to nearest, ties to even (finite inputs)
__host__ __device__ inline bf16_t f2bf_bits(unsigned u) { return (bf16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16); }
__device__ inline bf16_t f2bf(float x) { return f2bf_bits(__float_as_uint(x)); }
__device__ inline float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }
inline int up4(int n) { return (n + 3) / 4 * 4; }
inline int up16(int n) { return (n + 15) / 16 * 16; }

// ------------------------------------------------------------------------------------------
// Convolution as an implicit GEMM: out[o][p] = sum_{k,i} W[o][k*cin+i] * in[p + (k-2) dil][i]
// (periodic within each sample).  256 threads = 4 waves x 16 positions, all output rows.
// mode 0: z = acc + bias, act = bf16(relu(z)); mode 1 (adjoint through the previous ReLU):
// dz = bf16(acc * (zprev > 0)).
// ------------------------------------------------------------------------------------------
struct ConvArgs {
  const bf16_t* w; int kp, mrows;      // weights [mrows][kp], mrows in {16, 32, 48, 64}
  const bf16_t* in; int cin;           // input [P][cin], cin a multiple of 4
  int dil, L, P;
  const float* bias; int cout, cstride;
  float* z; bf16_t* act;               // mode 0 outputs ([P][cstride]); act may be null
  const float* zprev; bf16_t* dz;      // mode 1
};

// The weights [mrows][kp] (at most 64 x 320 BF16, 40 KB) are staged in LDS once per block with 16-byte
// loads, rows padded by 8 elements so that the 64 lanes' 8-byte fragment reads fall in distinct bank
// pairs; every wave then reads its A fragments there instead of from L2 (which cost 40 KB per wave).
constexpr int CONV_PAD = 8;
inline size_t conv_lds_bytes(int mrows, int kp) { return (size_t)mrows * (kp + CONV_PAD) * sizeof(bf16_t); }

template <int MODE, int MT>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
  extern __shared__ bf16_t wl[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int p = blockIdx.x * PB + w * 16 + l16;
  const int ldw = a.kp + CONV_PAD;
  int n = 0, t0 = 0;
  if (p < a.P) { n = p / a.L; t0 = p - n * a.L; }
  // Every global read of the block goes out before the first wait: the B fragments of the whole K
  // loop (kp <= 320 = 5 taps x 64 channels, so at most 20), the epilogue's operands (the bias, or
  // in the adjoint pass the previous layer's pre-activation, 4 consecutive rows per lane), and the
  // block's share of the weight staging (at most 64 x 320 / 8 / 256 = 10 16-byte chunks per thread).
  // Issued one after another as in a plain loop, each would cost its own memory round trip.
  constexpr int KMAX = 20, STAGE_MAX = 10;
  s4 bv[KMAX];
  #pragma unroll
  for (int t = 0; t < KMAX; ++t) {
    const int kk = 16 * t + 4 * kq;
    const int tap = kk / a.cin, ci = kk - tap * a.cin;
    s4 b = {0, 0, 0, 0};
    if (16 * t < a.kp && p < a.P && tap < KS) {
      int tt = (t0 + (tap - 2) * a.dil) % a.L;
      if (tt < 0) tt += a.L;
      b = *reinterpret_cast<const s4*>(a.in + ((size_t)n * a.L + tt) * a.cin + ci);
    }
    bv[t] = b;
  }
  // output rows o = mt*16 + 4 kq + e (e < 4); the row stride is a multiple of 4, so a lane's 4 rows
  // are all inside it or all outside, and they are 16-byte (FP32) / 8-byte (BF16) aligned
  f4 ep[4];
  #pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const int o = mt * 16 + 4 * kq;
    f4 v = {0.f, 0.f, 0.f, 0.f};
    if (mt < MT && p < a.P && o < a.cstride) {
      if (MODE == 0) {
        #pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = o + e < a.cout ? a.bias[o + e] : 0.f;
      } else {
        v = *reinterpret_cast<const f4*>(a.zprev + (size_t)p * a.cstride + o);
      }
    }
    ep[mt] = v;
  }
  {
    // kp is a multiple of 16, so every row is a whole number of 8-element (16-byte) chunks
    const int cpr = a.kp / 8, nchunk = a.mrows * cpr;
    // Indexes beyond the last chunk are clamped to it: those threads read and write the last
    // chunk again (the same bytes to the same place), which keeps both loops free of branches.
    uint4 st[STAGE_MAX];
    int off[STAGE_MAX];
    #pragma unroll
    for (int r = 0; r < STAGE_MAX; ++r) {
      const int q = min((int)threadIdx.x + 256 * r, nchunk - 1);
      const int row = q / cpr, c8 = q - row * cpr;
      st[r] = *reinterpret_cast<const uint4*>(a.w + (size_t)row * a.kp + 8 * c8);
      off[r] = row * ldw + 8 * c8;
    }
    #pragma unroll
    for (int r = 0; r < STAGE_MAX; ++r) *reinterpret_cast<uint4*>(wl + off[r]) = st[r];
  }
  __syncthreads();
  f4 acc[MT];
  #pragma unroll
  for (int t = 0; t < MT; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
  // MT row tiles (mrows / 16, a template argument) on independent accumulators; the next K step's
  // A fragments are read from LDS before this step's MFMAs issue.  (A loop over a run-time tile
  // count kept one accumulator and waited for each fragment read before its MFMA.)
  s4 av[MT];
  #pragma unroll
  for (int mt = 0; mt < MT; ++mt) av[mt] = *reinterpret_cast<const s4*>(wl + (size_t)(mt * 16 + l16) * ldw + 4 * kq);
  #pragma clang loop unroll(full)
  for (int t = 0; t < KMAX; ++t) {
    if (16 * t < a.kp) {
      s4 an[MT];
      // the step after the last reads row 0's first fragment again (in bounds, unused)
      const int kn = 16 * (t + 1) < a.kp ? 16 * (t + 1) : 0;
      #pragma unroll
      for (int mt = 0; mt < MT; ++mt)
        an[mt] = *reinterpret_cast<const s4*>(wl + (size_t)(mt * 16 + l16) * ldw + kn + 4 * kq);
      #pragma unroll
      for (int mt = 0; mt < MT; ++mt)
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[mt], bv[t], acc[mt], 0, 0, 0);
      #pragma unroll
      for (int mt = 0; mt < MT; ++mt) av[mt] = an[mt];
    }
  }
  if (p >= a.P) return;
  #pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int o = mt * 16 + 4 * kq;
    if (o >= a.cstride) continue;
    const size_t q = (size_t)p * a.cstride + o;
    unsigned short h[4];
    if (MODE == 0) {
      f4 zv;
      #pragma unroll
      for (int e = 0; e < 4; ++e) {
        zv[e] = o + e < a.cout ? acc[mt][e] + ep[mt][e] : 0.f;
        h[e] = f2bf(zv[e] > 0.f ? zv[e] : 0.f);
      }
      *reinterpret_cast<f4*>(a.z + q) = zv;
      if (a.act) *reinterpret_cast<s4*>(a.act + q) = s4{(short)h[0], (short)h[1], (short)h[2], (short)h[3]};
    } else {
      #pragma unroll
      for (int e = 0; e < 4; ++e) h[e] = f2bf(o + e < a.cout && ep[mt][e] > 0.f ? acc[mt][e] : 0.f);
      *reinterpret_cast<s4*>(a.dz + q) = s4{(short)h[0], (short)h[1], (short)h[2], (short)h[3]};
    }
  }
}

template <int MODE>
void launch_conv(const ConvArgs& a, int nblk, hipStream_t st) {
  const size_t lds = conv_lds_bytes(a.mrows, a.kp);
  switch (a.mrows / 16) {
    case 1: hipLaunchKernelGGL((k_conv<MODE, 1>), dim3(nblk), dim3(256), lds, st, a); break;
    case 2: hipLaunchKernelGGL((k_conv<MODE, 2>), dim3(nblk), dim3(256), lds, st, a); break;
    case 3: hipLaunchKernelGGL((k_conv<MODE, 3>), dim3(nblk), dim3(256), lds, st, a); break;
    default: hipLaunchKernelGGL((k_conv<MODE, 4>), dim3(nblk), dim3(256), lds, st, a); break;
  }
}

// ------------------------------------------------------------------------------------------
// Weight gradient: dW[o][k*cin+i] = sum_p dz[p][o] in[p + (k-2) dil][i]; one block per 64
// positions writes its partial sums (weights in [o][i][k] order, then the bias) for a fixed-order
// reduction.
// ------------------------------------------------------------------------------------------
struct ConvWArgs {
  const bf16_t* dz; int cop, cout;     // [P][cop]
  const bf16_t* in; int cin, cinr;     // [P][cin] (cinr real channels)
  int dil, L, P;
  int mrows, kp;                       // GEMM shape: rows o (padded to 16), columns k*cin+i
  float* part; int stride;             // [nblk][stride], stride = cout*cinr*KS + cout
};

// Both operands of this GEMM have the positions as their K dimension, which the [P][channels]
// layouts hold down a column: read straight from global memory every fragment was four 2-byte
// gathers.  The block stages its 64 positions in LDS transposed instead, dz as dzT[o][p] and the
// shifted input as inT[k*cin+i][p]: a thread takes one 4-channel chunk c (lanes along the channels,
// so each row is read by 16 consecutive lanes) of 4 consecutive positions, issues all of its loads
// before the first write, and writes each channel's 4 positions as one 8-byte store.  With rows
// CW_PITCH = 68 elements (34 dwords) apart, the 16 lanes of a sweep (4 chunks x 4 position quads)
// cover the 32 banks once, and an MFMA fragment (4 consecutive positions of one row) is one 8-byte
// read.  Requires cin <= 64 and mrows <= 64 (16 chunks), as the network's shapes are.
constexpr int CW_PITCH = PB + 4;
inline size_t convw_lds_bytes(int mrows, int kp) { return (size_t)(mrows + kp) * CW_PITCH * sizeof(bf16_t); }

__global__ __launch_bounds__(256) void k_convw(ConvWArgs a) {
  extern __shared__ bf16_t cw[];
  bf16_t* dzT = cw;                                   // [mrows][CW_PITCH]
  bf16_t* inT = cw + (size_t)a.mrows * CW_PITCH;      // [kp][CW_PITCH]
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int pbase = blockIdx.x * PB;
  const int c = threadIdx.x & 15, quad = threadIdx.x >> 4;     // chunk, positions 4 quad .. 4 quad + 3
  const int c4i = a.cin / 4;
  int nL[4], t[4];
  bool ok[4];
  {
    const int p0 = pbase + 4 * quad;
    int n = p0 / a.L, tt = p0 - n * a.L;
    #pragma unroll
    for (int i = 0; i < 4; ++i) {
      ok[i] = p0 + i < a.P;
      nL[i] = n * a.L; t[i] = tt;
      if (++tt == a.L) { tt = 0; ++n; }
    }
  }
  s4 vin[KS][4], vz[4];
  #pragma unroll
  for (int k = 0; k < KS; ++k) {
    const int off = (k - 2) * a.dil;
    #pragma unroll
    for (int i = 0; i < 4; ++i) {
      s4 x = {0, 0, 0, 0};
      if (ok[i] && c < c4i) {
        int tt = t[i] + off;
        while (tt < 0) tt += a.L;
        while (tt >= a.L) tt -= a.L;
        x = *reinterpret_cast<const s4*>(a.in + (size_t)(nL[i] + tt) * a.cin + 4 * c);
      }
      vin[k][i] = x;
    }
  }
  #pragma unroll
  for (int i = 0; i < 4; ++i) {
    s4 x = {0, 0, 0, 0};
    if (ok[i] && 4 * c < a.cop) x = *reinterpret_cast<const s4*>(a.dz + (size_t)(pbase + 4 * quad + i) * a.cop + 4 * c);
    vz[i] = x;
  }
  #pragma unroll
  for (int k = 0; k < KS; ++k)
    if (c < c4i) {
      #pragma unroll
      for (int j = 0; j < 4; ++j)
        *reinterpret_cast<s4*>(inT + (size_t)(k * a.cin + 4 * c + j) * CW_PITCH + 4 * quad) =
            s4{vin[k][0][j], vin[k][1][j], vin[k][2][j], vin[k][3][j]};
    }
  if (4 * c < a.mrows) {
    #pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bool live = 4 * c + j < a.cout;
      *reinterpret_cast<s4*>(dzT + (size_t)(4 * c + j) * CW_PITCH + 4 * quad) =
          live ? s4{vz[0][j], vz[1][j], vz[2][j], vz[3][j]} : s4{0, 0, 0, 0};
    }
  }
  // the K padding rows k*cin+i >= KS*cin (fewer than 16) are zero
  for (int q = threadIdx.x; q < (a.kp - KS * a.cin) * (PB / 4); q += 256) {
    const int r = KS * a.cin + q / (PB / 4), p4 = q % (PB / 4);
    *reinterpret_cast<s4*>(inT + (size_t)r * CW_PITCH + 4 * p4) = s4{0, 0, 0, 0};
  }
  __syncthreads();
  const int mt_n = a.mrows / 16, nt_n = a.kp / 16;
  float* part = a.part + (size_t)blockIdx.x * a.stride;
  for (int tile = w; tile < mt_n * nt_n; tile += 4) {
    const int mt = tile / nt_n, nt = tile - mt * nt_n;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    #pragma unroll
    for (int s = 0; s < PB; s += 16) {
      const s4 av = *reinterpret_cast<const s4*>(dzT + (size_t)(mt * 16 + l16) * CW_PITCH + s + 4 * kq);
      const s4 bv = *reinterpret_cast<const s4*>(inT + (size_t)(nt * 16 + l16) * CW_PITCH + s + 4 * kq);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    // D[4*kq+e][l16]: row o = mt*16 + 4*kq + e, column kk = nt*16 + l16
    const int kk = nt * 16 + l16, tap = kk / a.cin, ci = kk - tap * a.cin;
    #pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int oo = mt * 16 + 4 * kq + e;
      if (oo < a.cout && tap < KS && ci < a.cinr) part[((size_t)oo * a.cinr + ci) * KS + tap] = acc[e];
    }
  }
  if ((int)threadIdx.x < a.cout) {
    float sum = 0.f;
    #pragma unroll
    for (int p4 = 0; p4 < PB / 4; ++p4) {
      const s4 v = *reinterpret_cast<const s4*>(dzT + (size_t)threadIdx.x * CW_PITCH + 4 * p4);
      sum += (bf2f(v[0]) + bf2f(v[1])) + (bf2f(v[2]) + bf2f(v[3]));
    }
    part[(size_t)a.cout * a.cinr * KS + threadIdx.x] = sum;
  }
}

// Column sums of part[nblk][stride]: a block takes 64 columns, its 4 waves a quarter of the rows
// each (rows y, y + 4, ...), 16 loads in flight per thread; the quarters then add in a fixed order.
// One thread per column reading its rows in turn waited a memory round trip per row (60 us for 256
// rows).
constexpr int RED_COLS = 64, RED_SLICES = 4, RED_BATCH = 16;
__global__ __launch_bounds__(RED_COLS * RED_SLICES) void k_reduce(const float* part, int nblk, int stride, float* out) {
  __shared__ float red[RED_SLICES][RED_COLS];
  const int x = threadIdx.x % RED_COLS, y = threadIdx.x / RED_COLS;
  const int i = blockIdx.x * RED_COLS + x;
  float s = 0.f;
  if (i < stride) {
    for (int b0 = y; b0 < nblk; b0 += RED_SLICES * RED_BATCH) {
      float v[RED_BATCH];
      #pragma unroll
      for (int r = 0; r < RED_BATCH; ++r) {
        const int b = b0 + RED_SLICES * r;
        v[r] = b < nblk ? part[(size_t)b * stride + i] : 0.f;
      }
      #pragma unroll
      for (int r = 0; r < RED_BATCH; ++r) s += v[r];
    }
  }
  red[y][x] = s;
  __syncthreads();
  if (y == 0 && i < stride) out[i] = (red[0][x] + red[1][x]) + (red[2][x] + red[3][x]);
}

// loss = mean_p (z[p][0] - y[p])^2 ; dz[p][0] = bf16(2 (z - y) / P), other channels 0
__global__ __launch_bounds__(256) void k_loss(const float* z, int cs, const float* y, int P, bf16_t* dz, float* lossp) {
  __shared__ float red[256];
  const int p = blockIdx.x * 256 + threadIdx.x;
  float d2 = 0.f;
  if (p < P) {
    const float d = z[(size_t)p * cs] - y[p];
    d2 = d * d;
    dz[(size_t)p * cs] = f2bf(2.f * d / (float)P);
    for (int c = 1; c < cs; ++c) dz[(size_t)p * cs + c] = 0;
  }
  red[threadIdx.x] = d2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) lossp[blockIdx.x] = red[0];
}

__global__ void k_loss_total(const float* lossp, int nb, int P, float* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float s = 0.f;
    for (int b = 0; b < nb; ++b) s += lossp[b];
    *out = s / (float)P;
  }
}

__global__ void k_adam(float* p, const float* g, float* m1, float* m2, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float m = b1 * m1[i] + (1.f - b1) * gi;
  const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = m; m2[i] = v;
  p[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// BF16 copies of one layer's weights W[o][i][k] (at w): forward [mrows][kp] with column k*cip+i,
// and the adjoint [mrows_b][kpb] with column k*cop+o holding W[o][i][KS-1-k]
__global__ void k_pack(const float* w, int co, int ci, int cip, int cop, bf16_t* wf, int mrows, int kp, bf16_t* wb,
                       int mrows_b, int kpb) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < mrows * kp) {
    const int o = idx / kp, kk = idx - o * kp, k = kk / cip, i = kk - k * cip;
    const float v = (o < co && k < KS && i < ci) ? w[((size_t)o * ci + i) * KS + k] : 0.f;
    wf[idx] = f2bf(v);
  }
  if (wb && idx < mrows_b * kpb) {
    const int i = idx / kpb, kk = idx - i * kpb, k = kk / cop, o = kk - k * cop;
    const float v = (i < ci && k < KS && o < co) ? w[((size_t)o * ci + i) * KS + (KS - 1 - k)] : 0.f;
    wb[idx] = f2bf(v);
  }
}

// batch assembly: position p = b*L + t of sample idx[b]: act0 = bf16(c, features), target = y
__global__ void k_gather(const float* c, const float* y, const float* feat, int nf, int L, const int* idx, int B,
                         bf16_t* act0, int cs, float* target) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= B * L) return;
  const int b = p / L, t = p - b * L;
  const size_t s = (size_t)idx[b] * L + t;
  act0[(size_t)p * cs] = f2bf(c[s]);
  for (int f = 0; f < nf; ++f) act0[(size_t)p * cs + 1 + f] = f2bf(feat[(size_t)t * nf + f]);
  for (int f = 1 + nf; f < cs; ++f) act0[(size_t)p * cs + f] = 0;
  target[p] = y ? y[s] : c[s];
}

// Residual loss through a fixed matrix: one wave per (row i, sample b),
// out[b][i] = sum_t M[i][t] X[b][t] - sub[b][i] (sub optional); with lossp the block's sum of out^2,
// with dz the error bf16(scale * out) on channel 0 (other channels 0).
__global__ __launch_bounds__(256) void k_rowdot(const float* M, int L, const float* X, int xs, const float* sub,
                                                float* out, float* lossp, bf16_t* dz, int cs, float scale) {
  __shared__ float red[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * 4 + w, b = blockIdx.y;
  float v = 0.f;
  if (i < L) {
    const float* Mi = M + (size_t)i * L;
    const float* Xb = X + (size_t)b * L * xs;
    float s = 0.f;
    for (int t = lane; t < L; t += 64) s += Mi[t] * Xb[(size_t)t * xs];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (sub) s -= sub[(size_t)b * L + i];
    v = s;
    if (lane == 0) {
      if (out) out[(size_t)b * L + i] = s;
      if (dz) {
        const size_t q = ((size_t)b * L + i) * cs;
        dz[q] = f2bf(scale * s);
        for (int c = 1; c < cs; ++c) dz[q + c] = 0;
      }
    }
  }
  if (lossp) {
    if (lane == 0) red[w] = v * v;
    __syncthreads();
    if (threadIdx.x == 0) lossp[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
  }
}

// caller inputs x [P][cin] (FP32) -> bf16 [P][cs]
__global__ void k_tobf(const float* x, int cin, int P, bf16_t* act0, int cs) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  for (int f = 0; f < cs; ++f) act0[(size_t)p * cs + f] = f < cin ? f2bf(x[(size_t)p * cin + f]) : 0;
}

__global__ void k_take(const float* z, int cs, int P, float* out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < P) out[p] = z[(size_t)p * cs];
}

}  // namespace

struct GroupData {
  int n = 0, L = 0;
  float *c = nullptr, *y = nullptr, *feat = nullptr;
  float *S = nullptr, *ST = nullptr;   // residual loss |S yhat - c|^2 when set (L x L, angle order)
};

struct mcf_net {
  int layers = 0, cin = 0, ch = 0;
  int dil[NET_MAXL]{}, ci[NET_MAXL]{}, co[NET_MAXL]{}, cip[NET_MAXL]{}, cop[NET_MAXL]{};
  int kp[NET_MAXL]{}, mrows[NET_MAXL]{}, kpb[NET_MAXL]{}, mrows_b[NET_MAXL]{}, woff[NET_MAXL]{};
  int nparams = 0, adam_t = 0;
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  bf16_t* wf[NET_MAXL]{};
  bf16_t* wb[NET_MAXL]{};
  int cap = 0;                       // positions the workspace holds
  bf16_t* act[NET_MAXL + 1]{};
  float* z[NET_MAXL]{};
  bf16_t* dz[NET_MAXL]{};
  float *target = nullptr, *xin = nullptr, *part = nullptr, *lossp = nullptr, *loss = nullptr;
  float *resid = nullptr, *Sg = nullptr, *STg = nullptr;   // residual-loss scratch, matrices of loss_grad
  int Sg_L = 0;
  size_t part_cap = 0;
  int* idx = nullptr; int idx_cap = 0;
  hipStream_t stream = nullptr;
  std::vector<GroupData> groups;
  std::string err;
};

namespace {

int nfail(mcf_net* n, int code, const std::string& msg) {
  if (n) n->err = msg;
  return code;
}

#define NCHK(n, call)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return nfail(n, 4, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int layer_params(const mcf_net* n, int l) { return n->co[l] * n->ci[l] * KS + n->co[l]; }

int ensure_cap(mcf_net* n, int P) {
  if (P <= n->cap) return 0;
  const int cap = std::max(P, 2 * n->cap);
  for (int l = 0; l <= n->layers; ++l) { (void)hipFree(n->act[l]); n->act[l] = nullptr; }
  for (int l = 0; l < n->layers; ++l) {
    (void)hipFree(n->z[l]); (void)hipFree(n->dz[l]); n->z[l] = nullptr; n->dz[l] = nullptr;
  }
  (void)hipFree(n->target); (void)hipFree(n->xin); (void)hipFree(n->part); (void)hipFree(n->lossp);
  (void)hipFree(n->resid);
  n->cap = 0;
  NCHK(n, hipMalloc(&n->act[0], (size_t)cap * n->cip[0] * sizeof(bf16_t)));
  for (int l = 0; l < n->layers; ++l) {
    NCHK(n, hipMalloc(&n->z[l], (size_t)cap * n->cop[l] * sizeof(float)));
    NCHK(n, hipMalloc(&n->dz[l], (size_t)cap * n->cop[l] * sizeof(bf16_t)));
    NCHK(n, hipMalloc(&n->act[l + 1], (size_t)cap * n->cop[l] * sizeof(bf16_t)));
  }
  NCHK(n, hipMalloc(&n->target, (size_t)cap * sizeof(float)));
  NCHK(n, hipMalloc(&n->xin, (size_t)cap * n->cin * sizeof(float)));
  size_t stride = 0;
  for (int l = 0; l < n->layers; ++l) stride = std::max(stride, (size_t)layer_params(n, l));
  n->part_cap = (size_t)((cap + PB - 1) / PB) * stride;
  NCHK(n, hipMalloc(&n->part, n->part_cap * sizeof(float)));
  // k_loss writes ceil(P/256) partials, k_rowdot ceil(L/4) * B <= P/4 + B <= P/4 + P
  NCHK(n, hipMalloc(&n->lossp, ((size_t)cap / 4 + cap + 256) * sizeof(float)));
  NCHK(n, hipMalloc(&n->resid, (size_t)cap * sizeof(float)));
  n->cap = cap;
  return 0;
}

int pack_all(mcf_net* n) {
  for (int l = 0; l < n->layers; ++l) {
    const int tot = std::max(n->mrows[l] * n->kp[l], n->mrows_b[l] * n->kpb[l]);
    hipLaunchKernelGGL(k_pack, dim3((tot + 255) / 256), dim3(256), 0, n->stream, n->p + n->woff[l], n->co[l], n->ci[l],
                       n->cip[l], n->cop[l], n->wf[l], n->mrows[l], n->kp[l], n->wb[l], n->mrows_b[l], n->kpb[l]);
  }
  NCHK(n, hipGetLastError());
  return 0;
}

// forward over P positions already in act[0]
void enqueue_forward(mcf_net* n, int L, int P) {
  const dim3 grid((P + PB - 1) / PB);
  for (int l = 0; l < n->layers; ++l) {
    const bool last = l == n->layers - 1;
    ConvArgs a{n->wf[l], n->kp[l], n->mrows[l], n->act[l], n->cip[l], n->dil[l], L, P,
               n->p + n->woff[l] + n->co[l] * n->ci[l] * KS, n->co[l], n->cop[l], n->z[l],
               last ? nullptr : n->act[l + 1], nullptr, nullptr};
    launch_conv<0>(a, (int)grid.x, n->stream);
  }
}

// loss against target[P] and the full backward pass into g; loss -> *n->loss.  Without S the loss is
// mean (z - target)^2; with S (L x L, and its transpose ST) it is mean over positions of
// (S z - target)^2 per sample, whose error is 2 S^T (S z - target) / P.
void enqueue_backward(mcf_net* n, int L, int P, const float* S = nullptr, const float* ST = nullptr) {
  const int last = n->layers - 1;
  if (S) {
    const int B = P / L;
    const dim3 g((L + 3) / 4, B);
    hipLaunchKernelGGL(k_rowdot, g, dim3(256), 0, n->stream, S, L, n->z[last], n->cop[last], n->target, n->resid,
                       n->lossp, nullptr, n->cop[last], 0.f);
    hipLaunchKernelGGL(k_loss_total, dim3(1), dim3(64), 0, n->stream, n->lossp, (int)(g.x * g.y), P, n->loss);
    hipLaunchKernelGGL(k_rowdot, g, dim3(256), 0, n->stream, ST, L, n->resid, 1, nullptr, nullptr, nullptr,
                       n->dz[last], n->cop[last], 2.f / (float)P);
  } else {
    const int nb = (P + 255) / 256;
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, n->stream, n->z[last], n->cop[last], n->target, P, n->dz[last],
                       n->lossp);
    hipLaunchKernelGGL(k_loss_total, dim3(1), dim3(64), 0, n->stream, n->lossp, nb, P, n->loss);
  }
  const int nblk = (P + PB - 1) / PB;
  for (int l = last; l >= 0; --l) {
    const int stride = layer_params(n, l);
    ConvWArgs w{n->dz[l], n->cop[l], n->co[l], n->act[l], n->cip[l], n->ci[l], n->dil[l], L, P,
                up16(n->co[l]), up16(KS * n->cip[l]), n->part, stride};
    hipLaunchKernelGGL(k_convw, dim3(nblk), dim3(256), convw_lds_bytes(w.mrows, w.kp), n->stream, w);
    hipLaunchKernelGGL(k_reduce, dim3((stride + RED_COLS - 1) / RED_COLS), dim3(RED_COLS * RED_SLICES), 0, n->stream,
                       n->part, nblk, stride, n->g + n->woff[l]);
    if (l > 0) {
      ConvArgs a{n->wb[l], n->kpb[l], n->mrows_b[l], n->dz[l], n->cop[l], n->dil[l], L, P,
                 nullptr, n->ci[l], n->cip[l], nullptr, nullptr, n->z[l - 1], n->dz[l - 1]};
      launch_conv<1>(a, nblk, n->stream);
    }
  }
}

int upload_inputs(mcf_net* n, int nsamp, int L, const float* x, const float* y) {
  const int P = nsamp * L;
  if (int rc = ensure_cap(n, P)) return rc;
  NCHK(n, hipMemcpy(n->xin, x, (size_t)P * n->cin * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_tobf, dim3((P + 255) / 256), dim3(256), 0, n->stream, n->xin, n->cin, P, n->act[0], n->cip[0]);
  if (y) NCHK(n, hipMemcpy(n->target, y, (size_t)P * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

}  // namespace

extern "C" {

int mcf_net_create(int layers, int cin, int channels, const int* dilations, unsigned long long seed, mcf_net** out) {
  if (!out) return 1;
  *out = nullptr;
  if (layers < 1 || layers > NET_MAXL || cin < 1 || cin > 16 || channels < 1 || channels > 64 || !dilations) return 1;
  for (int l = 0; l < layers; ++l) if (dilations[l] < 1 || dilations[l] > 64) return 1;
  mcf_net* n = new mcf_net();
  n->layers = layers; n->cin = cin; n->ch = channels;
  int off = 0;
  for (int l = 0; l < layers; ++l) {
    n->dil[l] = dilations[l];
    n->ci[l] = l == 0 ? cin : channels;
    n->co[l] = l == layers - 1 ? 1 : channels;
    n->cip[l] = up4(n->ci[l]);
    n->cop[l] = up4(n->co[l]);
    n->kp[l] = up16(KS * n->cip[l]);
    n->mrows[l] = up16(n->co[l]);
    n->kpb[l] = up16(KS * n->cop[l]);
    n->mrows_b[l] = up16(n->ci[l]);
    n->woff[l] = off;
    off += layer_params(n, l);
  }
  n->nparams = off;
  if (hipStreamCreate(&n->stream) != hipSuccess) { delete n; return 4; }
  *out = n;
  NCHK(n, hipMalloc(&n->p, off * sizeof(float)));
  NCHK(n, hipMalloc(&n->g, off * sizeof(float)));
  NCHK(n, hipMalloc(&n->m1, off * sizeof(float)));
  NCHK(n, hipMalloc(&n->m2, off * sizeof(float)));
  NCHK(n, hipMalloc(&n->loss, sizeof(float)));
  for (int l = 0; l < layers; ++l) {
    NCHK(n, hipMalloc(&n->wf[l], (size_t)n->mrows[l] * n->kp[l] * sizeof(bf16_t)));
    NCHK(n, hipMalloc(&n->wb[l], (size_t)n->mrows_b[l] * n->kpb[l] * sizeof(bf16_t)));
  }
  // He-normal initialisation (std sqrt(2 / fan_in); the output layer sqrt(1 / fan_in)), zero biases
  std::mt19937_64 rng(seed);
  std::vector<float> hp(off, 0.f);
  for (int l = 0; l < layers; ++l) {
    const double fan = (double)n->ci[l] * KS;
    std::normal_distribution<double> nd(0.0, std::sqrt((l == layers - 1 ? 1.0 : 2.0) / fan));
    for (int i = 0; i < n->co[l] * n->ci[l] * KS; ++i) hp[n->woff[l] + i] = (float)nd(rng);
  }
  return mcf_net_set_params(n, hp.data());
}

void mcf_net_destroy(mcf_net* n) {
  if (!n) return;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  (void)hipFree(n->p); (void)hipFree(n->g); (void)hipFree(n->m1); (void)hipFree(n->m2); (void)hipFree(n->loss);
  for (int l = 0; l < NET_MAXL; ++l) { (void)hipFree(n->wf[l]); (void)hipFree(n->wb[l]); (void)hipFree(n->z[l]); (void)hipFree(n->dz[l]); }
  for (int l = 0; l <= NET_MAXL; ++l) (void)hipFree(n->act[l]);
  (void)hipFree(n->target); (void)hipFree(n->xin); (void)hipFree(n->part); (void)hipFree(n->lossp); (void)hipFree(n->idx);
  (void)hipFree(n->resid); (void)hipFree(n->Sg); (void)hipFree(n->STg);
  mcf_net_clear_groups(n);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

const char* mcf_net_last_error(const mcf_net* n) { return n ? n->err.c_str() : "null network"; }

int mcf_net_num_params(const mcf_net* n) { return n ? n->nparams : 0; }

int mcf_net_get_params(mcf_net* n, float* out) {
  if (!n || !out) return 1;
  NCHK(n, hipStreamSynchronize(n->stream));
  NCHK(n, hipMemcpy(out, n->p, n->nparams * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int mcf_net_set_params(mcf_net* n, const float* in) {
  if (!n || !in) return 1;
  NCHK(n, hipStreamSynchronize(n->stream));
  NCHK(n, hipMemcpy(n->p, in, n->nparams * sizeof(float), hipMemcpyHostToDevice));
  NCHK(n, hipMemset(n->m1, 0, n->nparams * sizeof(float)));
  NCHK(n, hipMemset(n->m2, 0, n->nparams * sizeof(float)));
  n->adam_t = 0;
  if (int rc = pack_all(n)) return rc;
  NCHK(n, hipStreamSynchronize(n->stream));
  return 0;
}

int mcf_net_clear_groups(mcf_net* n) {
  if (!n) return 1;
  for (GroupData& g : n->groups) {
    (void)hipFree(g.c); (void)hipFree(g.y); (void)hipFree(g.feat); (void)hipFree(g.S); (void)hipFree(g.ST);
  }
  n->groups.clear();
  return 0;
}

int mcf_net_add_group(mcf_net* n, int count, int L, const float* c, const float* y, const float* features,
                      const float* S) {
  if (!n || count <= 0 || L < 1 || !c || (!y && !S) || (n->cin > 1 && !features)) return nfail(n, 1, "invalid training group");
  GroupData g;
  g.n = count; g.L = L;
  NCHK(n, hipMalloc(&g.c, (size_t)count * L * sizeof(float)));
  NCHK(n, hipMemcpy(g.c, c, (size_t)count * L * sizeof(float), hipMemcpyHostToDevice));
  if (S) {
    std::vector<float> st((size_t)L * L);
    for (int i = 0; i < L; ++i)
      for (int t = 0; t < L; ++t) st[(size_t)t * L + i] = S[(size_t)i * L + t];
    NCHK(n, hipMalloc(&g.S, (size_t)L * L * sizeof(float)));
    NCHK(n, hipMalloc(&g.ST, (size_t)L * L * sizeof(float)));
    NCHK(n, hipMemcpy(g.S, S, (size_t)L * L * sizeof(float), hipMemcpyHostToDevice));
    NCHK(n, hipMemcpy(g.ST, st.data(), (size_t)L * L * sizeof(float), hipMemcpyHostToDevice));
  } else {
    NCHK(n, hipMalloc(&g.y, (size_t)count * L * sizeof(float)));
    NCHK(n, hipMemcpy(g.y, y, (size_t)count * L * sizeof(float), hipMemcpyHostToDevice));
  }
  if (n->cin > 1) {
    NCHK(n, hipMalloc(&g.feat, (size_t)L * (n->cin - 1) * sizeof(float)));
    NCHK(n, hipMemcpy(g.feat, features, (size_t)L * (n->cin - 1) * sizeof(float), hipMemcpyHostToDevice));
  }
  n->groups.push_back(g);
  return 0;
}

int mcf_net_train(mcf_net* n, int steps, int positions, float lr, unsigned long long seed, float* losses) {
  if (!n || steps < 0 || positions < 1 || !(lr > 0)) return nfail(n, 1, "invalid training request");
  if (n->groups.empty()) return nfail(n, 1, "no training data");
  if (steps == 0) return 0;
  // schedule on the host: group by size, samples uniformly within the group
  std::mt19937_64 rng(seed);
  std::vector<double> wts;
  for (const GroupData& g : n->groups) wts.push_back((double)g.n);
  std::discrete_distribution<int> pick(wts.begin(), wts.end());
  std::vector<int> sg(steps), sb(steps), soff(steps), all;
  int maxP = 0;
  for (int s = 0; s < steps; ++s) {
    const int gi = pick(rng);
    const GroupData& g = n->groups[gi];
    const int B = std::max(1, std::min(g.n, positions / g.L));
    std::uniform_int_distribution<int> u(0, g.n - 1);
    sg[s] = gi; sb[s] = B; soff[s] = (int)all.size();
    for (int b = 0; b < B; ++b) all.push_back(u(rng));
    maxP = std::max(maxP, B * g.L);
  }
  if (int rc = ensure_cap(n, maxP)) return rc;
  if ((int)all.size() > n->idx_cap) {
    (void)hipFree(n->idx);
    n->idx = nullptr;
    n->idx_cap = 0;
    NCHK(n, hipMalloc(&n->idx, all.size() * sizeof(int)));
    n->idx_cap = (int)all.size();
  }
  NCHK(n, hipMemcpy(n->idx, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice));
  float* dl = nullptr;
  NCHK(n, hipMalloc(&dl, steps * sizeof(float)));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int s = 0; s < steps; ++s) {
    const GroupData& g = n->groups[sg[s]];
    const int L = g.L, P = sb[s] * L;
    hipLaunchKernelGGL(k_gather, dim3((P + 255) / 256), dim3(256), 0, n->stream, g.c, g.y, g.feat, n->cin - 1, L,
                       n->idx + soff[s], sb[s], n->act[0], n->cip[0], n->target);
    enqueue_forward(n, L, P);
    enqueue_backward(n, L, P, g.S, g.ST);
    (void)hipMemcpyAsync(dl + s, n->loss, sizeof(float), hipMemcpyDeviceToDevice, n->stream);
    ++n->adam_t;
    const float c1 = 1.f - std::pow(b1, (float)n->adam_t), c2 = 1.f - std::pow(b2, (float)n->adam_t);
    hipLaunchKernelGGL(k_adam, dim3((n->nparams + 255) / 256), dim3(256), 0, n->stream, n->p, n->g, n->m1, n->m2,
                       n->nparams, lr, b1, b2, eps, c1, c2);
    if (int rc = pack_all(n)) { (void)hipFree(dl); return rc; }
  }
  hipError_t e = hipStreamSynchronize(n->stream);
  if (e == hipSuccess && losses) e = hipMemcpy(losses, dl, steps * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(dl);
  if (e != hipSuccess) return nfail(n, 4, std::string("training: ") + hipGetErrorString(e));
  return 0;
}

int mcf_net_predict(mcf_net* n, int nsamp, int L, const float* x, float* y) {
  if (!n || nsamp < 0 || L < 1 || (nsamp && (!x || !y))) return nfail(n, 1, "invalid prediction request");
  if (!nsamp) return 0;
  const int P = nsamp * L;
  if (int rc = upload_inputs(n, nsamp, L, x, nullptr)) return rc;
  enqueue_forward(n, L, P);
  const int last = n->layers - 1;
  hipLaunchKernelGGL(k_take, dim3((P + 255) / 256), dim3(256), 0, n->stream, n->z[last], n->cop[last], P, n->target);
  NCHK(n, hipGetLastError());
  NCHK(n, hipStreamSynchronize(n->stream));
  NCHK(n, hipMemcpy(y, n->target, (size_t)P * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int mcf_net_loss_grad(mcf_net* n, int nsamp, int L, const float* x, const float* y, const float* S, float* loss,
                      float* grad) {
  if (!n || nsamp < 1 || L < 1 || !x || !y || !loss || !grad) return nfail(n, 1, "invalid gradient request");
  const int P = nsamp * L;
  if (int rc = upload_inputs(n, nsamp, L, x, y)) return rc;
  if (S) {
    if (L != n->Sg_L) {
      (void)hipFree(n->Sg); (void)hipFree(n->STg);
      n->Sg = n->STg = nullptr; n->Sg_L = 0;
      NCHK(n, hipMalloc(&n->Sg, (size_t)L * L * sizeof(float)));
      NCHK(n, hipMalloc(&n->STg, (size_t)L * L * sizeof(float)));
      n->Sg_L = L;
    }
    std::vector<float> st((size_t)L * L);
    for (int i = 0; i < L; ++i)
      for (int t = 0; t < L; ++t) st[(size_t)t * L + i] = S[(size_t)i * L + t];
    NCHK(n, hipMemcpy(n->Sg, S, (size_t)L * L * sizeof(float), hipMemcpyHostToDevice));
    NCHK(n, hipMemcpy(n->STg, st.data(), (size_t)L * L * sizeof(float), hipMemcpyHostToDevice));
  }
  enqueue_forward(n, L, P);
  enqueue_backward(n, L, P, S ? n->Sg : nullptr, S ? n->STg : nullptr);
  NCHK(n, hipGetLastError());
  NCHK(n, hipStreamSynchronize(n->stream));
  NCHK(n, hipMemcpy(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost));
  NCHK(n, hipMemcpy(grad, n->g, n->nparams * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

// Average ms per launch of the forward convolution of a hidden layer over P = nsamp * L positions
// (HIP events on the network's stream), for the roofline of the MFMA kernel.
int mcf_net_profile(mcf_net* n, int nsamp, int L, int reps, float* ms_out) {
  if (!n || nsamp < 1 || L < 1 || reps < 1 || !ms_out || n->layers < 3) return nfail(n, 1, "invalid profile request");
  const int P = nsamp * L;
  if (int rc = ensure_cap(n, P)) return rc;
  NCHK(n, hipMemset(n->act[1], 0, (size_t)P * n->cip[1] * sizeof(bf16_t)));
  const int l = 1;
  ConvArgs a{n->wf[l], n->kp[l], n->mrows[l], n->act[l], n->cip[l], n->dil[l], L, P,
             n->p + n->woff[l] + n->co[l] * n->ci[l] * KS, n->co[l], n->cop[l], n->z[l], n->act[l + 1], nullptr, nullptr};
  hipEvent_t e0, e1;
  NCHK(n, hipEventCreate(&e0)); NCHK(n, hipEventCreate(&e1));
  launch_conv<0>(a, (P + PB - 1) / PB, n->stream);
  NCHK(n, hipEventRecord(e0, n->stream));
  for (int r = 0; r < reps; ++r) launch_conv<0>(a, (P + PB - 1) / PB, n->stream);
  NCHK(n, hipEventRecord(e1, n->stream));
  NCHK(n, hipEventSynchronize(e1));
  float ms = 0.f;
  NCHK(n, hipEventElapsedTime(&ms, e0, e1));
  *ms_out = ms / reps;
  (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
  return 0;
}

}  // extern "C"
