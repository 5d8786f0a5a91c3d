// Learned warm start: a 3-D U-Net that predicts lambda0 from the cell-type
// mask, R and the inflow profile, trained on the device.
//
// Network (levels L, base channels c0, c_l = c0 * 2^l):
//   encoder l:  conv3(cin_l -> c_l) ReLU, conv3(c_l -> c_l) ReLU  [skip l]
//               2x2x2 average pool to level l+1            (l < L-1)
//   decoder l:  nearest 2x upsample of level l+1, concat [skip_l, up],
//               conv3(c_l + c_{l+1} -> c_l) ReLU, conv3(c_l -> c_l) ReLU
//   head:       conv1(c_0 -> 1)
// Layout: activations channels-last [voxel][channel] in BF16, voxels x
// fastest.  3x3x3 convolutions are implicit GEMMs on MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation): rows = voxels, columns =
// output channels, K = 27 taps x input channels.  The data gradient is the
// same kernel with tap-flipped, transposed weights; the weight gradient is an
// MFMA GEMM reducing over voxels (split over work-groups, FP32 atomics).
// FP32 master weights, Adam.
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "qesmc_internal.h"

hipError_t qesmc_launch_divergence(qesmc_ctx *x, bool split, int alpha1);

namespace qesmc {
namespace unet {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;  // storage of one BF16 value

__device__ inline bf16_t f2bf(float f)
{
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (bf16_t)(u >> 16);
}
__device__ inline float bf2f(bf16_t h) { return __uint_as_float(((unsigned)h) << 16); }

__device__ inline bf16x8 as_bf16x8(uint4 u) { return __builtin_bit_cast(bf16x8, u); }

// ---------------------------------------------------------------------------
// 3x3x3 convolution, forward (also the data gradient)
// ---------------------------------------------------------------------------
struct ConvFwdArgs {
  const bf16_t *x;      // input [V][ldx], channels [0, cin)
  int ldx, cin, ktot;   // ktot = 27 * cin
  const bf16_t *w;      // packed weights [cout][kp]
  int kp;
  const float *bias;    // [cout] or null
  bf16_t *y;            // output [V][ldy] at channel offset ycoff
  int ldy, ycoff;
  int dx, dy, dz;
  int relu;
};

// A wave computes MT tiles of 16 voxels x NT tiles of 16 output channels.
// MFMA operand layout (16x16x32): lane l holds A[l%16][8*(l/16)+0..7],
// B[8*(l/16)+0..7][l%16]; the result D[4*(l/16)+i][l%16].
template <int NT, int MT>
__global__ __launch_bounds__(256) void k_conv_fwd(ConvFwdArgs a)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long V = (long long)a.dx * a.dy * a.dz;
  const long long mbase = ((long long)blockIdx.x * 4 + wave) * (16 * MT);
  if (mbase >= V) return;
  const int r = lane & 15, oct = lane >> 4;
  const int nb = blockIdx.y * NT;  // first output tile of this launch row
  int vx[MT], vy[MT], vz[MT];
  bool vin[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    long long m = mbase + 16 * t + r;
    vin[t] = m < V;
    if (!vin[t]) m = 0;
    vx[t] = (int)(m % a.dx);
    vy[t] = (int)((m / a.dx) % a.dy);
    vz[t] = (int)(m / ((long long)a.dx * a.dy));
  }
  f32x4 acc[MT][NT];
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[t][n] = f32x4{0.f, 0.f, 0.f, 0.f};

  int tap = 0, c = 8 * oct;
  while (c >= a.cin) {
    c -= a.cin;
    ++tap;
  }
  const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
  if (a.cin % 32 == 0) {
    // Each 32-wide K step lies inside one tap: taps outer, channel chunks
    // inner, same K order as below.  The A loads read a clamped in-bounds
    // address and the zero padding is selected after the load, so the loads
    // of several K steps can be in flight together (the small tiles of the
    // deep levels run long K loops with one output tile per wave).
    constexpr int kTapUnroll = MT * NT <= 2 ? 3 : 1;
#pragma unroll kTapUnroll
    for (int tap = 0; tap < 27; ++tap) {
      const int ox = tap % 3 - 1, oy = (tap / 3) % 3 - 1, oz = tap / 9 - 1;
      long long nvt[MT];
      bool okt[MT];
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        const int nx = vx[t] + ox, ny = vy[t] + oy, nz = vz[t] + oz;
        okt[t] = vin[t] && nx >= 0 && nx < a.dx && ny >= 0 && ny < a.dy && nz >= 0 && nz < a.dz;
        nvt[t] = okt[t] ? ((long long)nz * a.dy + ny) * a.dx + nx : 0;
      }
      for (int c0 = 0; c0 < a.cin; c0 += 32) {
        const int cc = c0 + 8 * oct;
        const int koct = tap * a.cin + cc;
        bf16x8 av[MT];
#pragma unroll
        for (int t = 0; t < MT; ++t) {
          const uint4 raw = *reinterpret_cast<const uint4 *>(a.x + nvt[t] * a.ldx + cc);
          av[t] = as_bf16x8(okt[t] ? raw : zero4);
        }
#pragma unroll
        for (int n = 0; n < NT; ++n) {
          const bf16x8 bv =
              as_bf16x8(*reinterpret_cast<const uint4 *>(a.w + (long long)(16 * (nb + n) + r) * a.kp + koct));
#pragma unroll
          for (int t = 0; t < MT; ++t)
            acc[t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t], bv, acc[t][n], 0, 0, 0);
        }
      }
    }
  } else
  for (int kc = 0; kc < a.kp; kc += 32) {
    const int koct = kc + 8 * oct;
    const bool kvalid = koct < a.ktot;
    const int ox = tap % 3 - 1, oy = (tap / 3) % 3 - 1, oz = tap / 9 - 1;
    bf16x8 av[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) {
      const int nx = vx[t] + ox, ny = vy[t] + oy, nz = vz[t] + oz;
      const bool ok = kvalid && vin[t] && nx >= 0 && nx < a.dx && ny >= 0 && ny < a.dy && nz >= 0 && nz < a.dz;
      const long long nv = ((long long)nz * a.dy + ny) * a.dx + nx;
      av[t] = as_bf16x8(ok ? *reinterpret_cast<const uint4 *>(a.x + nv * a.ldx + c) : zero4);
    }
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const bf16x8 bv = as_bf16x8(*reinterpret_cast<const uint4 *>(a.w + (long long)(16 * (nb + n) + r) * a.kp + koct));
#pragma unroll
      for (int t = 0; t < MT; ++t) acc[t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t], bv, acc[t][n], 0, 0, 0);
    }
    c += 32;
    while (c >= a.cin) {
      c -= a.cin;
      ++tap;
    }
  }
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const int col = 16 * (nb + n) + r;
      const float b = a.bias ? a.bias[col] : 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const long long m = mbase + 16 * t + 4 * oct + i;
        if (m >= V) continue;
        float v = acc[t][n][i] + b;
        if (a.relu) v = v > 0.f ? v : 0.f;
        as_global(a.y)[m * a.ldy + a.ycoff + col] = f2bf(v);
      }
    }
}

// ---------------------------------------------------------------------------
// 3x3x3 convolution, forward, LDS-staged (used when cin is 8 or a multiple
// of 16).  A work-group owns a brick of 8 x 8 x MT output voxels (4*MT
// M-tiles of 16 voxels, MT per wave, x fastest) and all cout.  For each
// chunk of CH input channels it stages the brick plus a one-voxel halo,
// (10 x 10 x (MT+2)) x CH BF16, in LDS (zero outside the volume), then
// every MFMA takes its A fragment from LDS at a per-tap constant offset:
// with CH = 16 one 16x16x32 MFMA covers two taps x 16 channels, with CH = 8
// four taps x 8 channels (27 taps padded to 28 with zeros).  K order
// (tap-major, k = tap*cin + c) is the packed weights', so each lane's B
// octet is 8 consecutive weights.
// ---------------------------------------------------------------------------
template <int NT, int MT, int CH>
__global__ __launch_bounds__(256) void k_conv_fwd_lds(ConvFwdArgs a)
{
  constexpr int HX = 10, HY = 10, HZ = MT + 2, HV = HX * HY * HZ;  // halo brick
  constexpr int OPT = CH / 8;                                      // lane octets per tap
  constexpr int TPM = 4 / OPT;                                     // taps per MFMA
  constexpr int NG = (27 + TPM - 1) / TPM;                         // MFMAs per channel chunk
  __shared__ __attribute__((aligned(16))) bf16_t lds[HV * CH];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, oct = lane >> 4;
  const int nbx = (a.dx + 7) / 8, nby = (a.dy + 7) / 8;
  const int bx = blockIdx.x % nbx, by = (blockIdx.x / nbx) % nby, bz = blockIdx.x / (nbx * nby);
  const int x0 = bx * 8, y0 = by * 8, z0 = bz * MT;
  // this lane's A rows: voxel of row r in each of the wave's MT M-tiles
  int abase[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    const int v = 16 * (wave * MT + t) + r;
    const int lx = v & 7, ly = (v >> 3) & 7, lz = v >> 6;
    abase[t] = (lz * HY + ly) * HX + lx;  // halo origin is (-1,-1,-1)
  }
  f32x4 acc[MT][NT];
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[t][n] = f32x4{0.f, 0.f, 0.f, 0.f};
  const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
  // lane octet -> (tap within the MFMA, channel octet within the chunk)
  const int otap = oct / OPT, och = 8 * (oct % OPT);
  for (int c0 = 0; c0 < a.cin; c0 += CH) {
    // with one output tile the chunk's weight fragments fit in registers:
    // load them before the staging so their latency overlaps it
    bf16x8 bpre[NT == 1 ? NG : 1];
    if constexpr (NT == 1) {
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        const int tap = g * TPM + otap;
        bpre[g] = as_bf16x8(tap < 27 ? *reinterpret_cast<const uint4 *>(a.w + (long long)r * a.kp +
                                                                         (long long)tap * a.cin + c0 + och)
                                     : zero4);
      }
    }
    __syncthreads();  // the previous chunk's readers are done
    for (int e = threadIdx.x; e < HV * OPT; e += 256) {
      const int hv = e / OPT, h = e % OPT;
      const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
      const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
      uint4 val = zero4;
      if (gx >= 0 && gx < a.dx && gy >= 0 && gy < a.dy && gz >= 0 && gz < a.dz)
        val = *reinterpret_cast<const uint4 *>(a.x + (((long long)gz * a.dy + gy) * a.dx + gx) * a.ldx + c0 + 8 * h);
      *reinterpret_cast<uint4 *>(lds + hv * CH + 8 * h) = val;
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      const int tap = g * TPM + otap;
      const bool live = tap < 27;
      const int tz = tap / 9, ty = (tap / 3) % 3, tx = tap % 3;
      const int toff = (tz * HY + ty) * HX + tx;
      bf16x8 av[MT];
#pragma unroll
      for (int t = 0; t < MT; ++t)
        av[t] = as_bf16x8(live ? *reinterpret_cast<const uint4 *>(lds + (abase[t] + toff) * CH + och) : zero4);
      const long long kw = (long long)tap * a.cin + c0 + och;  // weight column of this lane's octet
#pragma unroll
      for (int n = 0; n < NT; ++n) {
        bf16x8 bv;
        if constexpr (NT == 1)
          bv = bpre[g];
        else
          bv = as_bf16x8(live ? *reinterpret_cast<const uint4 *>(a.w + (long long)(16 * n + r) * a.kp + kw) : zero4);
#pragma unroll
        for (int t = 0; t < MT; ++t) acc[t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t], bv, acc[t][n], 0, 0, 0);
      }
    }
  }
  // epilogue: lane holds D[4*oct + i][r] of each tile
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const int col = 16 * n + r;
      const float b = a.bias ? a.bias[col] : 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int v = 16 * (wave * MT + t) + 4 * oct + i;
        const int gx = x0 + (v & 7), gy = y0 + ((v >> 3) & 7), gz = z0 + (v >> 6);
        if (gx >= a.dx || gy >= a.dy || gz >= a.dz) continue;
        float val = acc[t][n][i] + b;
        if (a.relu) val = val > 0.f ? val : 0.f;
        const long long m = ((long long)gz * a.dy + gy) * a.dx + gx;
        as_global(a.y)[m * a.ldy + a.ycoff + col] = f2bf(val);
      }
    }
}

// ---------------------------------------------------------------------------
// 3x3x3 convolution, weight gradient: dW[n][k] += sum_v g[v][n] * X[v+off(k)][c(k)]
// One wave: 16 output channels x (16*KT) columns k, over a voxel range.
// ---------------------------------------------------------------------------
struct ConvWgradArgs {
  const bf16_t *x;
  int ldx, cin, ktot;
  const bf16_t *g;      // gradient w.r.t. the pre-activation output [V][ldg]
  int ldg;
  float *dw;            // [cout][ktot]
  int dx, dy, dz;
  long long vchunk;     // voxels per wave (multiple of 32)
};

template <int KT>
__global__ __launch_bounds__(64) void k_conv_wgrad(ConvWgradArgs a)
{
  const int lane = threadIdx.x, r = lane & 15, oct = lane >> 4;
  const long long V = (long long)a.dx * a.dy * a.dz;
  const int n0 = blockIdx.y * 16;
  const int kk0 = blockIdx.x * 16 * KT;
  const long long vb = (long long)blockIdx.z * a.vchunk;
  if (vb >= V) return;
  const long long ve = vb + a.vchunk < V ? vb + a.vchunk : V;
  int cc[KT], ox[KT], oy[KT], oz[KT];
  bool kv[KT];
#pragma unroll
  for (int t = 0; t < KT; ++t) {
    const int kk = kk0 + 16 * t + r;
    kv[t] = kk < a.ktot;
    const int tap = kv[t] ? kk / a.cin : 0;
    cc[t] = kv[t] ? kk - tap * a.cin : 0;
    ox[t] = tap % 3 - 1;
    oy[t] = (tap / 3) % 3 - 1;
    oz[t] = tap / 9 - 1;
  }
  f32x4 acc[KT];
#pragma unroll
  for (int t = 0; t < KT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const long long plane = (long long)a.dx * a.dy;
  for (long long v0 = vb; v0 < ve; v0 += 32) {
    unsigned short ga[8];
    int px[8], py[8], pz[8];
    bool pv[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const long long v = v0 + 8 * oct + e;
      pv[e] = v < ve;
      const long long vv = pv[e] ? v : 0;
      ga[e] = pv[e] ? a.g[vv * a.ldg + n0 + r] : (unsigned short)0;
      px[e] = (int)(vv % a.dx);
      py[e] = (int)((vv / a.dx) % a.dy);
      pz[e] = (int)(vv / plane);
    }
    uint4 ua;
    ua.x = ga[0] | ((unsigned)ga[1] << 16);
    ua.y = ga[2] | ((unsigned)ga[3] << 16);
    ua.z = ga[4] | ((unsigned)ga[5] << 16);
    ua.w = ga[6] | ((unsigned)ga[7] << 16);
    const bf16x8 av = as_bf16x8(ua);
#pragma unroll
    for (int t = 0; t < KT; ++t) {
      unsigned short xb[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int nx = px[e] + ox[t], ny = py[e] + oy[t], nz = pz[e] + oz[t];
        const bool ok = pv[e] && kv[t] && nx >= 0 && nx < a.dx && ny >= 0 && ny < a.dy && nz >= 0 && nz < a.dz;
        const long long nv = ((long long)nz * a.dy + ny) * a.dx + nx;
        xb[e] = ok ? a.x[nv * a.ldx + cc[t]] : (unsigned short)0;
      }
      uint4 ub;
      ub.x = xb[0] | ((unsigned)xb[1] << 16);
      ub.y = xb[2] | ((unsigned)xb[3] << 16);
      ub.z = xb[4] | ((unsigned)xb[5] << 16);
      ub.w = xb[6] | ((unsigned)xb[7] << 16);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, as_bf16x8(ub), acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < KT; ++t) {
    const int kk = kk0 + 16 * t + r;
    if (!kv[t]) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) atomicAdd(a.dw + (long long)(n0 + 4 * oct + i) * a.ktot + kk, acc[t][i]);
  }
}

// bias gradient: db[n] += sum_v g[v][n]
__global__ __launch_bounds__(256) void k_bias_grad(const bf16_t *g, int ldg, int cout, long long V,
                                                   long long vchunk, float *db)
{
  const int groups = 256 / cout;
  const int n = threadIdx.x % cout, grp = threadIdx.x / cout;
  const long long vb = (long long)blockIdx.x * vchunk;
  const long long ve = vb + vchunk < V ? vb + vchunk : V;
  float s = 0.f;
  if (grp < groups)
    for (long long v = vb + grp; v < ve; v += groups) s += bf2f(g[v * ldg + n]);
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < cout) {
    float t = 0.f;
    for (int q = 0; q < groups; ++q) t += red[q * cout + threadIdx.x];
    atomicAdd(db + threadIdx.x, t);
  }
}

// g <- g * (y > 0), both [V][ld] slices of `ch` channels
__global__ __launch_bounds__(256) void k_relu_bwd(bf16_t *g, int ldg, int goff, const bf16_t *y, int ldy,
                                                  int yoff, int ch, long long V)
{
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= V * ch) return;
  const long long v = q / ch;
  const int c = (int)(q - v * ch);
  if (!(bf2f(y[v * ldy + yoff + c]) > 0.f)) g[v * ldg + goff + c] = 0;
}

// 2x2x2 average pool: out[V/8][ldo] <- in[V][ldi] (channels [ioff, ioff+ch))
__global__ __launch_bounds__(256) void k_pool_fwd(const bf16_t *in, int ldi, int ioff, bf16_t *out, int ldo, int ch,
                                                  int odx, int ody, int odz)
{
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long Vo = (long long)odx * ody * odz;
  if (q >= Vo * ch) return;
  const long long vo = q / ch;
  const int c = (int)(q - vo * ch);
  const int x = (int)(vo % odx), y = (int)((vo / odx) % ody), z = (int)(vo / ((long long)odx * ody));
  const int idx = 2 * odx, idy = 2 * ody;
  float s = 0.f;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const long long vi = ((long long)(2 * z + (b >> 2)) * idy + (2 * y + ((b >> 1) & 1))) * idx + 2 * x + (b & 1);
    s += bf2f(in[vi * ldi + ioff + c]);
  }
  out[vo * ldo + c] = f2bf(s * 0.125f);
}

// gin[child] += gout[parent] / 8
__global__ __launch_bounds__(256) void k_pool_bwd_add(const bf16_t *gout, int ldo, bf16_t *gin, int ldi, int ioff,
                                                      int ch, int odx, int ody, int odz)
{
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long Vi = 8LL * odx * ody * odz;
  if (q >= Vi * ch) return;
  const long long vi = q / ch;
  const int c = (int)(q - vi * ch);
  const int idx = 2 * odx, idy = 2 * ody;
  const int x = (int)(vi % idx), y = (int)((vi / idx) % idy), z = (int)(vi / ((long long)idx * idy));
  const long long vo = ((long long)(z >> 1) * ody + (y >> 1)) * odx + (x >> 1);
  bf16_t &d = gin[vi * ldi + ioff + c];
  d = f2bf(bf2f(d) + 0.125f * bf2f(gout[vo * ldo + c]));
}

// nearest 2x upsample: out[V][ldo] (channels [ooff, ooff+ch)) <- in[V/8][ldi]
__global__ __launch_bounds__(256) void k_up_fwd(const bf16_t *in, int ldi, bf16_t *out, int ldo, int ooff, int ch,
                                                int dx, int dy, int dz)
{
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long V = (long long)dx * dy * dz;
  if (q >= V * ch) return;
  const long long v = q / ch;
  const int c = (int)(q - v * ch);
  const int x = (int)(v % dx), y = (int)((v / dx) % dy), z = (int)(v / ((long long)dx * dy));
  const long long vp = ((long long)(z >> 1) * (dy >> 1) + (y >> 1)) * (dx >> 1) + (x >> 1);
  out[v * ldo + ooff + c] = in[vp * ldi + c];
}

// gin[parent] = sum of gout over its 8 children
__global__ __launch_bounds__(256) void k_up_bwd(const bf16_t *gout, int ldo, int ooff, bf16_t *gin, int ldi, int ch,
                                                int pdx, int pdy, int pdz)
{
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long Vp = (long long)pdx * pdy * pdz;
  if (q >= Vp * ch) return;
  const long long vp = q / ch;
  const int c = (int)(q - vp * ch);
  const int x = (int)(vp % pdx), y = (int)((vp / pdx) % pdy), z = (int)(vp / ((long long)pdx * pdy));
  const int dx = 2 * pdx, dy = 2 * pdy;
  float s = 0.f;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const long long v = ((long long)(2 * z + (b >> 2)) * dy + (2 * y + ((b >> 1) & 1))) * dx + 2 * x + (b & 1);
    s += bf2f(gout[v * ldo + ooff + c]);
  }
  gin[vp * ldi + c] = f2bf(s);
}

// The same three data-movement steps with 8 channels (16 bytes) per thread
// and 32-bit index math, for channel counts, strides and offsets that are
// multiples of 8 and V * ch / 8 < 2^32 (the host checks): same arithmetic in
// the same order per channel, so the same results.
__device__ inline float bf_at(const uint4 &t, int e)
{
  const unsigned w = e < 2 ? t.x : e < 4 ? t.y : e < 6 ? t.z : t.w;
  return __uint_as_float((e & 1) ? (w & 0xffff0000u) : (w << 16));
}

__global__ __launch_bounds__(256) void k_pool_fwd8(const bf16_t *in, int ldi, int ioff, bf16_t *out, int ldo, int ch,
                                                   int odx, int ody, int odz)
{
  const unsigned g = (unsigned)ch >> 3;
  const unsigned q = blockIdx.x * 256u + threadIdx.x;
  if (q >= (unsigned)odx * ody * odz * g) return;
  const unsigned vo = q / g;
  const unsigned c8 = (q - vo * g) * 8u;
  const unsigned x = vo % odx, y = (vo / odx) % ody, z = vo / ((unsigned)odx * ody);
  const unsigned idx = 2u * odx, idy = 2u * ody;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const unsigned vi = ((2u * z + (b >> 2)) * idy + (2u * y + ((b >> 1) & 1))) * idx + 2u * x + (b & 1);
    const uint4 t = *reinterpret_cast<const uint4 *>(in + (size_t)vi * ldi + ioff + c8);
#pragma unroll
    for (int e = 0; e < 8; ++e) s[e] += bf_at(t, e);
  }
  unsigned o[4];
#pragma unroll
  for (int e = 0; e < 4; ++e)
    o[e] = (unsigned)f2bf(s[2 * e] * 0.125f) | ((unsigned)f2bf(s[2 * e + 1] * 0.125f) << 16);
  *reinterpret_cast<uint4 *>(out + (size_t)vo * ldo + c8) = make_uint4(o[0], o[1], o[2], o[3]);
}

__global__ __launch_bounds__(256) void k_up_fwd8(const bf16_t *in, int ldi, bf16_t *out, int ldo, int ooff, int ch,
                                                 int dx, int dy, int dz)
{
  const unsigned g = (unsigned)ch >> 3;
  const unsigned q = blockIdx.x * 256u + threadIdx.x;
  if (q >= (unsigned)dx * dy * dz * g) return;
  const unsigned v = q / g;
  const unsigned c8 = (q - v * g) * 8u;
  const unsigned x = v % dx, y = (v / dx) % dy, z = v / ((unsigned)dx * dy);
  const unsigned vp = ((z >> 1) * ((unsigned)dy >> 1) + (y >> 1)) * ((unsigned)dx >> 1) + (x >> 1);
  *reinterpret_cast<uint4 *>(out + (size_t)v * ldo + ooff + c8) =
      *reinterpret_cast<const uint4 *>(in + (size_t)vp * ldi + c8);
}

__global__ __launch_bounds__(256) void k_head_fwd8(const bf16_t *x, int ldx, int ch, const float *w, float *y,
                                                   long long V)
{
  const long long v = (long long)blockIdx.x * 256 + threadIdx.x;
  if (v >= V) return;
  float s = w[ch];
  for (int c8 = 0; c8 < ch; c8 += 8) {
    const uint4 t = *reinterpret_cast<const uint4 *>(x + v * ldx + c8);
#pragma unroll
    for (int e = 0; e < 8; ++e) s += w[c8 + e] * bf_at(t, e);
  }
  y[v] = s;
}

// head: y[v] = sum_c W[c] x[v][c] + b
__global__ __launch_bounds__(256) void k_head_fwd(const bf16_t *x, int ldx, int ch, const float *w, float *y,
                                                  long long V)
{
  const long long v = (long long)blockIdx.x * 256 + threadIdx.x;
  if (v >= V) return;
  float s = w[ch];
  for (int c = 0; c < ch; ++c) s += w[c] * bf2f(x[v * ldx + c]);
  y[v] = s;
}

// head backward: gx[v][c] = gy[v] W[c]; dW[c] += sum gy x; db += sum gy
__global__ __launch_bounds__(256) void k_head_bwd(const bf16_t *x, int ldx, int ch, const float *w, const float *gy,
                                                  bf16_t *gx, float *dw, long long V)
{
  const long long v = (long long)blockIdx.x * 256 + threadIdx.x;
  __shared__ float red[256];
  float acc[33];
  for (int c = 0; c <= ch; ++c) acc[c] = 0.f;
  if (v < V) {
    const float g = gy[v];
    for (int c = 0; c < ch; ++c) {
      gx[v * ldx + c] = f2bf(g * w[c]);
      acc[c] = g * bf2f(x[v * ldx + c]);
    }
    acc[ch] = g;
  }
  for (int c = 0; c <= ch; ++c) {
    red[threadIdx.x] = acc[c];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(dw + c, red[0]);
    __syncthreads();
  }
}

// MSE loss over the volume: gy = 2 (y - t) / V, loss += (y - t)^2 / V
__global__ __launch_bounds__(256) void k_mse(const float *y, const float *t, float inv_scale, float *gy, long long V,
                                             float *loss)
{
  const long long v = (long long)blockIdx.x * 256 + threadIdx.x;
  float e = 0.f;
  if (v < V) {
    const float d = y[v] - t[v] * inv_scale;
    gy[v] = 2.f * d / (float)V;
    e = d * d / (float)V;
  }
  __shared__ float red[256];
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// Block-mean loss term.  The error e = y - t (k_mse leaves gy = 2 e / V) is
// averaged over blocks of S^3 voxels; loss += weight * mean_b(ebar_b^2) and
// gy += weight * 2 ebar_b / (NB * S^3) for every voxel of block b.  Smooth
// errors (which SOR removes slowest) are penalised by (1 + weight), errors
// that change sign within a block by about 1.  One work-group per block.
__global__ __launch_bounds__(256) void k_block_err(const float *gy, float half_v, int dx, int dy, int S,
                                                   float w_over_nb, float *eblk, float *loss)
{
  const int nbx = dx / S, nby = dy / S;
  const int b = blockIdx.x;
  const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
  const int n = S * S * S;
  float acc = 0.f;
  for (int q = threadIdx.x; q < n; q += 256) {
    const int x = q % S, y = (q / S) % S, z = q / (S * S);
    const long long v = ((long long)(bz * S + z) * dy + (by * S + y)) * dx + (bx * S + x);
    acc += gy[v] * half_v;
  }
  __shared__ float red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float m = red[0] / (float)n;
    eblk[b] = m;
    if (loss) atomicAdd(loss, w_over_nb * m * m);
  }
}

__global__ __launch_bounds__(256) void k_block_grad(float *gy, const float *eblk, int dx, int dy, int dz, int S,
                                                    float coef)
{
  const long long V = (long long)dx * dy * dz;
  const long long v = (long long)blockIdx.x * 256 + threadIdx.x;
  if (v >= V) return;
  const int x = (int)(v % dx), y = (int)((v / dx) % dy), z = (int)(v / ((long long)dx * dy));
  const int b = ((z / S) * (dy / S) + (y / S)) * (dx / S) + x / S;
  gy[v] += coef * eblk[b];
}

// acc += sum of t^2
__global__ __launch_bounds__(256) void k_sumsq(const float *t, long long n, double *acc)
{
  double s = 0.0;
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256)
    s += (double)t[q] * t[q];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(acc, red[0]);
}

__global__ __launch_bounds__(256) void k_adam(float *p, const float *g, float *m, float *v, long long n, float lr,
                                              float b1, float b2, float eps, float c1, float c2)
{
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  const float gq = g[q];
  const float mq = b1 * m[q] + (1.f - b1) * gq;
  const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
  m[q] = mq;
  v[q] = vq;
  p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

// packed BF16 weights from the FP32 master W[cout][27*cin]:
//   forward  wt[n][kp],  wt[n][tap*cin + c] = W[n][tap*cin + c]
//   dgrad    wd[c][kpd], wd[c][tap'*cout + n] = W[n][(26 - tap')*cin + c]
__global__ __launch_bounds__(256) void k_pack_w(const float *W, int cin, int cout, int kp, int kpd, bf16_t *wt,
                                                bf16_t *wd)
{
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long nt = (long long)cout * kp, nd = (long long)cin * kpd;
  if (q < nt) {
    const int n = (int)(q / kp), k = (int)(q - (long long)n * kp);
    wt[q] = k < 27 * cin ? f2bf(W[(long long)n * 27 * cin + k]) : (bf16_t)0;
  } else if (q < nt + nd) {
    const long long r = q - nt;
    const int c = (int)(r / kpd), k = (int)(r - (long long)c * kpd);
    bf16_t val = 0;
    if (k < 27 * cout) {
      const int tp = k / cout, n = k - tp * cout;
      val = f2bf(W[(long long)n * 27 * cin + (26 - tp) * cin + c]);
    }
    wd[r] = val;
  }
}

// fp32 host volume -> BF16 device volume
__global__ __launch_bounds__(256) void k_f2bf(const float *src, bf16_t *dst, long long n)
{
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q < n) dst[q] = f2bf(src[q]);
}

// Per-problem scales: s = rms(R) over updated cells, u = max inflow speed of
// the boundary column (0,0).  sc[0] = s, sc[1] = u.
__global__ __launch_bounds__(256) void k_scales(Dims d, const float *R, const float *u0, const float *v0,
                                                double *acc, float *sc, int stage)
{
  if (stage == 0) {
    const int W = d.NX - 2, H = d.NY - 2;
    const long long n = (long long)W * H * (d.NZ - 2);
    double s = 0.0;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
      const long long k = q / ((long long)W * H);
      const long long rem = q - k * W * H;
      const long long j = rem / W, i = rem - j * W;
      const float r = R[(i + 1) + (j + 1) * d.sy + (k + 1) * d.sz];
      s += (double)r * r;
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int t = 128; t > 0; t >>= 1) {
      if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
      __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(acc, red[0]);
  } else if (blockIdx.x == 0 && threadIdx.x == 0) {
    const long long n = (long long)(d.NX - 2) * (d.NY - 2) * (d.NZ - 2);
    double ms = n > 0 ? acc[0] / (double)n : 0.0;
    float s = (float)sqrt(ms);
    sc[0] = s > 1e-30f ? s : 1.f;
    float um = 0.f;
    for (int k = 0; k < d.nz; ++k) {
      const float a = u0[k * d.fz], b = v0[k * d.fz];
      um = fmaxf(um, sqrtf(a * a + b * b));
    }
    sc[1] = um > 1e-30f ? um : 1.f;
  }
}

// Input features of the padded volume (dx, dy, dz): net voxel (x, y, z) is
// cell (x, y, z+1).  ch0 fluid mask, ch1 R/s, ch2-3 inflow (u, v)/speed at
// that height, ch4 normalised height, ch5-7 zero.
__global__ __launch_bounds__(256) void k_features(Dims d, const int *flag, const float *R, const float *u0,
                                                  const float *v0, const float *sc, int ncx, int ncy, int ncz,
                                                  int dx, int dy, int dz, bf16_t *feat)
{
  const long long v = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long V = (long long)dx * dy * dz;
  if (v >= V) return;
  const int x = (int)(v % dx), y = (int)((v / dx) % dy), z = (int)(v / ((long long)dx * dy));
  float f[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (x < ncx && y < ncy && z < ncz) {
    const int k = z + 1;
    const long long c = x + (long long)y * d.sy + (long long)k * d.sz;
    f[0] = flag[c] == 1 ? 1.f : 0.f;
    f[1] = R[c] / sc[0];
    f[2] = u0[(long long)k * d.fz] / sc[1];
    f[3] = v0[(long long)k * d.fz] / sc[1];
    f[4] = (z + 0.5f) / (float)ncz;
  }
  uint4 u;
  u.x = f2bf(f[0]) | ((unsigned)f2bf(f[1]) << 16);
  u.y = f2bf(f[2]) | ((unsigned)f2bf(f[3]) << 16);
  u.z = f2bf(f[4]) | ((unsigned)f2bf(f[5]) << 16);
  u.w = f2bf(f[6]) | ((unsigned)f2bf(f[7]) << 16);
  *reinterpret_cast<uint4 *>(feat + v * 8) = u;
}

// target[v] = lambda(cell)/s (0 outside the problem)
__global__ __launch_bounds__(256) void k_target(Dims d, const float *lam, const float *sc, int ncx, int ncy, int ncz,
                                                int dx, int dy, int dz, float *tgt)
{
  const long long v = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long V = (long long)dx * dy * dz;
  if (v >= V) return;
  const int x = (int)(v % dx), y = (int)((v / dx) % dy), z = (int)(v / ((long long)dx * dy));
  float t = 0.f;
  if (x < ncx && y < ncy && z < ncz) t = lam[x + (long long)y * d.sy + (long long)(z + 1) * d.sz] / sc[0];
  tgt[v] = t;
}

// lambda0 of the solver from the prediction: updated cells get out*s, every
// other cell 0, then the k = 0 ghost mirrors k = 1.
__global__ __launch_bounds__(256) void k_install(Dims d, const float *out, const float *sc, float tscale, int dx,
                                                 int dy, float *lam)
{
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long n = (long long)d.NX * d.NY * d.NZ;
  if (c >= n) return;
  const int k = (int)(c / d.sz);
  const long long rem = c - (long long)k * d.sz;
  const int j = (int)(rem / d.sy), i = (int)(rem - (long long)j * d.sy);
  const int kk = k == 0 ? 1 : k;  // ghost layer mirrors layer 1
  float val = 0.f;
  if (i >= 1 && i <= d.NX - 2 && j >= 1 && j <= d.NY - 2 && kk >= 1 && kk <= d.NZ - 2 && d.NZ >= 3)
    val = out[((long long)(kk - 1) * dy + j) * dx + i] * tscale * sc[0];
  lam[c] = val;
}

}  // namespace unet
}  // namespace qesmc

using namespace qesmc;
using namespace qesmc::unet;

namespace {

struct Conv {
  int cin, cout, kp, kpd;
  long long woff, boff;  // offsets in the parameter vector
  bf16_t *wt = nullptr, *wd = nullptr;
};

struct Level {
  int dx, dy, dz;
  long long V;
  int c;        // channels of this level
  int ldcat;    // channels of CAT (c + c_{l+1}, or c at the bottom)
  bf16_t *A = nullptr, *CAT = nullptr, *P = nullptr, *E = nullptr, *D = nullptr;
  bf16_t *gA = nullptr, *gCAT = nullptr, *gP = nullptr, *gE = nullptr, *gD = nullptr;
};

inline int round_up(int a, int b) { return (a + b - 1) / b * b; }

}  // namespace

struct qesmc_unet {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {};
  int L = 4, c0 = 16;
  int ncx = 0, ncy = 0, ncz = 0;
  int dx = 0, dy = 0, dz = 0;
  std::vector<Conv> enc1, enc2, dec1, dec2;  // per level
  long long head_off = 0, nparams = 0;
  float *p = nullptr, *gp = nullptr, *m1 = nullptr, *m2 = nullptr;
  int adam_t = 0;
  std::vector<Level> lv;
  bf16_t *feat = nullptr;        // scratch input [V0][8]
  float *out = nullptr, *gout = nullptr, *tgt = nullptr, *hostbuf_dev = nullptr;
  float *sc = nullptr;           // scales of the current problem
  double *dacc = nullptr;
  float *loss = nullptr;
  std::vector<bf16_t *> sfeat;   // dataset
  std::vector<float *> stgt;
  double conv_flops = 0.0;       // forward FLOPs of all convolutions
  float tscale = 0.f;            // output unit (0: derive from the samples)
  int loss_block = 0;            // block-mean loss term: block edge (0: off)
  float loss_weight = 0.f;       //   and its weight
  float *eblk = nullptr;         //   block means of the error
};

namespace {

#define UCHK(expr)                                   \
  do {                                               \
    hipError_t _e = (expr);                          \
    if (_e != hipSuccess) return map_err(_e);        \
  } while (0)

int map_err(hipError_t e)
{
  if (e == hipErrorOutOfMemory) return QESMC_ERR_OOM;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return QESMC_ERR_NO_DEVICE;
  return QESMC_ERR_HIP;
}

template <typename T>
hipError_t dmalloc(T **p, long long n)
{
  *p = nullptr;
  return hipMalloc((void **)p, sizeof(T) * (size_t)(n > 0 ? n : 1));
}

inline dim3 g1(long long n) { return dim3((unsigned)((n + 255) / 256)); }

// the 8-channel data-movement kernels apply: channels, strides and offset in
// whole 16-byte groups, and every index fits 32 bits
inline bool vec8(int ch, int ld_a, int ld_b, int off, long long V)
{
  return ch % 8 == 0 && ld_a % 8 == 0 && ld_b % 8 == 0 && off % 8 == 0 && V * (long long)(ld_a > ld_b ? ld_a : ld_b) < (1LL << 31);
}

hipError_t launch_conv_fwd(hipStream_t s, const ConvFwdArgs &a, int cout)
{
  const long long V = (long long)a.dx * a.dy * a.dz;
  const int NT = cout / 16;
  // LDS-staged kernel: cin 8 or a multiple of 16, 16-byte aligned channel
  // rows, and enough 8x8x4 bricks to fill the GPU (the large, shallow levels;
  // the deep levels have too few voxels for bricks and use the kernel above)
  const long long bxy = (long long)((a.dx + 7) / 8) * ((a.dy + 7) / 8);
  if ((a.cin == 8 || a.cin % 16 == 0) && a.ldx % 8 == 0 && cout % 16 == 0 && bxy * ((a.dz + 3) / 4) >= 1024) {
    auto lds = [&](auto kern, int MT) {
      const long long nb = bxy * ((a.dz + MT - 1) / MT);
      hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(256), 0, s, a);
      return hipGetLastError();
    };
    if (a.cin == 8) {
      switch (NT) {
      case 1: return lds(k_conv_fwd_lds<1, 4, 8>, 4);
      case 2: return lds(k_conv_fwd_lds<2, 4, 8>, 4);
      default: break;
      }
    } else {
      switch (NT) {
      case 1: return lds(k_conv_fwd_lds<1, 4, 16>, 4);
      case 2: return lds(k_conv_fwd_lds<2, 4, 16>, 4);
      case 3: return lds(k_conv_fwd_lds<3, 4, 16>, 4);
      case 4: return lds(k_conv_fwd_lds<4, 4, 16>, 4);
      case 6: return lds(k_conv_fwd_lds<6, 2, 16>, 2);
      case 8: return lds(k_conv_fwd_lds<8, 2, 16>, 2);
      case 12: return lds(k_conv_fwd_lds<12, 2, 16>, 2);
      default: break;
      }
    }
  }
  auto go = [&](auto kern, int MT, int rows = 1) {
    const long long waves = (V + 16LL * MT - 1) / (16LL * MT);
    hipLaunchKernelGGL(kern, dim3((unsigned)((waves + 3) / 4), (unsigned)rows), dim3(256), 0, s, a);
  };
  // Small volumes (the deep levels, fewer than 2048 waves with the tiles
  // below): one output tile per wave and fewer voxel tiles, so that up to
  // 4096 waves cover the GPU.
  const long long want = 4096;
  const int mt_default = NT == 1 ? 8 : NT == 2 ? 4 : NT <= 4 ? 2 : 1;
  if ((V + 16LL * mt_default - 1) / (16LL * mt_default) < 2048) {
    const long long w4 = (V + 63) / 64 * NT, w2 = (V + 31) / 32 * NT;
    if (w4 >= want) go(k_conv_fwd<1, 4>, 4, NT);
    else if (w2 >= want) go(k_conv_fwd<1, 2>, 2, NT);
    else go(k_conv_fwd<1, 1>, 1, NT);
    return hipGetLastError();
  }
  switch (NT) {
  case 1: go(k_conv_fwd<1, 8>, 8); break;
  case 2: go(k_conv_fwd<2, 4>, 4); break;
  case 3: go(k_conv_fwd<3, 2>, 2); break;
  case 4: go(k_conv_fwd<4, 2>, 2); break;
  case 6: go(k_conv_fwd<6, 1>, 1); break;
  case 8: go(k_conv_fwd<8, 1>, 1); break;
  case 12: go(k_conv_fwd<12, 1>, 1); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_wgrad(hipStream_t s, const ConvWgradArgs &a0, int cout)
{
  ConvWgradArgs a = a0;
  const long long V = (long long)a.dx * a.dy * a.dz;
  constexpr int KT = 4;
  const int kblocks = (a.ktot + 16 * KT - 1) / (16 * KT);
  const int nblocks = cout / 16;
  // about 8192 waves in flight, voxel chunks of >= 256
  long long splits = 8192 / (kblocks * nblocks);
  if (splits < 1) splits = 1;
  long long chunk = (V + splits - 1) / splits;
  chunk = (chunk + 31) / 32 * 32;
  if (chunk < 256) chunk = 256;
  splits = (V + chunk - 1) / chunk;
  a.vchunk = chunk;
  hipLaunchKernelGGL(k_conv_wgrad<KT>, dim3(kblocks, nblocks, (unsigned)splits), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_bias_grad(hipStream_t s, const bf16_t *g, int ldg, int cout, long long V, float *db)
{
  const long long chunk = 4096;
  hipLaunchKernelGGL(k_bias_grad, dim3((unsigned)((V + chunk - 1) / chunk)), dim3(256), 0, s, g, ldg, cout, V,
                     chunk, db);
  return hipGetLastError();
}

// One conv layer forward: x (ldx, cin) -> y (ldy at ycoff), bias + ReLU
hipError_t conv_forward(qesmc_unet *u, const Conv &cv, const Level &lv, const bf16_t *x, int ldx, bf16_t *y,
                        int ldy, int ycoff)
{
  ConvFwdArgs a{x, ldx, cv.cin, 27 * cv.cin, cv.wt, cv.kp, u->p + cv.boff, y, ldy, ycoff, lv.dx, lv.dy, lv.dz, 1};
  return launch_conv_fwd(u->stream, a, cv.cout);
}

// Backward through one conv layer whose ReLU output y (ldy, yoff) received
// the gradient g (ldg, goff) (masked in place); x is its input.  Adds dW, db;
// if gx is given, writes dL/dx into it (ldgx, cin channels at offset 0).
hipError_t conv_backward(qesmc_unet *u, const Conv &cv, const Level &lv, const bf16_t *x, int ldx,
                         const bf16_t *y, int ldy, int yoff, bf16_t *g, int ldg, int goff, bf16_t *gx, int ldgx)
{
  hipStream_t s = u->stream;
  const long long V = lv.V;
  hipLaunchKernelGGL(k_relu_bwd, g1(V * cv.cout), dim3(256), 0, s, g, ldg, goff, y, ldy, yoff, cv.cout, V);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  ConvWgradArgs wa{x, ldx, cv.cin, 27 * cv.cin, g + goff, ldg, u->gp + cv.woff, lv.dx, lv.dy, lv.dz, 0};
  if ((e = launch_wgrad(s, wa, cv.cout)) != hipSuccess) return e;
  if ((e = launch_bias_grad(s, g + goff, ldg, cv.cout, V, u->gp + cv.boff)) != hipSuccess) return e;
  if (gx) {
    ConvFwdArgs a{g + goff, ldg, cv.cout, 27 * cv.cout, cv.wd, cv.kpd, nullptr, gx, ldgx, 0, lv.dx, lv.dy, lv.dz, 0};
    if ((e = launch_conv_fwd(s, a, cv.cin)) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t pack_weights(qesmc_unet *u)
{
  auto pk = [&](Conv &c) {
    const long long n = (long long)c.cout * c.kp + (long long)c.cin * c.kpd;
    hipLaunchKernelGGL(k_pack_w, g1(n), dim3(256), 0, u->stream, u->p + c.woff, c.cin, c.cout, c.kp, c.kpd, c.wt,
                       c.wd);
  };
  for (int l = 0; l < u->L; ++l) {
    pk(u->enc1[l]);
    pk(u->enc2[l]);
  }
  for (int l = 0; l + 1 < u->L; ++l) {
    pk(u->dec1[l]);
    pk(u->dec2[l]);
  }
  return hipGetLastError();
}

hipError_t forward(qesmc_unet *u, const bf16_t *feat)
{
  hipStream_t s = u->stream;
  hipError_t e;
  const int L = u->L;
  for (int l = 0; l < L; ++l) {
    Level &lv = u->lv[l];
    const bf16_t *in = l == 0 ? feat : lv.P;
    const int ldin = l == 0 ? QESMC_UNET_IN_CHANNELS : u->lv[l - 1].c;
    if ((e = conv_forward(u, u->enc1[l], lv, in, ldin, lv.A, lv.c, 0)) != hipSuccess) return e;
    if ((e = conv_forward(u, u->enc2[l], lv, lv.A, lv.c, lv.CAT, lv.ldcat, 0)) != hipSuccess) return e;
    if (l + 1 < L) {
      Level &nx = u->lv[l + 1];
      if (vec8(lv.c, lv.ldcat, lv.c, 0, nx.V))
        hipLaunchKernelGGL(k_pool_fwd8, g1(nx.V * lv.c / 8), dim3(256), 0, s, lv.CAT, lv.ldcat, 0, nx.P, lv.c, lv.c,
                           nx.dx, nx.dy, nx.dz);
      else
        hipLaunchKernelGGL(k_pool_fwd, g1(nx.V * lv.c), dim3(256), 0, s, lv.CAT, lv.ldcat, 0, nx.P, lv.c, lv.c,
                           nx.dx, nx.dy, nx.dz);
    }
  }
  for (int l = L - 2; l >= 0; --l) {
    Level &lv = u->lv[l];
    Level &dn = u->lv[l + 1];
    const bf16_t *src = (l == L - 2) ? dn.CAT : dn.D;
    const int ldsrc = (l == L - 2) ? dn.ldcat : dn.c;
    if (vec8(dn.c, ldsrc, lv.ldcat, lv.c, lv.V))
      hipLaunchKernelGGL(k_up_fwd8, g1(lv.V * dn.c / 8), dim3(256), 0, s, src, ldsrc, lv.CAT, lv.ldcat, lv.c, dn.c,
                         lv.dx, lv.dy, lv.dz);
    else
      hipLaunchKernelGGL(k_up_fwd, g1(lv.V * dn.c), dim3(256), 0, s, src, ldsrc, lv.CAT, lv.ldcat, lv.c, dn.c, lv.dx,
                         lv.dy, lv.dz);
    if ((e = conv_forward(u, u->dec1[l], lv, lv.CAT, lv.ldcat, lv.E, lv.c, 0)) != hipSuccess) return e;
    if ((e = conv_forward(u, u->dec2[l], lv, lv.E, lv.c, lv.D, lv.c, 0)) != hipSuccess) return e;
  }
  Level &top = u->lv[0];
  if (vec8(top.c, top.c, top.c, 0, top.V))
    hipLaunchKernelGGL(k_head_fwd8, g1(top.V), dim3(256), 0, s, top.D, top.c, top.c, u->p + u->head_off, u->out,
                       top.V);
  else
    hipLaunchKernelGGL(k_head_fwd, g1(top.V), dim3(256), 0, s, top.D, top.c, top.c, u->p + u->head_off, u->out,
                       top.V);
  return hipGetLastError();
}

// gradients of all parameters (added into u->gp, which the caller zeroes)
// from u->gout (dL/d output)
hipError_t backward(qesmc_unet *u, const bf16_t *feat)
{
  hipStream_t s = u->stream;
  hipError_t e;
  const int L = u->L;
  Level &top = u->lv[0];
  hipLaunchKernelGGL(k_head_bwd, g1(top.V), dim3(256), 0, s, top.D, top.c, top.c, u->p + u->head_off, u->gout,
                     top.gD, u->gp + u->head_off, top.V);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  for (int l = 0; l + 1 < L; ++l) {
    Level &lv = u->lv[l];
    Level &dn = u->lv[l + 1];
    if ((e = conv_backward(u, u->dec2[l], lv, lv.E, lv.c, lv.D, lv.c, 0, lv.gD, lv.c, 0, lv.gE, lv.c)) != hipSuccess)
      return e;
    if ((e = conv_backward(u, u->dec1[l], lv, lv.CAT, lv.ldcat, lv.E, lv.c, 0, lv.gE, lv.c, 0, lv.gCAT, lv.ldcat)) !=
        hipSuccess)
      return e;
    bf16_t *gsrc = (l == L - 2) ? dn.gCAT : dn.gD;
    const int ldg = (l == L - 2) ? dn.ldcat : dn.c;
    hipLaunchKernelGGL(k_up_bwd, g1(dn.V * dn.c), dim3(256), 0, s, lv.gCAT, lv.ldcat, lv.c, gsrc, ldg, dn.c, dn.dx,
                       dn.dy, dn.dz);
  }
  for (int l = L - 1; l >= 0; --l) {
    Level &lv = u->lv[l];
    if (l + 1 < L) {
      Level &dn = u->lv[l + 1];
      hipLaunchKernelGGL(k_pool_bwd_add, g1(lv.V * lv.c), dim3(256), 0, s, dn.gP, lv.c, lv.gCAT, lv.ldcat, 0, lv.c,
                         dn.dx, dn.dy, dn.dz);
    }
    if ((e = conv_backward(u, u->enc2[l], lv, lv.A, lv.c, lv.CAT, lv.ldcat, 0, lv.gCAT, lv.ldcat, 0, lv.gA, lv.c)) !=
        hipSuccess)
      return e;
    const bf16_t *in = l == 0 ? feat : lv.P;
    const int ldin = l == 0 ? QESMC_UNET_IN_CHANNELS : u->lv[l - 1].c;
    if ((e = conv_backward(u, u->enc1[l], lv, in, ldin, lv.A, lv.c, 0, lv.gA, lv.c, 0, l > 0 ? lv.gP : nullptr,
                           ldin)) != hipSuccess)
      return e;
  }
  return hipSuccess;
}

void free_all(qesmc_unet *u)
{
  auto f = [](auto *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
  };
  for (auto *v : {&u->enc1, &u->enc2, &u->dec1, &u->dec2})
    for (auto &c : *v) {
      f(c.wt);
      f(c.wd);
    }
  for (auto &l : u->lv) {
    f(l.A), f(l.CAT), f(l.P), f(l.E), f(l.D), f(l.gA), f(l.gCAT), f(l.gP), f(l.gE), f(l.gD);
  }
  f(u->p), f(u->gp), f(u->m1), f(u->m2), f(u->feat), f(u->out), f(u->gout), f(u->tgt), f(u->hostbuf_dev);
  f(u->sc), f(u->dacc), f(u->loss), f(u->eblk);
  for (auto &q : u->sfeat) f(q);
  for (auto &q : u->stgt) f(q);
  u->sfeat.clear();
  u->stgt.clear();
}

// Loss of the current output against target t into u->loss, gradient into
// u->gout: mean squared error plus the optional block-mean term.
hipError_t launch_loss(qesmc_unet *u, const float *t, bool grad)
{
  hipStream_t s = u->stream;
  const long long V0 = (long long)u->dx * u->dy * u->dz;
  hipLaunchKernelGGL(k_mse, g1(V0), dim3(256), 0, s, u->out, t, 1.f / u->tscale, u->gout, V0, u->loss);
  if (u->loss_block > 0 && u->loss_weight > 0.f) {
    const int S = u->loss_block;
    const long long nb = (long long)(u->dx / S) * (u->dy / S) * (u->dz / S);
    hipLaunchKernelGGL(k_block_err, dim3((unsigned)nb), dim3(256), 0, s, u->gout, 0.5f * (float)V0, u->dx, u->dy, S,
                       u->loss_weight / (float)nb, u->eblk, u->loss);
    if (grad)
      hipLaunchKernelGGL(k_block_grad, g1(V0), dim3(256), 0, s, u->gout, u->eblk, u->dx, u->dy, u->dz, S,
                         2.f * u->loss_weight / ((float)nb * (float)(S * S * S)));
  }
  return hipGetLastError();
}

// features + scales of ctx's resident problem into u->feat, u->sc
int problem_features(qesmc_unet *u, qesmc_ctx *x)
{
  if (!x->have_geom || !x->have_inflow) return QESMC_ERR_STATE;
  if (x->d.nx - 1 != u->ncx || x->d.ny - 1 != u->ncy || x->d.nz - 2 != u->ncz) return QESMC_ERR_ARG;
  UCHK(hipStreamSynchronize(x->stream));
  // R on the x-fastest layout, on the solver's stream
  UCHK(qesmc_launch_divergence(x, false, 1));
  UCHK(hipStreamSynchronize(x->stream));
  hipStream_t s = u->stream;
  UCHK(hipMemsetAsync(u->dacc, 0, sizeof(double), s));
  hipLaunchKernelGGL(k_scales, dim3(512), dim3(256), 0, s, x->d, x->d_R, x->d_u0, x->d_v0, u->dacc, u->sc, 0);
  hipLaunchKernelGGL(k_scales, dim3(1), dim3(64), 0, s, x->d, x->d_R, x->d_u0, x->d_v0, u->dacc, u->sc, 1);
  const long long V = (long long)u->dx * u->dy * u->dz;
  hipLaunchKernelGGL(k_features, g1(V), dim3(256), 0, s, x->d, x->d_flag, x->d_R, x->d_u0, x->d_v0, u->sc, u->ncx,
                     u->ncy, u->ncz, u->dx, u->dy, u->dz, u->feat);
  UCHK(hipGetLastError());
  return QESMC_OK;
}

}  // namespace

extern "C" {

int qesmc_unet_create(qesmc_unet **out, int device, const qesmc_unet_config *cfg)
{
  if (!out || !cfg) return QESMC_ERR_ARG;
  *out = nullptr;
  if (cfg->levels < 2 || cfg->levels > 5 || cfg->base_channels < 16 || cfg->base_channels > 32 ||
      cfg->base_channels % 16 ||
      cfg->base_channels * (1 << (cfg->levels - 1)) > 128 || cfg->ncx < 1 || cfg->ncy < 1 || cfg->ncz < 1)
    return QESMC_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return QESMC_ERR_NO_DEVICE;
  UCHK(hipSetDevice(device));
  qesmc_unet *u = new (std::nothrow) qesmc_unet();
  if (!u) return QESMC_ERR_OOM;
  u->device = device;
  u->L = cfg->levels;
  u->c0 = cfg->base_channels;
  u->ncx = cfg->ncx;
  u->ncy = cfg->ncy;
  u->ncz = cfg->ncz;
  u->tscale = cfg->target_scale > 0.f ? cfg->target_scale : 0.f;
  const int mult = 1 << (u->L - 1);
  u->dx = round_up(cfg->ncx, mult);
  u->dy = round_up(cfg->ncy, mult);
  u->dz = round_up(cfg->ncz, mult);
  // layers and parameter offsets
  long long off = 0;
  auto mk = [&](int cin, int cout) {
    Conv c;
    c.cin = cin;
    c.cout = cout;
    c.kp = round_up(27 * cin, 32);
    c.kpd = round_up(27 * cout, 32);
    c.woff = off;
    off += 27LL * cin * cout;
    c.boff = off;
    off += cout;
    return c;
  };
  u->lv.resize(u->L);
  for (int l = 0; l < u->L; ++l) {
    const int c = u->c0 << l;
    const int cin = l == 0 ? QESMC_UNET_IN_CHANNELS : (u->c0 << (l - 1));
    u->enc1.push_back(mk(cin, c));
    u->enc2.push_back(mk(c, c));
  }
  u->dec1.resize(u->L - 1);
  u->dec2.resize(u->L - 1);
  for (int l = u->L - 2; l >= 0; --l) {
    const int c = u->c0 << l;
    u->dec1[l] = mk(c + (c << 1), c);
    u->dec2[l] = mk(c, c);
  }
  u->head_off = off;
  off += u->c0 + 1;
  u->nparams = off;

  int rc = QESMC_OK;
  auto fail = [&](hipError_t e) {
    if (e != hipSuccess) rc = map_err(e);
    return e != hipSuccess;
  };
  if (fail(hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking)) || fail(hipEventCreate(&u->ev[0])) ||
      fail(hipEventCreate(&u->ev[1]))) {
    qesmc_unet_destroy(u);
    return rc;
  }
  bool bad = fail(dmalloc(&u->p, off)) || fail(dmalloc(&u->gp, off)) || fail(dmalloc(&u->m1, off)) ||
             fail(dmalloc(&u->m2, off));
  for (auto *v : {&u->enc1, &u->enc2, &u->dec1, &u->dec2})
    for (auto &c : *v)
      bad = bad || fail(dmalloc(&c.wt, (long long)c.cout * c.kp)) || fail(dmalloc(&c.wd, (long long)c.cin * c.kpd));
  double flops = 0.0;
  for (int l = 0; l < u->L && !bad; ++l) {
    Level &lv = u->lv[l];
    lv.dx = u->dx >> l;
    lv.dy = u->dy >> l;
    lv.dz = u->dz >> l;
    lv.V = (long long)lv.dx * lv.dy * lv.dz;
    lv.c = u->c0 << l;
    lv.ldcat = (l + 1 < u->L) ? lv.c + (lv.c << 1) : lv.c;
    const long long V = lv.V;
    bad = fail(dmalloc(&lv.A, V * lv.c)) || fail(dmalloc(&lv.CAT, V * lv.ldcat)) || fail(dmalloc(&lv.E, V * lv.c)) ||
          fail(dmalloc(&lv.D, V * lv.c)) || fail(dmalloc(&lv.gA, V * lv.c)) || fail(dmalloc(&lv.gCAT, V * lv.ldcat)) ||
          fail(dmalloc(&lv.gE, V * lv.c)) || fail(dmalloc(&lv.gD, V * lv.c));
    if (!bad && l > 0) {
      const int cp = u->c0 << (l - 1);
      bad = fail(dmalloc(&lv.P, V * cp)) || fail(dmalloc(&lv.gP, V * cp));
    }
    flops += 2.0 * V * 27.0 * (u->enc1[l].cin * (double)u->enc1[l].cout + u->enc2[l].cin * (double)u->enc2[l].cout);
    if (l + 1 < u->L)
      flops += 2.0 * V * 27.0 * (u->dec1[l].cin * (double)u->dec1[l].cout + u->dec2[l].cin * (double)u->dec2[l].cout);
  }
  u->conv_flops = flops;
  const long long V0 = (long long)u->dx * u->dy * u->dz;
  bad = bad || fail(dmalloc(&u->feat, V0 * QESMC_UNET_IN_CHANNELS)) || fail(dmalloc(&u->out, V0)) ||
        fail(dmalloc(&u->gout, V0)) || fail(dmalloc(&u->tgt, V0)) ||
        fail(dmalloc(&u->hostbuf_dev, V0 * QESMC_UNET_IN_CHANNELS)) || fail(dmalloc(&u->sc, 2)) ||
        fail(dmalloc(&u->dacc, 1)) || fail(dmalloc(&u->loss, 1));
  if (bad) {
    qesmc_unet_destroy(u);
    return rc == QESMC_OK ? QESMC_ERR_OOM : rc;
  }
  // He-normal initialisation of the conv weights, zero biases (host RNG)
  std::vector<float> hp((size_t)off, 0.f);
  std::mt19937_64 rng(cfg->seed);
  auto init = [&](const Conv &c) {
    std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (27.0 * c.cin)));
    for (long long q = 0; q < 27LL * c.cin * c.cout; ++q) hp[(size_t)(c.woff + q)] = (float)nd(rng);
  };
  for (int l = 0; l < u->L; ++l) {
    init(u->enc1[l]);
    init(u->enc2[l]);
  }
  for (int l = u->L - 2; l >= 0; --l) {
    init(u->dec1[l]);
    init(u->dec2[l]);
  }
  std::normal_distribution<double> nh(0.0, std::sqrt(1.0 / u->c0));
  for (int c = 0; c < u->c0; ++c) hp[(size_t)(u->head_off + c)] = (float)nh(rng);
  rc = qesmc_unet_set_params(u, hp.data(), off);
  if (rc != QESMC_OK) {
    qesmc_unet_destroy(u);
    return rc;
  }
  *out = u;
  return QESMC_OK;
}

int qesmc_unet_destroy(qesmc_unet *u)
{
  if (!u) return QESMC_OK;
  (void)hipSetDevice(u->device);
  if (u->stream) (void)hipStreamSynchronize(u->stream);
  free_all(u);
  for (auto &e : u->ev)
    if (e) (void)hipEventDestroy(e);
  if (u->stream) (void)hipStreamDestroy(u->stream);
  delete u;
  return QESMC_OK;
}

int qesmc_unet_num_params(const qesmc_unet *u, long long *count)
{
  if (!u || !count) return QESMC_ERR_ARG;
  *count = u->nparams;
  return QESMC_OK;
}

int qesmc_unet_dims(const qesmc_unet *u, int *dx, int *dy, int *dz)
{
  if (!u || !dx || !dy || !dz) return QESMC_ERR_ARG;
  *dx = u->dx;
  *dy = u->dy;
  *dz = u->dz;
  return QESMC_OK;
}

int qesmc_unet_get_params(qesmc_unet *u, float *host, long long count)
{
  if (!u || !host || count != u->nparams) return QESMC_ERR_ARG;
  UCHK(hipSetDevice(u->device));
  UCHK(hipMemcpyAsync(host, u->p, sizeof(float) * count, hipMemcpyDeviceToHost, u->stream));
  UCHK(hipStreamSynchronize(u->stream));
  return QESMC_OK;
}

int qesmc_unet_set_params(qesmc_unet *u, const float *host, long long count)
{
  if (!u || !host || count != u->nparams) return QESMC_ERR_ARG;
  UCHK(hipSetDevice(u->device));
  UCHK(hipMemcpyAsync(u->p, host, sizeof(float) * count, hipMemcpyHostToDevice, u->stream));
  UCHK(hipMemsetAsync(u->m1, 0, sizeof(float) * count, u->stream));
  UCHK(hipMemsetAsync(u->m2, 0, sizeof(float) * count, u->stream));
  u->adam_t = 0;
  UCHK(pack_weights(u));
  UCHK(hipStreamSynchronize(u->stream));
  return QESMC_OK;
}

int qesmc_unet_forward(qesmc_unet *u, const float *input, float *output)
{
  if (!u || !input || !output) return QESMC_ERR_ARG;
  UCHK(hipSetDevice(u->device));
  const long long V0 = (long long)u->dx * u->dy * u->dz;
  const long long n = V0 * QESMC_UNET_IN_CHANNELS;
  UCHK(hipMemcpyAsync(u->hostbuf_dev, input, sizeof(float) * n, hipMemcpyHostToDevice, u->stream));
  hipLaunchKernelGGL(k_f2bf, g1(n), dim3(256), 0, u->stream, u->hostbuf_dev, u->feat, n);
  UCHK(hipGetLastError());
  UCHK(forward(u, u->feat));
  UCHK(hipMemcpyAsync(output, u->out, sizeof(float) * V0, hipMemcpyDeviceToHost, u->stream));
  UCHK(hipStreamSynchronize(u->stream));
  return QESMC_OK;
}

int qesmc_unet_backward(qesmc_unet *u, const float *grad_output, float *grad_params)
{
  if (!u || !grad_output || !grad_params) return QESMC_ERR_ARG;
  UCHK(hipSetDevice(u->device));
  const long long V0 = (long long)u->dx * u->dy * u->dz;
  UCHK(hipMemcpyAsync(u->gout, grad_output, sizeof(float) * V0, hipMemcpyHostToDevice, u->stream));
  UCHK(hipMemsetAsync(u->gp, 0, sizeof(float) * u->nparams, u->stream));
  UCHK(backward(u, u->feat));
  UCHK(hipMemcpyAsync(grad_params, u->gp, sizeof(float) * u->nparams, hipMemcpyDeviceToHost, u->stream));
  UCHK(hipStreamSynchronize(u->stream));
  return QESMC_OK;
}

int qesmc_unet_add_sample(qesmc_unet *u, qesmc_ctx *x)
{
  if (!u || !x) return QESMC_ERR_ARG;
  UCHK(hipSetDevice(u->device));
  int rc = problem_features(u, x);
  if (rc != QESMC_OK) return rc;
  const long long V0 = (long long)u->dx * u->dy * u->dz;
  bf16_t *f = nullptr;
  float *t = nullptr;
  if (hipMalloc((void **)&f, sizeof(bf16_t) * V0 * QESMC_UNET_IN_CHANNELS) != hipSuccess) return QESMC_ERR_OOM;
  if (hipMalloc((void **)&t, sizeof(float) * V0) != hipSuccess) {
    (void)hipFree(f);
    return QESMC_ERR_OOM;
  }
  UCHK(hipMemcpyAsync(f, u->feat, sizeof(bf16_t) * V0 * QESMC_UNET_IN_CHANNELS, hipMemcpyDeviceToDevice, u->stream));
  hipLaunchKernelGGL(k_target, g1(V0), dim3(256), 0, u->stream, x->d, x->d_lam, u->sc, u->ncx, u->ncy, u->ncz, u->dx,
                     u->dy, u->dz, t);
  UCHK(hipGetLastError());
  UCHK(hipStreamSynchronize(u->stream));
  u->sfeat.push_back(f);
  u->stgt.push_back(t);
  return QESMC_OK;
}

int qesmc_unet_num_samples(const qesmc_unet *u, int *count)
{
  if (!u || !count) return QESMC_ERR_ARG;
  *count = (int)u->sfeat.size();
  return QESMC_OK;
}

int qesmc_unet_clear_samples(qesmc_unet *u)
{
  if (!u) return QESMC_ERR_ARG;
  (void)hipSetDevice(u->device);
  (void)hipStreamSynchronize(u->stream);
  for (auto &q : u->sfeat) (void)hipFree(q);
  for (auto &q : u->stgt) (void)hipFree(q);
  u->sfeat.clear();
  u->stgt.clear();
  return QESMC_OK;
}

int qesmc_unet_train(qesmc_unet *u, int steps, float lr, int first, int last, unsigned long long seed,
                     float *mean_loss, double *ms)
{
  if (!u || steps < 0 || first < 0 || last > (int)u->sfeat.size() || first >= last) return QESMC_ERR_ARG;
  UCHK(hipSetDevice(u->device));
  hipStream_t s = u->stream;
  const long long V0 = (long long)u->dx * u->dy * u->dz;
  if (!(u->tscale > 0.f)) {
    // output unit from the data: RMS of the stored targets lambda / rms(R)
    UCHK(hipMemsetAsync(u->dacc, 0, sizeof(double), s));
    for (size_t q = 0; q < u->stgt.size(); ++q)
      hipLaunchKernelGGL(k_sumsq, dim3(256), dim3(256), 0, s, u->stgt[q], V0, u->dacc);
    double ss = 0.0;
    UCHK(hipMemcpyAsync(&ss, u->dacc, sizeof(double), hipMemcpyDeviceToHost, s));
    UCHK(hipStreamSynchronize(s));
    const double rms = std::sqrt(ss / ((double)V0 * u->stgt.size()));
    u->tscale = rms > 0.0 ? (float)rms : 1.f;
  }
  std::mt19937_64 rng(seed);
  std::uniform_int_distribution<int> pick(first, last - 1);
  UCHK(hipMemsetAsync(u->loss, 0, sizeof(float), s));
  UCHK(hipEventRecord(u->ev[0], s));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int it = 0; it < steps; ++it) {
    const int q = pick(rng);
    UCHK(forward(u, u->sfeat[q]));
    UCHK(launch_loss(u, u->stgt[q], true));
    UCHK(hipMemsetAsync(u->gp, 0, sizeof(float) * u->nparams, s));
    UCHK(backward(u, u->sfeat[q]));
    ++u->adam_t;
    const float c1 = 1.f - std::pow(b1, (float)u->adam_t), c2 = 1.f - std::pow(b2, (float)u->adam_t);
    hipLaunchKernelGGL(k_adam, g1(u->nparams), dim3(256), 0, s, u->p, u->gp, u->m1, u->m2, u->nparams, lr, b1, b2,
                       eps, c1, c2);
    UCHK(pack_weights(u));
  }
  UCHK(hipEventRecord(u->ev[1], s));
  float l = 0.f;
  UCHK(hipMemcpyAsync(&l, u->loss, sizeof(float), hipMemcpyDeviceToHost, s));
  UCHK(hipStreamSynchronize(s));
  if (mean_loss) *mean_loss = steps > 0 ? l / steps : 0.f;
  if (ms) {
    float t = 0.f;
    (void)hipEventElapsedTime(&t, u->ev[0], u->ev[1]);
    *ms = t;
  }
  return QESMC_OK;
}

int qesmc_unet_set_loss(qesmc_unet *u, int block, float weight)
{
  if (!u || block < 0 || !(weight >= 0.f)) return QESMC_ERR_ARG;
  if (block > 0 && ((block & (block - 1)) || u->dx % block || u->dy % block || u->dz % block)) return QESMC_ERR_ARG;
  UCHK(hipSetDevice(u->device));
  if (u->eblk) {
    UCHK(hipStreamSynchronize(u->stream));
    UCHK(hipFree(u->eblk));
    u->eblk = nullptr;
  }
  u->loss_block = weight > 0.f ? block : 0;
  u->loss_weight = u->loss_block ? weight : 0.f;
  if (u->loss_block) {
    const long long nb = (long long)(u->dx / block) * (u->dy / block) * (u->dz / block);
    UCHK(hipMalloc((void **)&u->eblk, sizeof(float) * nb));
  }
  return QESMC_OK;
}

int qesmc_unet_target_scale(const qesmc_unet *u, float *scale)
{
  if (!u || !scale) return QESMC_ERR_ARG;
  *scale = u->tscale;
  return QESMC_OK;
}

int qesmc_unet_eval(qesmc_unet *u, int first, int last, float *mean_loss)
{
  if (!u || !mean_loss || first < 0 || last > (int)u->sfeat.size() || first >= last) return QESMC_ERR_ARG;
  if (!(u->tscale > 0.f)) return QESMC_ERR_STATE;  // no output unit yet: train or configure it
  UCHK(hipSetDevice(u->device));
  hipStream_t s = u->stream;
  UCHK(hipMemsetAsync(u->loss, 0, sizeof(float), s));
  for (int q = first; q < last; ++q) {
    UCHK(forward(u, u->sfeat[q]));
    UCHK(launch_loss(u, u->stgt[q], false));
  }
  float l = 0.f;
  UCHK(hipMemcpyAsync(&l, u->loss, sizeof(float), hipMemcpyDeviceToHost, s));
  UCHK(hipStreamSynchronize(s));
  *mean_loss = l / (last - first);
  return QESMC_OK;
}

int qesmc_unet_predict(qesmc_unet *u, qesmc_ctx *x, double *ms, double *tflops)
{
  if (!u || !x) return QESMC_ERR_ARG;
  if (!(u->tscale > 0.f)) return QESMC_ERR_STATE;
  UCHK(hipSetDevice(u->device));
  UCHK(hipStreamSynchronize(x->stream));
  const auto t0 = std::chrono::steady_clock::now();
  int rc = problem_features(u, x);
  if (rc != QESMC_OK) return rc;
  hipStream_t s = u->stream;
  UCHK(hipEventRecord(u->ev[0], s));
  UCHK(forward(u, u->feat));
  UCHK(hipEventRecord(u->ev[1], s));
  hipLaunchKernelGGL(k_install, g1(x->ncell), dim3(256), 0, s, x->d, u->out, u->sc, u->tscale, u->dx, u->dy,
                     x->d_lam);
  UCHK(hipGetLastError());
  UCHK(hipStreamSynchronize(s));
  const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  float t = 0.f;
  (void)hipEventElapsedTime(&t, u->ev[0], u->ev[1]);
  if (ms) *ms = wall;
  if (tflops) *tflops = t > 0.f ? u->conv_flops / (t * 1e-3) / 1e12 : 0.0;
  return QESMC_OK;
}

}  // extern "C"
