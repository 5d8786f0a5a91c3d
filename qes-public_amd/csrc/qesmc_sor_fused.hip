// Red-black SOR with both colour passes of an iteration in one launch, on
// the colour-split layout with the wall-mask stencil.
//
// Iteration t reads lambda from buffer set A = t & 1 and writes buffer set
// B = A ^ 1 (each set holds both colours), so no work-group ever reads a
// value another work-group of the same launch writes.  The result is the
// reference's in-place schedule (src/winds/Solver_CPU_RB.cpp:109-162)
// exactly: a red cell sees the old black values, a black cell the new red
// ones, and after the iteration lambda(i,j,0) = lambda(i,j,1) — written here
// straight into B's plane 0 when a k = 1 cell is relaxed.  Cells that are
// never updated (the domain's ring, planes 0 and NZ-1) hold the same values
// in both sets.
//
// Work decomposition.  A work-group owns a tile of 124 half-cells (q) x 8
// rows (j) x kb planes of the black colour and of the red colour; each lane
// holds two neighbouring half-cells (one vector load per array).  Marching
// up the planes, it first relaxes the red cells of plane p over the tile
// plus a halo (128 x 10: two half-cells and one row each side, recomputed
// from A exactly as their owner does) into a 3-plane LDS ring, then the
// black cells of plane p-1, whose six red neighbours are all in the ring.
// Per iteration a cell costs its own lambda read and write, R, the mask and
// one read of the other colour (as a neighbour of a red cell): 13 bytes,
// against 17 per colour pass when each pass is its own launch.  Tiles are numbered q-fastest, then
// j, then plane chunk, and XCD x (work-group id % 8) takes the x-th eighth
// of them, so the halo rows and columns a work-group re-reads sit in the L2
// of the work-groups that own them.
#include "qesmc_sor_tile.h"

namespace qesmc {

namespace {

constexpr int kXcds = 8;
constexpr int kTq = 128;         // half-cells of a tile row relaxed in red (two per lane)
constexpr int kTo = kTq - 4;     // of which the tile owns the middle ones (lanes 1..62)
constexpr int kTj = 8;           // rows per tile (two per wave)
constexpr int kRw = kTq;         // ring row length
constexpr int kRh = kTj + 2;     // ring rows (halo row each side)

// Work-group barrier that orders LDS only: __syncthreads() also fences
// global memory, which makes every wave wait for all of its outstanding
// global loads and stores at each barrier.
__device__ __forceinline__ void lds_barrier()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ inline float wave_max(float v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Inputs of two neighbouring red cells of one row (half-cells q, q+1 of a
// lane): one vector load per array; the x-neighbour beyond the pair (q-1
// for s = 0, q+2 for s = 1) is one extra scalar load.
struct RedPair {
  float o[2], r[2], c[2], yp[2], ym[2], zp[2], zm[2], ext;
  unsigned m;
  long long x;
  int s;
  bool valid, rowupd, upd[2];
};
// Inputs of two black cells of one row.
struct BlackPair {
  float o[2], r[2];
  unsigned m;
  long long x;
  int s;
  bool valid, upd[2];
};
// Everything one plane step reads from global memory: this lane's red pairs
// of plane p (three in waves 0 and 1, two in waves 2 and 3) and its two black
// pairs of plane p-1, with the two planes' wall values.
struct StepIn {
  float cvr[6], cvb[6];
  RedPair pr[3];
  BlackPair bp[2];
};

__device__ __forceinline__ float2 ld2(const float *__restrict__ p, long long x)
{
  return *reinterpret_cast<const float2 *>(p + x);
}

__device__ __forceinline__ unsigned ld_mask2(const unsigned char *__restrict__ p, long long x)
{
  return *reinterpret_cast<const unsigned short *>(p + x);
}

// Loads of plane step p, addresses clamped in bounds (to index 0 where a
// value is not needed) and no branches but the wave-uniform one, so that
// all of them are in flight together.  q0 is even and HP a multiple of 4:
// a pair never straddles a row and its vector loads are aligned.  Lane L
// holds half-cells q0 + 2L and q0 + 2L + 1 of its rows.
__device__ __forceinline__ StepIn load_step(const Dims &d, int p, int k0, int j0, int q0, int wave, int lane,
                                           const float *__restrict__ a0, const float *__restrict__ a1,
                                           const float *__restrict__ R0, const float *__restrict__ R1,
                                           const unsigned char *__restrict__ mask0,
                                           const unsigned char *__restrict__ mask1,
                                           const unsigned *__restrict__ canon)
{
  StepIn L;
  const long long HP = d.HP, sp = d.splane;
  // in the same batch as the cell loads: a wait for a later load would also
  // wait for every earlier one (the counter is in order)
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    L.cvr[a] = __uint_as_float(canon[a * d.NZ + p]);
    L.cvb[a] = __uint_as_float(canon[a * d.NZ + (p >= 1 ? p - 1 : 0)]);
  }
  const int npairs = wave < 2 ? 3 : 2;  // ring rows wave, wave + 4, wave + 8 (< 10)
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    RedPair &P = L.pr[r];
    if (r >= npairs) {  // wave-uniform
      P.valid = P.rowupd = P.upd[0] = P.upd[1] = false;
      continue;
    }
    const int j = j0 - 1 + wave + 4 * r;
    const int q = q0 + 2 * lane;
    P.s = (j + p) & 1;
    P.valid = j < d.NY && q >= 0 && q < d.HP;
    P.rowupd = P.valid && j >= 1 && j <= d.NY - 2 && p >= 1 && p <= d.NZ - 2;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int i = 2 * (q + e) + P.s;
      P.upd[e] = P.rowupd && i >= 1 && i <= d.NX - 2;
    }
    const long long x = P.valid ? ((long long)p * d.NY + j) * HP + q : 0;
    const long long xn = P.rowupd ? x : 0;
    P.x = x;
    const float2 o = ld2(a0, x), rv = ld2(R0, x), c = ld2(a1, xn), yp = ld2(a1, P.rowupd ? x + HP : 0),
                 ym = ld2(a1, P.rowupd ? x - HP : 0), zp = ld2(a1, P.rowupd ? x + sp : 0),
                 zm = ld2(a1, P.rowupd ? x - sp : 0);
    P.m = ld_mask2(mask0, x);
    const bool need = P.s ? P.upd[1] : P.upd[0];
    P.ext = a1[need ? (P.s ? x + 2 : x - 1) : 0];
    P.o[0] = o.x, P.o[1] = o.y, P.r[0] = rv.x, P.r[1] = rv.y, P.c[0] = c.x, P.c[1] = c.y;
    P.yp[0] = yp.x, P.yp[1] = yp.y, P.ym[0] = ym.x, P.ym[1] = ym.y;
    P.zp[0] = zp.x, P.zp[1] = zp.y, P.zm[0] = zm.x, P.zm[1] = zm.y;
  }
  const int pb = p - 1;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    BlackPair &B = L.bp[a];
    const int j = j0 + 2 * wave + a;
    const int q = q0 + 2 * lane;
    B.s = (j + pb + 1) & 1;
    B.valid = lane >= 1 && lane <= 62 && pb >= k0 && j <= d.NY - 2 && q >= 0 && q < d.HP;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int i = 2 * (q + e) + B.s;
      B.upd[e] = B.valid && i >= 1 && i <= d.NX - 2;
    }
    const long long x = B.valid ? ((long long)pb * d.NY + j) * HP + q : 0;
    B.x = x;
    const float2 o = ld2(a1, x), rv = ld2(R1, x);
    B.m = ld_mask2(mask1, x);
    B.o[0] = o.x, B.o[1] = o.y, B.r[0] = rv.x, B.r[1] = rv.y;
  }
  return L;
}

__device__ __forceinline__ float relax_masked(unsigned m, const float *cv, float xp, float xm, float yp, float ym,
                                              float zp, float zm, float r, float o, float omega, float one_m_omega)
{
  return relax((m & 1u) ? cv[0] : 0.f, (m & 2u) ? cv[1] : 0.f, (m & 4u) ? cv[2] : 0.f, (m & 8u) ? cv[3] : 0.f,
               (m & 16u) ? cv[4] : 0.f, (m & 32u) ? cv[5] : 0.f, xp, xm, yp, ym, zp, zm, r, o, omega, one_m_omega);
}

// Plane step p: red cells of plane p into the ring (and to B where this
// tile owns them), then black cells of plane p-1 from the ring.
__device__ __forceinline__ void compute_step(const Dims &d, int p, int k0, int k1, int wave, int lane,
                                             const StepIn &L, float (*ring)[kRh][kRw], float *__restrict__ b0,
                                             float *__restrict__ b1, float omega, float one_m_omega, float &dl,
                                             bool &bad)
{
  const long long sp = d.splane;
  const int pb = p - 1;  // black plane relaxed in this step
  const int cc = 2 * lane;  // ring column of the lane's first half-cell
  // ---- red plane p into the ring
  {
    float(*rp)[kRw] = ring[p % 3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const RedPair &P = L.pr[r];
      const int rr = wave + 4 * r;
      if (rr >= kRh) break;  // wave-uniform
      float val[2], nv[2];
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const float xm = P.s ? P.c[e] : (e ? P.c[0] : P.ext);
        const float xp = P.s ? (e ? P.ext : P.c[1]) : P.c[e];
        nv[e] = relax_masked((P.m >> (8 * e)) & 0xffu, L.cvr, xp, xm, P.yp[e], P.ym[e], P.zp[e], P.zm[e], P.r[e],
                             P.o[e], omega, one_m_omega);
        val[e] = P.upd[e] ? nv[e] : P.valid ? P.o[e] : 0.f;
      }
      rp[rr][cc] = val[0];
      rp[rr][cc + 1] = val[1];
      if (P.rowupd && lane >= 1 && lane <= 62 && rr >= 1 && rr <= kTj && p >= k0 && p < k1) {
        *reinterpret_cast<float2 *>(b0 + P.x) = make_float2(val[0], val[1]);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          if (!P.upd[e]) continue;
          if (p == 1) b1[P.x - sp + e] = nv[e];  // lambda(i,j,0) = lambda(i,j,1), black plane 0
          if (!isfinite(nv[e])) bad = true;
          const float dd = fabsf(nv[e] - P.o[e]);
          dl = (dd > dl) ? dd : dl;  // NaN never wins, as in the reference's max test
        }
      }
    }
  }
  lds_barrier();
  // ---- black plane p-1 from the ring (planes p-2, p-1, p)
  if (pb >= k0) {
    const float(*r0)[kRw] = ring[pb % 3];
    const float(*rz)[kRw] = ring[(pb + 1) % 3];
    const float(*rm)[kRw] = ring[(pb + 2) % 3];  // plane pb-1
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const BlackPair &B = L.bp[a];
      if (!B.valid) continue;
      const int rr = 1 + 2 * wave + a;
      float val[2];
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int c = cc + e;
        const float nv = relax_masked((B.m >> (8 * e)) & 0xffu, L.cvb, r0[rr][c + B.s], r0[rr][c - 1 + B.s],
                                      r0[rr + 1][c], r0[rr - 1][c], rz[rr][c], rm[rr][c], B.r[e], B.o[e], omega,
                                      one_m_omega);
        val[e] = B.upd[e] ? nv : B.o[e];
        if (B.upd[e]) {
          if (pb == 1) b0[B.x - sp + e] = nv;  // lambda(i,j,0) = lambda(i,j,1), red plane 0
          if (!isfinite(nv)) bad = true;
          const float dd = fabsf(nv - B.o[e]);
          dl = (dd > dl) ? dd : dl;
        }
      }
      *reinterpret_cast<float2 *>(b1 + B.x) = make_float2(val[0], val[1]);
    }
  }
  lds_barrier();  // the next red plane overwrites ring slot (p + 1) % 3 = (pb - 1) % 3
}

}  // namespace

__global__ __launch_bounds__(256) void k_sor_fused(Dims d, int kb, int qt, int jt, long long ntiles,
                                                   const float *__restrict__ a0, const float *__restrict__ a1,
                                                   float *__restrict__ b0, float *__restrict__ b1,
                                                   const float *__restrict__ R0, const float *__restrict__ R1,
                                                   const unsigned char *__restrict__ mask0,
                                                   const unsigned char *__restrict__ mask1,
                                                   const unsigned *__restrict__ canon, float omega,
                                                   float one_m_omega, float tol, unsigned *err, int t,
                                                   int itermax, unsigned *status, unsigned *cur)
{
  __shared__ float ring[3][kRh][kRw];
  __shared__ float wmax[4];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (t >= itermax) return;
  {
    // gate: iteration t runs iff max(err row t) > tol (row 0 = 1.0f)
    const float g = wave_max(__uint_as_float(err[(long long)t * kErrSlots + lane]));
    if (!gate_pass(g, t, tol, status)) return;  // identical in every wave of the grid
  }
  const long long per = (long long)gridDim.x / kXcds;  // the grid is a multiple of 8
  const long long tile = (long long)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  if (tile >= ntiles) return;  // whole work-group, before any barrier
  if (tile == 0 && threadIdx.x == 0) *cur = (unsigned)((t + 1) & 1);
  const int tq = (int)(tile % qt);
  const int tj = (int)((tile / qt) % jt);
  const int kc = (int)(tile / ((long long)qt * jt));
  const int q0 = tq * kTo - 2;  // owns half-cells [q0 + 2, q0 + 126)
  const int j0 = 1 + tj * kTj;
  const int k0 = 1 + kc * kb;
  const int k1 = k0 + kb < d.NZ - 1 ? k0 + kb : d.NZ - 1;  // black planes [k0, k1)

  float dl = 0.f;
  bool bad = false;
  const int pbeg = k0 - 1;  // >= 0
  const int pend = k1;      // <= NZ-1

  // One plane step at a time.  Prefetching step p + 1 during step p (two
  // load buffers) was measured slower: the extra registers cut occupancy,
  // and this kernel needs the waves more.
#pragma unroll 1
  for (int p = pbeg; p <= pend; ++p) {
    const StepIn L = load_step(d, p, k0, j0, q0, wave, lane, a0, a1, R0, R1, mask0, mask1, canon);
    compute_step(d, p, k0, k1, wave, lane, L, ring, b0, b1, omega, one_m_omega, dl, bad);
  }
  if (bad) atomicOr(status, ST_NONFINITE);
  const float wm = wave_max(dl);
  if (lane == 0) wmax[wave] = wm;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float bm = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
    if (bm > 0.f) atomicMax(err + (long long)(t + 1) * kErrSlots + (blockIdx.x % kErrSlots), __float_as_uint(bm));
  }
}

}  // namespace qesmc

using namespace qesmc;

// Launch iteration t of the fused path: reads buffer set t & 1, writes the other.
hipError_t qesmc_launch_fused_iteration(qesmc_ctx *x, int t, int itermax, float omega, float tol)
{
  const Dims &d = x->d;
  if (d.NZ - 2 < 1 || d.NY - 2 < 1) return hipSuccess;
  const int kb = x->kplanes < 1 ? 1 : x->kplanes;
  const int qt = (d.HP + kTo - 1) / kTo;
  const int jt = (d.NY - 2 + kTj - 1) / kTj;
  const int nkc = (d.NZ - 2 + kb - 1) / kb;
  const long long ntiles = (long long)qt * jt * nkc;
  const long long blocks = (ntiles + kXcds - 1) / kXcds * kXcds;
  float *const *A = (t & 1) ? x->d_lc2 : x->d_lc;
  float *const *B = (t & 1) ? x->d_lc : x->d_lc2;
  hipLaunchKernelGGL(k_sor_fused, dim3((unsigned)blocks), dim3(256), 0, x->stream, d, kb, qt, jt, ntiles, A[0],
                     A[1], B[0], B[1], x->d_Rc[0], x->d_Rc[1], x->d_mask[0], x->d_mask[1], x->d_canon, omega,
                     1.0f - omega, tol, x->d_err, t, itermax, x->d_status, x->d_cur);
  return hipGetLastError();
}
