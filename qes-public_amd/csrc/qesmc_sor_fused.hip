// Red-black SOR with both colour passes of an iteration in one launch, on
// the colour-split layout with the wall-mask stencil (QESMC_PATH_FUSED).
//
// Iteration t reads lambda from buffer set A = t & 1 and writes buffer set
// B = A ^ 1 (each set holds both colours), so no wave ever reads a value
// another wave of the same launch writes.  The result is the reference's
// in-place schedule (src/winds/Solver_CPU_RB.cpp:109-162) exactly: a red
// cell sees the old black values, a black cell the new red ones, and after
// the iteration lambda(i,j,0) = lambda(i,j,1) — written here straight into
// B's plane 0 when a k = 1 cell is relaxed.  Cells that are never updated
// (the domain's ring, planes 0 and NZ-1) hold the same values in both sets.
//
// Work decomposition: one wave per column tile, marching along j.  A wave
// owns W <= 62 consecutive half-cells q of a row (lanes 1..W; lanes 0 and
// W+1 hold the half-cells just outside, whose red values it recomputes as
// their owner does), the K planes [k0, k0+K) and the rows [j0, j1) of a
// segment.  Step j (for j = j0-1 .. j1) relaxes the red cells of row j over
// planes k0-1 .. k0+K (a halo plane each side) from the black values of A
// in rows j-1..j+1, then the black cells of row j-1 over the wave's own
// planes, whose six red neighbours (rows j-2..j, planes k-1..k+1, and the
// x-neighbours in the adjacent lanes) are all in registers.  Each black row
// of A is loaded once, as row j+1, and kept while the march passes it.
// There is no LDS and no barrier: the halos are recomputed, and the tiles
// that share them (neighbouring k-chunks first) are dispatched together so
// that the re-reads mostly hit in L2.  Per iteration a cell costs its own
// lambda read and write, R and the mask: 13 bytes, against 2 x 17 per cell
// pair for two split colour passes, plus the halo re-reads (K+4 black planes
// loaded per K relaxed, W+2 lanes per W, L+2 red rows per L).  As built it
// is bound by instruction issue, not by bytes: the halo relaxations, lane
// shuffles and the register rotation of the march about double the split
// kernel's arithmetic per update (loading the next step during the current
// one changed nothing), so the split kernel stays the default.
#include "qesmc_sor_tile.h"

namespace qesmc {

namespace {

constexpr int kXcds = 8;  // MI355X: 8 XCDs, work-groups dispatched round-robin
constexpr int kWaves = 4; // waves (tiles) per work-group

__device__ inline float wave_max(float v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ float relax_masked(unsigned m, const float *cv, float xp, float xm, float yp, float ym,
                                              float zp, float zm, float r, float o, float omega, float one_m_omega)
{
  return relax((m & 1u) ? cv[0] : 0.f, (m & 2u) ? cv[1] : 0.f, (m & 4u) ? cv[2] : 0.f, (m & 8u) ? cv[3] : 0.f,
               (m & 16u) ? cv[4] : 0.f, (m & 32u) ? cv[5] : 0.f, xp, xm, yp, ym, zp, zm, r, o, omega, one_m_omega);
}

}  // namespace

template <int K>
__global__ __launch_bounds__(256) void k_sor_march(Dims d, int W, int qt, int L, int nkc, long long ntiles,
                                                   const float *__restrict__ a0, const float *__restrict__ a1,
                                                   float *__restrict__ b0, float *__restrict__ b1,
                                                   const float *__restrict__ R0, const float *__restrict__ R1,
                                                   const unsigned char *__restrict__ mask0,
                                                   const unsigned char *__restrict__ mask1,
                                                   const unsigned *__restrict__ canon8, float omega,
                                                   float one_m_omega, float tol, unsigned *err, int t,
                                                   int itermax, unsigned *status, unsigned *cur)
{
  constexpr int NB = K + 4;  // black-old plane slots: planes k0-2 .. k0+K+1
  constexpr int NR = K + 2;  // red plane slots: planes k0-1 .. k0+K
  __shared__ float wmax[kWaves];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (t >= itermax) return;
  {
    // gate: iteration t runs iff max(err row t) > tol (row 0 = 1.0f)
    const float g = wave_max(__uint_as_float(err[(long long)t * kErrSlots + lane]));
    if (!gate_pass(g, t, tol, status)) return;  // identical in every wave of the grid
  }
  // XCD x (work-group id % 8) takes the x-th eighth of the work-group range;
  // tiles are numbered k-chunk fastest, so the 4 waves of a work-group take
  // neighbouring k-chunks of one column and share their halo planes
  const long long per = (long long)gridDim.x / kXcds;  // the grid is a multiple of 8
  const long long blk = (long long)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  if (blk == 0 && threadIdx.x == 0) *cur = (unsigned)((t + 1) & 1);
  const long long tile = blk * kWaves + wave;

  float dl = 0.f;
  bool bad = false;
  if (tile < ntiles) {  // wave-uniform, and no barrier inside
    const int tk = (int)(tile % nkc);
    const long long tr = tile / nkc;
    const int tq = (int)(tr % qt);
    const int tj = (int)(tr / qt);
    const int k0 = 1 + tk * K;
    const int k1 = k0 + K < d.NZ - 1 ? k0 + K : d.NZ - 1;  // owned planes [k0, k1)
    const int j0 = 1 + tj * L;
    const int j1 = j0 + L < d.NY - 1 ? j0 + L : d.NY - 1;  // owned rows [j0, j1)
    const int q = tq * W - 1 + lane;
    const bool qv = q >= 0 && q < d.HP;
    const unsigned qc = qv ? (unsigned)q : 0u;
    const bool own = qv && lane >= 1 && lane <= W;
    // the column i = 2q + s of this lane is updated (s: the cell's offset
    // within its half-cell pair, (j + k + colour) & 1)
    const bool col0 = qv && 2 * q >= 1 && 2 * q <= d.NX - 2;
    const bool col1 = qv && 2 * q + 1 <= d.NX - 2;
    // 32-bit element offsets (the launcher refuses arrays of 2^31 elements
    // or more).  Every address is clamped into the domain, so loads need no
    // predicate; values loaded for positions outside it are never used by
    // an updated cell.
    const unsigned HPu = (unsigned)d.HP, sp = (unsigned)d.splane;
    auto row_off = [&](int j) -> unsigned {
      const int jc = j < 0 ? 0 : j > d.NY - 1 ? d.NY - 1 : j;
      return (unsigned)jc * HPu + qc;
    };
    auto plane_off = [&](int k) -> unsigned {
      const int kc = k < 0 ? 0 : k > d.NZ - 1 ? d.NZ - 1 : k;
      return (unsigned)kc * sp;
    };

    // black values of A, rows j-1 (bm), j (bc), j+1 (bp), planes k0-2+p
    float bm[NB], bc[NB], bp[NB];
    auto load_black = [&](int j, float *dst) {
      const unsigned r = row_off(j);
#pragma unroll
      for (int p = 0; p < NB; ++p) dst[p] = a1[r + plane_off(k0 - 2 + p)];
    };
    load_black(j0 - 2, bm);
    load_black(j0 - 1, bc);
    // new red values, rows j-2 (rm2), j-1 (rm1), j (rn), planes k0-1+p
    float rm2[NR], rm1[NR], rn[NR];
#pragma unroll
    for (int p = 0; p < NR; ++p) rm2[p] = rm1[p] = rn[p] = 0.f;

#pragma unroll 1
    for (int j = j0 - 1; j <= j1; ++j) {
      load_black(j + 1, bp);
      // red inputs of row j (planes k0-1 .. k0+K); j0-1 >= 0 and j1 <= NY-1
      const unsigned rj = row_off(j);
      float ro[NR], rr[NR];
      unsigned rmk[NR];
#pragma unroll
      for (int p = 0; p < NR; ++p) {
        const unsigned x = rj + plane_off(k0 - 1 + p);
        ro[p] = a0[x];
        rr[p] = R0[x];
        rmk[p] = mask0[x];
      }
      // black inputs of row j-1 (owned planes)
      const int jb = j - 1;
      const bool brow = jb >= j0 && jb < j1;
      const unsigned rb = row_off(jb);
      float br[K];
      unsigned bmk[K];
#pragma unroll
      for (int p = 0; p < K; ++p) {
        const unsigned x = rb + plane_off(k0 + p);
        br[p] = R1[x];
        bmk[p] = mask1[x];
      }
      // ---- red cells of row j, plane by plane, each followed by the black
      // cell of row j-1 in the same plane (its red neighbours: rows j-2..j
      // of this plane, planes k-1 and k+1 of row j-1 from the last step)
      const bool rowu = j >= 1 && j <= d.NY - 2;
      const bool rown = j >= j0 && j < j1;
      // the wall values of a plane: scalar loads every step (cached),
      // rather than all planes' held in scalar registers for the whole march
      const unsigned *cg = canon8;
      asm volatile("" : "+s"(cg));
      const __attribute__((address_space(4))) unsigned *cn = (const __attribute__((address_space(4))) unsigned *)cg;
#pragma unroll
      for (int p = 0; p < NR; ++p) {
        const int k = k0 - 1 + p;
        const int kc = k < d.NZ ? k : d.NZ - 1;
        float cv[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) cv[a] = __uint_as_float(cn[kc * 8 + a]);
        {
          const int s = (j + k) & 1;
          const bool upd = (rowu && k >= 1 && k <= d.NZ - 2) && (s ? col1 : col0);
          const float c = bc[p + 1];
          const float cdn = __shfl_up(c, 1, 64);    // half-cell q-1
          const float cup = __shfl_down(c, 1, 64);  // half-cell q+1
          const float xm = s ? c : cdn;
          const float xp = s ? cup : c;
          const float nv = relax_masked(rmk[p], cv, xp, xm, bp[p + 1], bm[p + 1], bc[p + 2], bc[p], rr[p], ro[p],
                                        omega, one_m_omega);
          rn[p] = upd ? nv : ro[p];
          if (rown && k >= k0 && k < k1) {  // wave-uniform; k >= k0: never the halo plane
            const bool cnt = own && upd;
            const float dd = cnt ? fabsf(nv - ro[p]) : 0.f;
            dl = (dd > dl) ? dd : dl;  // NaN never wins, as in the reference's max test
            bad = bad || (cnt && !isfinite(nv));
            const unsigned x = rj + plane_off(k);
            if (own) b0[x] = rn[p];
            // lambda(i,j,0) = lambda(i,j,1), black plane 0
            if (p == 1 && k0 == 1 && cnt) b1[x - sp] = nv;
          }
        }
        if (p >= 1 && p <= K && brow && k < k1) {  // wave-uniform
          const int s = (jb + k + 1) & 1;
          const bool upd = s ? col1 : col0;  // row and plane are updated ones
          const float c = rm1[p];
          const float cdn = __shfl_up(c, 1, 64);
          const float cup = __shfl_down(c, 1, 64);
          const float xm = s ? c : cdn;
          const float xp = s ? cup : c;
          const float o = bm[p + 1];
          const float nv = relax_masked(bmk[p - 1], cv, xp, xm, rn[p], rm2[p], rm1[p + 1], rm1[p - 1], br[p - 1], o,
                                        omega, one_m_omega);
          const bool cnt = own && upd;
          const float dd = cnt ? fabsf(nv - o) : 0.f;
          dl = (dd > dl) ? dd : dl;
          bad = bad || (cnt && !isfinite(nv));
          const unsigned x = rb + plane_off(k);
          if (own) b1[x] = upd ? nv : o;
          if (p == 1 && k0 == 1 && cnt) b0[x - sp] = nv;  // lambda(i,j,0) = lambda(i,j,1), red plane 0
        }
      }
      // ---- the march moves one row
#pragma unroll
      for (int p = 0; p < NB; ++p) {
        bm[p] = bc[p];
        bc[p] = bp[p];
      }
#pragma unroll
      for (int p = 0; p < NR; ++p) {
        rm2[p] = rm1[p];
        rm1[p] = rn[p];
      }
    }
  }
  if (bad) atomicOr(status, ST_NONFINITE);
  const float wm = wave_max(dl);
  if (lane == 0) wmax[wave] = wm;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float bmax = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
    if (bmax > 0.f) atomicMax(err + (long long)(t + 1) * kErrSlots + (blockIdx.x % kErrSlots), __float_as_uint(bmax));
  }
}

}  // namespace qesmc

using namespace qesmc;

namespace {

// The march's tiling: lanes per tile, k planes, rows per segment.
struct MarchTiling {
  int K, W, qt, nkc, L;
  long long ntiles, blocks;
};

MarchTiling march_tiling(const Dims &d, int kplanes, int n_cu)
{
  MarchTiling m{};
  // 16 planes would need 256 VGPRs (one wave per SIMD): 8 at most
  m.K = kplanes >= 8 ? 8 : 4;
  m.qt = (d.HP + 61) / 62;
  m.W = (d.HP + m.qt - 1) / m.qt;
  m.nkc = (d.NZ - 2 + m.K - 1) / m.K;
  // rows per segment: the most (at most 64) that still give >= 16 waves per CU
  auto tiles = [&](int l) { return (long long)m.qt * m.nkc * ((d.NY - 2 + l - 1) / l); };
  m.L = 64;
  while (m.L > 4 && tiles(m.L) < 16LL * n_cu) m.L /= 2;
  m.ntiles = tiles(m.L);
  m.blocks = (m.ntiles + kWaves - 1) / kWaves;
  m.blocks = (m.blocks + kXcds - 1) / kXcds * kXcds;
  return m;
}

}  // namespace

// The wall values layer-major for the march's scalar loads: canon8[k*8 + a]
// = canon[a*NZ + k] (words 6 and 7 unused).
__global__ __launch_bounds__(256) void k_canon_transpose(int NZ, const unsigned *__restrict__ canon,
                                                         unsigned *__restrict__ canon8)
{
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= NZ * 8) return;
  const int k = e >> 3, a = e & 7;
  canon8[e] = a < 6 ? canon[a * NZ + k] : 0u;
}

hipError_t qesmc_fused_prepare(qesmc_ctx *x)
{
  const int n = x->d.NZ * 8;
  hipLaunchKernelGGL(k_canon_transpose, dim3((n + 255) / 256), dim3(256), 0, x->stream, x->d.NZ, x->d_canon,
                     x->d_canon8);
  return hipGetLastError();
}

// Launch iteration t of the fused path: reads buffer set t & 1, writes the other.
hipError_t qesmc_launch_fused_iteration(qesmc_ctx *x, int t, int itermax, float omega, float tol)
{
  const Dims &d = x->d;
  if (d.NZ - 2 < 1 || d.NY - 2 < 1) return hipSuccess;
  if (x->nsplit >= (1LL << 31)) return hipErrorInvalidValue;  // the kernel's 32-bit offsets
  const MarchTiling m = march_tiling(d, x->kplanes, x->n_cu);
  float *const *A = (t & 1) ? x->d_lc2 : x->d_lc;
  float *const *B = (t & 1) ? x->d_lc : x->d_lc2;
#define QESMC_MARCH_LAUNCH(KK)                                                                                 \
  hipLaunchKernelGGL(k_sor_march<KK>, dim3((unsigned)m.blocks), dim3(64 * kWaves), 0, x->stream, d, m.W, m.qt, \
                     m.L, m.nkc, m.ntiles, A[0], A[1], B[0], B[1], x->d_Rc[0], x->d_Rc[1], x->d_mask[0],        \
                     x->d_mask[1], x->d_canon8, omega, 1.0f - omega, tol, x->d_err, t, itermax, x->d_status,     \
                     x->d_cur)
  if (m.K == 8) QESMC_MARCH_LAUNCH(8);
  else QESMC_MARCH_LAUNCH(4);
#undef QESMC_MARCH_LAUNCH
  return hipGetLastError();
}
