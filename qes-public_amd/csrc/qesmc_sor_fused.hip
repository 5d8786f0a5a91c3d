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
// The halos are recomputed rather than exchanged, and the tiles that share
// them are dispatched together so that the re-reads mostly hit in L2: the
// four waves of a work-group are x-neighbours (a 58-lane row segment is not
// line aligned, so neighbours share cache lines), the k-neighbours a few
// work-groups on.  Per iteration a cell costs its own lambda read and
// write, R and the mask: 13 bytes per update, against 17 in a split colour
// pass (which also reads the other colour's lambda), plus the halo re-reads
// (K+4 black planes loaded per K relaxed, W+2 lanes per W, L+2 red rows per
// L).  The factor
// omega / (e+f+g+h+m+n) of a cell depends only on its layer and its 6-bit
// mask, so each wave tabulates it for its planes in LDS (with the same
// expression, hence the same value) instead of dividing per cell; with the
// coefficients selected by bit masks this took a quarter of the march's
// vector instructions away.  On grids beyond the Infinity Cache the march
// is bound by its HBM traffic, about 1.25 x 13 bytes per update with the
// halos, against 1.17 x 17 for the split passes:
// QESMC_PATH_AUTO picks it there and the split kernel where the arrays stay
// in the cache, whose bandwidth the split passes use better.
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

// Wall value a of a layer (its bit pattern cb) where mask bit a is set, +0
// where it is not: the bit sign-extended to a word and ANDed with the value.
template <int A>
__device__ __forceinline__ float wall_coef(unsigned m, unsigned cb)
{
  return __uint_as_float(cb & (unsigned)((int)(m << (31 - A)) >> 31));
}

// relax() of qesmc_sor_tile.h with its factor omega / (e+f+g+h+m+n) read
// from a table (wd, computed by the same expression) instead of divided per
// cell: the same operations on the same values, so the same result.
__device__ __forceinline__ float relax_tab(unsigned m, const unsigned *cb, float wd, float xp, float xm, float yp,
                                           float ym, float zp, float zm, float r, float o, float one_m_omega)
{
  const float ec = wall_coef<0>(m, cb[0]), fc = wall_coef<1>(m, cb[1]), gc = wall_coef<2>(m, cb[2]),
              hc = wall_coef<3>(m, cb[3]), mc = wall_coef<4>(m, cb[4]), nc = wall_coef<5>(m, cb[5]);
  return wd * (ec * xp + fc * xm + gc * yp + hc * ym + mc * zp + nc * zm - r) + one_m_omega * o;
}

}  // namespace

template <int K>
__global__ __launch_bounds__(256, 3) void k_sor_march(Dims d, int W, int qt, int L, int nkc, long long ntiles,
                                                   const float *__restrict__ a0, const float *__restrict__ a1,
                                                   float *__restrict__ b0, float *__restrict__ b1,
                                                   const float *__restrict__ R0, const float *__restrict__ R1,
                                                   const unsigned char *__restrict__ mask0,
                                                   const unsigned char *__restrict__ mask1,
                                                   const unsigned *__restrict__ canon8, float omega_r,
                                                   float omo_r, float omega_b, float omo_b, float tol, unsigned *err,
                                                   int t,
                                                   int itermax, unsigned *status, unsigned *cur)
{
  constexpr int NB = K + 4;  // black-old plane slots: planes k0-2 .. k0+K+1
  constexpr int NR = K + 2;  // red plane slots: planes k0-1 .. k0+K
  __shared__ float wmax[kWaves];
  // per wave, colour (red, black) and relaxed plane k0-1+p: omega / (sum of
  // the wall values a cell's mask selects), for each of the 64 masks, and
  // the six wall values
  __shared__ float wd_tab[kWaves][2][NR][64];
  __shared__ unsigned cv_tab[kWaves][NR][8];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (t >= itermax) return;
  {
    // gate: iteration t runs iff max(err row t) > tol (row 0 = 1.0f)
    const float g = wave_max(__uint_as_float(err[(long long)t * kErrSlots + lane]));
    if (!gate_pass(g, t, tol, status)) return;  // identical in every wave of the grid
  }
  // XCD x (work-group id % 8) takes the x-th eighth of the work-group range;
  // tiles are numbered x-tile fastest, so the 4 waves of a work-group are
  // x-neighbours in one k-chunk and row segment and share the cache lines at
  // their edges; the k-neighbours (sharing halo planes) follow qt tiles on
  const long long per = (long long)gridDim.x / kXcds;  // the grid is a multiple of 8
  const long long blk = (long long)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  if (blk == 0 && threadIdx.x == 0) *cur = (unsigned)((t + 1) & 1);
  const long long tile = blk * kWaves + wave;
  const int tq = (int)(tile % qt);
  const long long tr = tile / qt;
  const int tk = (int)(tr % nkc);
  const long long tj = tr / nkc;
  const bool valid = tile < ntiles;
  {
    // this wave's tables (planes clamped into the domain; a tile past the
    // end fills harmless values it never reads)
    const int k0t = 1 + tk * K;
    for (int e = lane; e < NR * 64; e += 64) {
      const int p = e >> 6, mm = e & 63;
      int k = k0t - 1 + p;
      k = k < 0 ? 0 : k > d.NZ - 1 ? d.NZ - 1 : k;
      float v[6];
#pragma unroll
      for (int a = 0; a < 6; ++a) v[a] = ((mm >> a) & 1) ? __uint_as_float(canon8[k * 8 + a]) : 0.f;
      const float sum = v[0] + v[1] + v[2] + v[3] + v[4] + v[5];
      wd_tab[wave][0][p][mm] = omega_r / sum;
      wd_tab[wave][1][p][mm] = omega_b / sum;
    }
    for (int e = lane; e < NR * 8; e += 64) {
      const int p = e >> 3;
      int k = k0t - 1 + p;
      k = k < 0 ? 0 : k > d.NZ - 1 ? d.NZ - 1 : k;
      cv_tab[wave][p][e & 7] = canon8[k * 8 + (e & 7)];
    }
  }
  __syncthreads();

  float dl = 0.f;
  bool bad = false;
  if (valid) {  // wave-uniform, and no barrier inside
    const int k0 = 1 + tk * K;
    const int k1 = k0 + K < d.NZ - 1 ? k0 + K : d.NZ - 1;  // owned planes [k0, k1)
    const int j0 = 1 + (int)tj * L;
    const int j1 = j0 + L < d.NY - 1 ? j0 + L : d.NY - 1;  // owned rows [j0, j1)
    // lanes past the right halo (W + 2 <= 64) repeat its address: they load
    // nothing the wave does not already fetch, and own nothing
    const int q = tq * W - 1 + (lane < W + 2 ? lane : W + 1);
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
#pragma unroll
      for (int p = 0; p < NR; ++p) {
        const int k = k0 - 1 + p;
        // the plane's wall values (LDS broadcast reads), read here rather
        // than hoisted for all planes at once, which would hold 6 more
        // VGPRs per plane through the step
        asm volatile("" ::: "memory");
        unsigned cv[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) cv[a] = cv_tab[wave][p][a];
        {
          const int s = (j + k) & 1;
          const bool upd = (rowu && k >= 1 && k <= d.NZ - 2) && (s ? col1 : col0);
          const float c = bc[p + 1];
          const float cdn = __shfl_up(c, 1, 64);    // half-cell q-1
          const float cup = __shfl_down(c, 1, 64);  // half-cell q+1
          const float xm = s ? c : cdn;
          const float xp = s ? cup : c;
          const float nv = relax_tab(rmk[p], cv, wd_tab[wave][0][p][rmk[p] & 63u], xp, xm, bp[p + 1], bm[p + 1],
                                     bc[p + 2], bc[p], rr[p], ro[p], omo_r);
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
          const float nv = relax_tab(bmk[p - 1], cv, wd_tab[wave][1][p][bmk[p - 1] & 63u], xp, xm, rn[p], rm2[p],
                                     rm1[p + 1], rm1[p - 1], br[p - 1], o, omo_b);
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

constexpr int kMarchWavesPerCu = 8;  // segment rule below

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
  // rows per segment: the most (at most 64) that still give >= 8 wave tiles
  // per CU (measured: 16 gave 4-row segments on 400 x 400 x 120, 72.7 us per
  // iteration against 64 us with 8-row ones; 1000 x 1000 x 150 within 1 %)
  auto tiles = [&](int l) { return (long long)m.qt * m.nkc * ((d.NY - 2 + l - 1) / l); };
  m.L = 64;
  while (m.L > 4 && tiles(m.L) < (long long)kMarchWavesPerCu * n_cu) m.L /= 2;
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
hipError_t qesmc_launch_fused_iteration(qesmc_ctx *x, int t, int itermax, float omega_red, float omega_black,
                                        float tol)
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
                     x->d_mask[1], x->d_canon8, omega_red, 1.0f - omega_red, omega_black, 1.0f - omega_black, tol,  \
                     x->d_err, t, itermax, x->d_status, x->d_cur)
  if (m.K == 8) QESMC_MARCH_LAUNCH(8);
  else QESMC_MARCH_LAUNCH(4);
#undef QESMC_MARCH_LAUNCH
  return hipGetLastError();
}
