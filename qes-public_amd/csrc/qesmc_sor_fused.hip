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
// Work decomposition.  A work-group owns a tile of 64 half-cells (q) x 8
// rows (j) x kb planes of the black colour and of the red colour.  Marching
// up the planes, it first relaxes the red cells of plane p over the tile
// plus a one-cell halo (66 x 10, the halo recomputed from A exactly as its
// owner does) into a 3-plane LDS ring, then the black cells of plane p-1,
// whose six red neighbours are all in the ring.  Per iteration a cell costs
// its own lambda read and write, R, the mask and one read of the other
// colour (as a neighbour of a red cell): 13 bytes, against 17 per colour
// pass when each pass is its own launch.  Tiles are numbered q-fastest, then
// j, then plane chunk, and XCD x (work-group id % 8) takes the x-th eighth
// of them, so the halo rows and columns a work-group re-reads sit in the L2
// of the work-groups that own them.
#include "qesmc_sor_tile.h"

namespace qesmc {

namespace {

constexpr int kXcds = 8;
constexpr int kTq = 64;          // half-cells per tile row (one per lane)
constexpr int kTj = 8;           // rows per tile (two per wave)
constexpr int kRw = kTq + 2;     // ring row length (halo column each side)
constexpr int kRh = kTj + 2;     // ring rows (halo row each side)

// Work-group barrier that orders LDS only: __syncthreads() also fences
// global memory, which makes every wave wait for all of its outstanding
// global loads (the prefetched next step) and stores at each barrier.
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

// Everything one plane step reads from global memory: the inputs of this
// lane's three red cells of plane p and of its two black cells of plane p-1.
struct StepIn {
  float cvr[6], cvb[6];  // wall values of the red plane p and the black plane p-1
  float o[3], rv[3], nb[3][6];
  unsigned mk[3];
  long long xr[3];
  bool valid[3], upd[3];
  float ob[2], rb[2];
  unsigned mb[2];
  long long xb[2];
  bool ub[2];
  int sb[2];
};

// Loads of plane step p, addresses clamped in bounds and no branches, so
// that all of them are in flight together.
__device__ __forceinline__ StepIn load_step(const Dims &d, int p, int k0, int j0, int q0, int wave, int lane,
                                           const int *rr_, const int *cc_, const bool *in_,
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
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int j = j0 - 1 + rr_[r];
    const int q = q0 - 1 + cc_[r];
    L.valid[r] = in_[r] && j >= 0 && j < d.NY && q >= 0 && q < d.HP;
    const int sr = (j + p) & 1;
    const int i = 2 * q + sr;
    L.upd[r] = L.valid[r] && i >= 1 && i <= d.NX - 2 && j >= 1 && j <= d.NY - 2 && p >= 1 && p <= d.NZ - 2;
    const long long x = L.valid[r] ? ((long long)p * d.NY + j) * HP + q : 0;
    const bool u = L.upd[r];
    L.xr[r] = x;
    L.o[r] = a0[x];
    L.rv[r] = R0[x];
    L.mk[r] = mask0[x];
    L.nb[r][0] = a1[u ? x + sr : x];
    L.nb[r][1] = a1[u ? x - 1 + sr : x];
    L.nb[r][2] = a1[u ? x + HP : x];
    L.nb[r][3] = a1[u ? x - HP : x];
    L.nb[r][4] = a1[u ? x + sp : x];
    L.nb[r][5] = a1[u ? x - sp : x];
  }
  const int pb = p - 1;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int j = j0 + 2 * wave + a;
    const int q = q0 + lane;
    L.sb[a] = (j + pb + 1) & 1;
    const int i = 2 * q + L.sb[a];
    L.ub[a] = pb >= k0 && j <= d.NY - 2 && q < d.HP && i >= 1 && i <= d.NX - 2;
    const long long x = L.ub[a] ? ((long long)pb * d.NY + j) * HP + q : 0;
    L.xb[a] = x;
    L.ob[a] = a1[x];
    L.rb[a] = R1[x];
    L.mb[a] = mask1[x];
  }
  return L;
}

// Plane step p: red cells of plane p into the ring (and to B where this
// tile owns them), then black cells of plane p-1 from the ring.
__device__ __forceinline__ void compute_step(const Dims &d, int p, int k0, int k1, int wave, int lane,
                                             const int *rr_, const int *cc_, const bool *in_, const StepIn &L,
                                             float (*ring)[kRh][kRw], float *__restrict__ b0,
                                             float *__restrict__ b1, float omega, float one_m_omega, float &dl, bool &bad)
{
  const long long sp = d.splane;
  const int pb = p - 1;  // black plane relaxed in this step
  const float *o = L.o, *rv = L.rv, *ob = L.ob, *rb = L.rb;
  const unsigned *mk = L.mk, *mb = L.mb;
  const long long *xr = L.xr, *xb = L.xb;
  const bool *valid = L.valid, *upd = L.upd, *ub = L.ub;
  const int *sb = L.sb;
  const float(*nb)[6] = L.nb;
  // ---- red plane p into the ring
  {
    const float *cv = L.cvr;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const unsigned m = mk[r];
      const float nv = relax((m & 1u) ? cv[0] : 0.f, (m & 2u) ? cv[1] : 0.f, (m & 4u) ? cv[2] : 0.f,
                             (m & 8u) ? cv[3] : 0.f, (m & 16u) ? cv[4] : 0.f, (m & 32u) ? cv[5] : 0.f,
                             nb[r][0], nb[r][1], nb[r][2], nb[r][3], nb[r][4], nb[r][5], rv[r], o[r], omega,
                             one_m_omega);
      const float val = upd[r] ? nv : valid[r] ? o[r] : 0.f;
      if (in_[r]) ring[p % 3][rr_[r]][cc_[r]] = val;
      if (upd[r] && rr_[r] >= 1 && rr_[r] <= kTj && cc_[r] >= 1 && cc_[r] <= kTq && p >= k0 && p < k1) {
        b0[xr[r]] = nv;
        if (p == 1) b1[xr[r] - sp] = nv;  // lambda(i,j,0) = lambda(i,j,1), black plane 0
        if (!isfinite(nv)) bad = true;
        const float dd = fabsf(nv - o[r]);
        dl = (dd > dl) ? dd : dl;  // NaN never wins, as in the reference's max test
      }
    }
  }
  lds_barrier();
  // ---- black plane p-1 from the ring (planes p-2, p-1, p)
  if (pb >= k0) {
    const float *cv = L.cvb;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      if (ub[a]) {
        const int rr = 1 + 2 * wave + a;
        const int cc = lane + 1;
        const int s = sb[a];
        const unsigned m = mb[a];
        const float(*r0)[kRw] = ring[pb % 3];
        const float nv = relax((m & 1u) ? cv[0] : 0.f, (m & 2u) ? cv[1] : 0.f, (m & 4u) ? cv[2] : 0.f,
                               (m & 8u) ? cv[3] : 0.f, (m & 16u) ? cv[4] : 0.f, (m & 32u) ? cv[5] : 0.f,
                               r0[rr][cc + s], r0[rr][cc - 1 + s], r0[rr + 1][cc], r0[rr - 1][cc],
                               ring[(pb + 1) % 3][rr][cc], ring[(pb + 2) % 3][rr][cc], rb[a], ob[a], omega,
                               one_m_omega);
        b1[xb[a]] = nv;
        if (pb == 1) b0[xb[a] - sp] = nv;  // lambda(i,j,0) = lambda(i,j,1), red plane 0
        if (!isfinite(nv)) bad = true;
        const float dd = fabsf(nv - ob[a]);
        dl = (dd > dl) ? dd : dl;
      }
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
    if (!(g > tol)) return;  // identical in every wave of the grid
  }
  const long long per = (long long)gridDim.x / kXcds;  // the grid is a multiple of 8
  const long long tile = (long long)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  if (tile >= ntiles) return;  // whole work-group, before any barrier
  if (tile == 0 && threadIdx.x == 0) *cur = (unsigned)((t + 1) & 1);
  const int tq = (int)(tile % qt);
  const int tj = (int)((tile / qt) % jt);
  const int kc = (int)(tile / ((long long)qt * jt));
  const int q0 = tq * kTq;
  const int j0 = 1 + tj * kTj;
  const int k0 = 1 + kc * kb;
  const int k1 = k0 + kb < d.NZ - 1 ? k0 + kb : d.NZ - 1;  // black planes [k0, k1)
  const long long HP = d.HP, sp = d.splane;

  float dl = 0.f;
  bool bad = false;
  const int pbeg = k0 - 1;  // >= 0
  const int pend = k1;      // <= NZ-1

  // Ring positions of this lane's red cells, three rounds of the four waves:
  // rows 0..9 x columns 1..64, and in the third round waves 2 and 3 take the
  // halo columns 0 and 65 (lanes 0..9).
  int rr_[3], cc_[3];
  bool in_[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    rr_[r] = wave + 4 * r;
    cc_[r] = lane + 1;
    in_[r] = rr_[r] < kRh;
  }
  if (wave >= 2) {
    rr_[2] = lane;
    cc_[2] = wave == 2 ? 0 : kRw - 1;
    in_[2] = lane < kRh;
  }

  // One plane step at a time.  Prefetching step p + 1 during step p (two
  // load buffers) was measured slower: the extra registers cut occupancy
  // from 5 to 3 waves per SIMD, and this kernel needs the waves more.
#pragma unroll 1
  for (int p = pbeg; p <= pend; ++p) {
    const StepIn L = load_step(d, p, k0, j0, q0, wave, lane, rr_, cc_, in_, a0, a1, R0, R1, mask0, mask1, canon);
    compute_step(d, p, k0, k1, wave, lane, rr_, cc_, in_, L, ring, b0, b1, omega, one_m_omega, dl, bad);
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
  const int qt = (d.HP + kTq - 1) / kTq;
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
