// HIP kernels of the mass-consistency step for gfx950 (wave64).
//
// Arithmetic parity: the library is compiled with -ffp-contract=off, and
// every expression below keeps the operand order and float/double
// promotions of the reference's CPU red-black solver
// (src/winds/Solver_CPU_RB.cpp), so the device results equal the CPU ones
// bit for bit (IEEE fp32/fp64, correctly rounded division).
#include "qesmc_internal.h"

namespace qesmc {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// max over a 256-thread block laid out as (64, 4): wave64 butterfly through
// DPP/permute, then one LDS slot per wave.  Every thread must call it.
__device__ inline float block_max_64x4(float v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  __shared__ float wmax[4];
  if (threadIdx.x == 0) wmax[threadIdx.y] = v;
  __syncthreads();
  float r = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
  return r;
}

__device__ inline double block_max_64x4_d(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  __shared__ double wmax[4];
  if (threadIdx.x == 0) wmax[threadIdx.y] = v;
  __syncthreads();
  return fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
}

// |new - old| as the reference's stop test sees it: a NaN never wins the
// max (`if (error > max_error)` is false for NaN, Solver_CPU_RB.cpp:175).
__device__ inline float stop_delta(float nv, float ov)
{
  float dl = fabsf(nv - ov);
  return (dl > 0.f) ? dl : 0.f;
}

// Iteration gate: iteration t runs iff max(err row t) > tol, with row 0
// holding 1.0f, the reference's initial max_error (Solver_CPU_RB.cpp:91,95).
// Lane l of each wave reads slot l of the row (blockDim.x == 64).
__device__ inline bool gate_open(const unsigned *err, int t, float tol, unsigned *status)
{
  float v = __uint_as_float(err[(long long)t * kErrSlots + threadIdx.x]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return gate_pass(v, t, tol, status);
}

// Block's max|dlambda| into its striped slot of row t+1.
__device__ inline void publish_delta(unsigned *err, int t, float bm)
{
  if (threadIdx.x == 0 && threadIdx.y == 0 && bm > 0.f) {
    const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    atomicMax(err + (long long)(t + 1) * kErrSlots + (b % kErrSlots), __float_as_uint(bm));
  }
}

// ---------------------------------------------------------------------------
// R: divergence of u0 (Solver_CPU_RB.cpp:56-80), k in [1, nz-3]
// ---------------------------------------------------------------------------
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_divergence(Dims d, float dx, float dy,
                                                    const float *__restrict__ dz,
                                                    const float *__restrict__ e,
                                                    const float *__restrict__ f,
                                                    const float *__restrict__ g,
                                                    const float *__restrict__ h,
                                                    const float *__restrict__ m,
                                                    const float *__restrict__ n,
                                                    const float *__restrict__ u0,
                                                    const float *__restrict__ v0,
                                                    const float *__restrict__ w0, double pre,
                                                    float *__restrict__ R, float *__restrict__ R0,
                                                    float *__restrict__ R1)
{
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z + 1;
  if (i >= d.NX || j >= d.NY) return;
  const long long c = i + j * d.sy + k * d.sz;
  const long long fc = i + j * d.fy + k * d.fz;
  const float dzk = dz[k], dzp = dz[k + 1], dzm = dz[k - 1];
  const float r = pre * (((e[c] * u0[fc + 1] - f[c] * u0[fc]) * dx)
                         + ((g[c] * v0[fc + d.fy] - h[c] * v0[fc]) * dy)
                         + ((m[c] * dzk * 0.5 * (dzk + dzp) * w0[fc + d.fz]
                             - n[c] * dzk * 0.5f * (dzk + dzm) * w0[fc])));
  if (SPLIT) {
    const int colour = (i + j + k) & 1;
    const long long s = (long long)(k * d.NY + j) * d.HP + (i >> 1);
    (colour ? R1 : R0)[s] = r;
  } else {
    R[c] = r;
  }
}

// ---------------------------------------------------------------------------
// Plain path: one colour of red-black SOR in place on the x-fastest layout
// (Solver_CPU_RB.cpp:109-151).  Thread t of a row handles i = 2t + s, the
// cells of this colour.  The k = 0 mirror (:154-162) is folded in: the only
// reader of lambda(i,j,0) is cell (i,j,1) itself, so it is written right
// after that cell is relaxed.  max|dlambda| (:164-180) is formed in registers.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rb_plain(Dims d, int colour, int kplanes,
                                                  const float *__restrict__ e,
                                                  const float *__restrict__ f,
                                                  const float *__restrict__ g,
                                                  const float *__restrict__ h,
                                                  const float *__restrict__ m,
                                                  const float *__restrict__ n,
                                                  const float *__restrict__ R, float *lam,
                                                  float omega, float one_m_omega, float tol,
                                                  unsigned *err, int t, int itermax,
                                                  unsigned *status)
{
  if (t >= itermax || !gate_open(err, t, tol, status)) return;  // uniform per launch
  const int j = blockIdx.y * 4 + threadIdx.y + 1;
  const int tq = blockIdx.x * 64 + threadIdx.x;
  const int k0 = 1 + blockIdx.z * kplanes;
  const int k1 = min(k0 + kplanes, d.NZ - 1);
  float dl = 0.f;
  for (int k = k0; k < k1; ++k) {
    const int s = (colour + j + k) & 1;
    const int i = 2 * tq + s;
    if (j <= d.NY - 2 && i >= 1 && i <= d.NX - 2) {
      const long long c = i + j * d.sy + k * d.sz;
      const float ec = e[c], fc = f[c], gc = g[c], hc = h[c], mc = m[c], nc = n[c];
      const float old = lam[c];
      const float nv = (omega / (ec + fc + gc + hc + mc + nc))
                           * (ec * lam[c + 1] + fc * lam[c - 1] + gc * lam[c + d.sy]
                              + hc * lam[c - d.sy] + mc * lam[c + d.sz] + nc * lam[c - d.sz] - R[c])
                       + one_m_omega * old;
      lam[c] = nv;
      if (k == 1) lam[c - d.sz] = nv;
      if (!isfinite(nv)) atomicOr(status, ST_NONFINITE);
      dl = fmaxf(dl, stop_delta(nv, old));
    }
  }
  publish_delta(err, t, block_max_64x4(dl));
}

// ---------------------------------------------------------------------------
// Split path.  Colour c of cell (i,j,k) is (i+j+k)&1; both colours store
// cell i at half-index q = i>>1 of row (j,k):  idx = (k*NY + j)*HP + q.
// A colour-c cell at q has i = 2q + s, s = (j+k+c)&1, and its neighbours in
// the other colour's array are: x-1 -> q-1+s, x+1 -> q+s, y/z -> same q.
// ---------------------------------------------------------------------------

// Wall value of each direction on each layer: the largest coefficient over
// the updated cells (stair-step coefficients are 0 or that value).
// gridDim.x work-groups per plane (k = blockIdx.y + 1) stride over the
// plane's updated cells; 6 atomics per work-group on per-plane addresses.
__global__ __launch_bounds__(256) void k_canon_reduce(Dims d, const float *__restrict__ e,
                                                      const float *__restrict__ f,
                                                      const float *__restrict__ g,
                                                      const float *__restrict__ h,
                                                      const float *__restrict__ m,
                                                      const float *__restrict__ n,
                                                      unsigned *canon, unsigned *status)
{
  const int k = blockIdx.y + 1;
  const int W = d.NX - 2;
  const int nplane = W * (d.NY - 2);
  const float *arr[6] = {e, f, g, h, m, n};
  float mx[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  unsigned bad = 0;
  for (int idx = blockIdx.x * 256 + threadIdx.y * 64 + threadIdx.x; idx < nplane; idx += gridDim.x * 256) {
    const int jr = idx / W;
    const long long c = (idx - jr * W + 1) + (long long)(jr + 1) * d.sy + k * d.sz;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      const float v = arr[a][c];
      // negative, NaN or Inf coefficients cannot be encoded as a wall bit
      if (!(v >= 0.f) || isinf(v)) bad = 1;
      else mx[a] = v > mx[a] ? v : mx[a];
    }
  }
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    const float bm = block_max_64x4(mx[a]);
    if (threadIdx.x == 0 && threadIdx.y == 0 && bm > 0.f) atomicMax(canon + a * d.NZ + k, __float_as_uint(bm));
    __syncthreads();
  }
  if (bad) atomicOr(status, ST_INCOMPRESSIBLE);
}

// Per updated cell: the 6-bit wall mask (bit a set <=> coefficient a equals
// its layer's wall value), in the split layout; flags coefficients that are
// neither 0 nor the wall value, and zero diagonals (the reference would
// divide by zero at Solver_CPU_RB.cpp:116).
__global__ __launch_bounds__(256) void k_build_mask(Dims d, const float *__restrict__ e,
                                                    const float *__restrict__ f,
                                                    const float *__restrict__ g,
                                                    const float *__restrict__ h,
                                                    const float *__restrict__ m,
                                                    const float *__restrict__ n,
                                                    const unsigned *__restrict__ canon,
                                                    unsigned char *__restrict__ mask0,
                                                    unsigned char *__restrict__ mask1,
                                                    unsigned *status)
{
  const int i = blockIdx.x * 64 + threadIdx.x + 1;
  const int j = blockIdx.y * 4 + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > d.NX - 2 || j > d.NY - 2) return;
  const long long c = i + (long long)j * d.sy + k * d.sz;
  const float v[6] = {e[c], f[c], g[c], h[c], m[c], n[c]};
  unsigned mk = 0, st = 0;
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    const float w = __uint_as_float(canon[a * d.NZ + k]);
    if (v[a] == w && w != 0.f) mk |= 1u << a;
    else if (v[a] != 0.f) st |= ST_INCOMPRESSIBLE;
  }
  if (((((v[0] + v[1]) + v[2]) + v[3]) + v[4]) + v[5] == 0.f) st |= ST_ZERO_DIAG;
  if (st) atomicOr(status, st);
  const int colour = (i + j + k) & 1;
  const long long s = (long long)(k * d.NY + j) * d.HP + (i >> 1);
  (colour ? mask1 : mask0)[s] = (unsigned char)mk;
}

// Full coefficients in the split layout (used when the mask encoding does
// not hold, e.g. fractional cut-cell coefficients).
__global__ __launch_bounds__(256) void k_split_coefs(Dims d, const float *__restrict__ src,
                                                     float *__restrict__ dst0,
                                                     float *__restrict__ dst1)
{
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.NX || j >= d.NY) return;
  const long long c = i + (long long)j * d.sy + k * d.sz;
  const long long s = (long long)(k * d.NY + j) * d.HP + (i >> 1);
  (((i + j + k) & 1) ? dst1 : dst0)[s] = src[c];
}

// natural <-> split copies of lambda (every cell, all layers)
__global__ __launch_bounds__(256) void k_pack(Dims d, const float *__restrict__ lam,
                                              float *__restrict__ l0, float *__restrict__ l1)
{
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.NX || j >= d.NY) return;
  const long long c = i + (long long)j * d.sy + k * d.sz;
  const long long s = (long long)(k * d.NY + j) * d.HP + (i >> 1);
  (((i + j + k) & 1) ? l1 : l0)[s] = lam[c];
}

// split -> natural from the buffer set the fused path left its last iterate in
__global__ __launch_bounds__(256) void k_unpack_selected(Dims d, float *__restrict__ lam,
                                                         const float *__restrict__ a0,
                                                         const float *__restrict__ a1,
                                                         const float *__restrict__ b0,
                                                         const float *__restrict__ b1,
                                                         const unsigned *__restrict__ cur)
{
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.NX || j >= d.NY) return;
  const long long c = i + (long long)j * d.sy + k * d.sz;
  const long long s = (long long)(k * d.NY + j) * d.HP + (i >> 1);
  const bool odd = (i + j + k) & 1;
  lam[c] = (*cur ? (odd ? b1 : b0) : (odd ? a1 : a0))[s];
}

__global__ __launch_bounds__(256) void k_unpack(Dims d, float *__restrict__ lam,
                                                const float *__restrict__ l0,
                                                const float *__restrict__ l1)
{
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.NX || j >= d.NY) return;
  const long long c = i + (long long)j * d.sy + k * d.sz;
  const long long s = (long long)(k * d.NY + j) * d.HP + (i >> 1);
  lam[c] = (((i + j + k) & 1) ? l1 : l0)[s];
}

// ---------------------------------------------------------------------------
// k = 0 mirror over the whole plane (Solver_CPU_RB.cpp:154-162); run once
// after the loop for the never-updated ring cells (interior columns were
// mirrored in the sweeps).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mirror(Dims d, float *__restrict__ lam)
{
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= d.NX || j >= d.NY) return;
  const long long c = i + (long long)j * d.sy;
  lam[c] = lam[c + d.sz];
}

// ---------------------------------------------------------------------------
// Velocity correction (Solver_CPU_RB.cpp:194-260): u = u0 on every face,
// corrected on k in [1,nz-3], j in [1,ny-2], i in [1,nx-2].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_correct(Dims d, float dx, float dy,
                                                 const float *__restrict__ dz,
                                                 const float *__restrict__ f,
                                                 const float *__restrict__ h,
                                                 const float *__restrict__ n,
                                                 const float *__restrict__ u0,
                                                 const float *__restrict__ v0,
                                                 const float *__restrict__ w0,
                                                 const float *__restrict__ lam, float p1, float p2,
                                                 float *__restrict__ u, float *__restrict__ v,
                                                 float *__restrict__ w)
{
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const long long fc = i + j * d.fy + k * d.fz;
  float uu = u0[fc], vv = v0[fc], ww = w0[fc];
  if (k >= 1 && k < d.nz - 2 && j >= 1 && j < d.ny - 1 && i >= 1 && i < d.nx - 1) {
    const long long c = i + j * d.sy + k * d.sz;
    const float lc = lam[c];
    uu = uu + p1 * f[c] * dx * (lc - lam[c - 1]);
    vv = vv + p1 * h[c] * dy * (lc - lam[c - d.sy]);
    ww = ww + p2 * n[c] * dz[k] * (lc - lam[c - d.sz]);
  }
  u[fc] = uu;
  v[fc] = vv;
  w[fc] = ww;
}

// Solid zeroing (Solver_CPU_RB.cpp:262-284), k in [1, nz-2].  Runs after the
// correction kernel has finished (the reference GPU kernel does both in one
// launch and races on neighbour faces).
__global__ __launch_bounds__(256) void k_zero_solid(Dims d, const int *__restrict__ flag,
                                                    float *__restrict__ u, float *__restrict__ v,
                                                    float *__restrict__ w)
{
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z + 1;
  if (i >= d.NX || j >= d.NY) return;
  const long long c = i + j * d.sy + k * d.sz;
  const int fl = flag[c];
  if (fl == 0 || fl == 2) {
    const long long fc = i + j * d.fy + k * d.fz;
    u[fc] = 0.f;
    u[fc + 1] = 0.f;
    v[fc] = 0.f;
    v[fc + d.fy] = 0.f;
    w[fc] = 0.f;
    w[fc + d.fz] = 0.f;
  }
}

// FP64 L-infinity divergence of the corrected field over interior fluid
// cells; reference operator (R/(-2 a1^2)) or physical (vertical / dz_k).
__global__ __launch_bounds__(256) void k_div_linf(Dims d, double dx, double dy,
                                                  const float *__restrict__ dz,
                                                  const float *__restrict__ e,
                                                  const float *__restrict__ f,
                                                  const float *__restrict__ g,
                                                  const float *__restrict__ h,
                                                  const float *__restrict__ m,
                                                  const float *__restrict__ n,
                                                  const int *__restrict__ flag,
                                                  const float *__restrict__ u,
                                                  const float *__restrict__ v,
                                                  const float *__restrict__ w, int physical,
                                                  unsigned long long *out)
{
  const int k = blockIdx.y + 1;
  const int W = d.NX - 2;
  const int nplane = W * (d.NY - 2);
  const double dzk = dz[k];
  double mx = 0.0;
  for (int idx = blockIdx.x * 256 + threadIdx.y * 64 + threadIdx.x; idx < nplane; idx += gridDim.x * 256) {
    const int jr = idx / W;
    const int i = idx - jr * W + 1, j = jr + 1;
    const long long c = i + (long long)j * d.sy + k * d.sz;
    if (flag[c] != 1) continue;
    const long long fc = i + (long long)j * d.fy + k * d.fz;
    double dv;
    if (physical) {
      dv = ((double)u[fc + 1] - u[fc]) / dx + ((double)v[fc + d.fy] - v[fc]) / dy
           + ((double)w[fc + d.fz] - w[fc]) / dzk;
    } else {
      dv = (((double)e[c] * u[fc + 1] - (double)f[c] * u[fc]) * dx)
           + (((double)g[c] * v[fc + d.fy] - (double)h[c] * v[fc]) * dy)
           + ((double)m[c] * dzk * 0.5 * (dzk + (double)dz[k + 1]) * w[fc + d.fz]
              - (double)n[c] * dzk * 0.5 * (dzk + (double)dz[k - 1]) * w[fc]);
    }
    dv = fabs(dv);
    if (dv > mx) mx = dv;  // NaN never wins
  }
  const double bm = block_max_64x4_d(mx);
  if (threadIdx.x == 0 && threadIdx.y == 0 && bm > 0.0)
    atomicMax(out, (unsigned long long)__double_as_longlong(bm));
}

__global__ void k_set_u32(unsigned *p, unsigned v) { *p = v; }

// p[0 .. count) = v (a kernel rather than hipMemsetAsync, so that the solve
// schedule holds no memset nodes if a caller captures it in a graph)
__global__ __launch_bounds__(256) void k_fill_u32(unsigned *__restrict__ p, long long count, unsigned v)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = blockIdx.x * 256LL + threadIdx.x; q < count; q += stride) p[q] = v;
}

// The largest entry of one convergence row read as floats, NaN never winning
// (the host's reading of a row).
__device__ inline float row_max_dev(const unsigned *row)
{
  float mx = 0.f;
  for (int q = 0; q < kErrSlots; ++q) {
    const float v = __uint_as_float(row[q]);
    mx = v > mx ? v : mx;
  }
  return mx;
}

// Iterations that ran: the first row t in [0, launched) whose max is not
// above tol (iteration t is gated on row t), else launched.  first is
// pre-set to launched; rows are independent, one thread per row.
__global__ __launch_bounds__(256) void k_rows_first_stop(const unsigned *__restrict__ rows, int launched, float tol,
                                                         unsigned *__restrict__ first)
{
  for (int t = blockIdx.x * 256 + threadIdx.x; t < launched; t += gridDim.x * 256)
    if (!(row_max_dev(rows + (long long)t * kErrSlots) > tol)) atomicMin(first, (unsigned)t);
}

// out[0] = iterations (first[0]; with the production stop rule the end it
// recorded, if below), out[1] = bits of the max of row out[0]: the last
// executed iteration's max|dlambda| (the sentinel 1.0 if none ran).
__global__ void k_rows_summary(const unsigned *__restrict__ rows, const unsigned *__restrict__ first,
                               const unsigned *__restrict__ status, int launched, unsigned *__restrict__ out)
{
  unsigned count = *first;
  if (__uint_as_float(status[kStopExtend]) > 0.f) {
    const unsigned end = status[kStopAt];
    count = end < (unsigned)launched ? end : (unsigned)launched;
  }
  out[0] = count;
  out[1] = __float_as_uint(row_max_dev(rows + (long long)count * kErrSlots));
}

// ---------------------------------------------------------------------------
// host-side launch wrappers
// ---------------------------------------------------------------------------
static inline dim3 grid_xy(long long nxthreads, long long nythreads, long long nzblocks)
{
  return dim3((unsigned)((nxthreads + 63) / 64), (unsigned)((nythreads + 3) / 4), (unsigned)nzblocks);
}
static const dim3 kBlock(64, 4, 1);
static const unsigned kPlaneBlocks = 8;  // work-groups per plane of the plane reductions

}  // namespace qesmc

using namespace qesmc;

hipError_t qesmc_launch_fused_iteration(qesmc_ctx *x, int t, int itermax, float omega_red, float omega_black,
                                        float tol);
hipError_t qesmc_launch_split_iteration(qesmc_ctx *x, bool masked, int t, int itermax, float omega_red,
                                        float omega_black, float tol);

// Entry points used by the API translation unit.
hipError_t qesmc_launch_divergence(qesmc_ctx *x, bool split, int alpha1)
{
  const Dims &d = x->d;
  if (d.NZ - 2 < 1) return hipSuccess;
  const double pre = -2.0f * pow((double)alpha1, 2.0);
  dim3 gr = grid_xy(d.NX, d.NY, d.NZ - 2);
  if (split)
    hipLaunchKernelGGL(k_divergence<true>, gr, kBlock, 0, x->stream, d, x->dx, x->dy, x->d_dz,
                       x->d_coef[0], x->d_coef[1], x->d_coef[2], x->d_coef[3], x->d_coef[4],
                       x->d_coef[5], x->d_u0, x->d_v0, x->d_w0, pre, x->d_R, x->d_Rc[0], x->d_Rc[1]);
  else
    hipLaunchKernelGGL(k_divergence<false>, gr, kBlock, 0, x->stream, d, x->dx, x->dy, x->d_dz,
                       x->d_coef[0], x->d_coef[1], x->d_coef[2], x->d_coef[3], x->d_coef[4],
                       x->d_coef[5], x->d_u0, x->d_v0, x->d_w0, pre, x->d_R, x->d_Rc[0], x->d_Rc[1]);
  return hipGetLastError();
}

hipError_t qesmc_launch_stencil_encode(qesmc_ctx *x)
{
  const Dims &d = x->d;
  if (d.NZ - 2 < 1 || d.NX < 3 || d.NY < 3) return hipSuccess;
  dim3 gr = grid_xy(d.NX - 2, d.NY - 2, d.NZ - 2);
  hipLaunchKernelGGL(k_canon_reduce, dim3(kPlaneBlocks, d.NZ - 2), kBlock, 0, x->stream, d, x->d_coef[0], x->d_coef[1],
                     x->d_coef[2], x->d_coef[3], x->d_coef[4], x->d_coef[5], x->d_canon, x->d_status);
  hipLaunchKernelGGL(k_build_mask, gr, kBlock, 0, x->stream, d, x->d_coef[0], x->d_coef[1],
                     x->d_coef[2], x->d_coef[3], x->d_coef[4], x->d_coef[5], x->d_canon,
                     x->d_mask[0], x->d_mask[1], x->d_status);
  return hipGetLastError();
}

hipError_t qesmc_launch_split_coefs(qesmc_ctx *x)
{
  const Dims &d = x->d;
  dim3 gr = grid_xy(d.NX, d.NY, d.NZ);
  for (int a = 0; a < 6; ++a)
    hipLaunchKernelGGL(k_split_coefs, gr, kBlock, 0, x->stream, d, x->d_coef[a], x->d_cc[0][a],
                       x->d_cc[1][a]);
  return hipGetLastError();
}

hipError_t qesmc_launch_pack(qesmc_ctx *x, bool to_split)
{
  const Dims &d = x->d;
  dim3 gr = grid_xy(d.NX, d.NY, d.NZ);
  if (to_split)
    hipLaunchKernelGGL(k_pack, gr, kBlock, 0, x->stream, d, x->d_lam, x->d_lc[0], x->d_lc[1]);
  else
    hipLaunchKernelGGL(k_unpack, gr, kBlock, 0, x->stream, d, x->d_lam, x->d_lc[0], x->d_lc[1]);
  return hipGetLastError();
}

hipError_t qesmc_launch_unpack_selected(qesmc_ctx *x)
{
  const Dims &d = x->d;
  hipLaunchKernelGGL(k_unpack_selected, grid_xy(d.NX, d.NY, d.NZ), kBlock, 0, x->stream, d, x->d_lam, x->d_lc[0],
                     x->d_lc[1], x->d_lc2[0], x->d_lc2[1], x->d_cur);
  return hipGetLastError();
}

// Iteration t: the red pass at omega_red, the black pass at omega_black.
hipError_t qesmc_launch_iteration(qesmc_ctx *x, int path, bool masked, int t, int itermax,
                                  float omega_red, float omega_black, float tol)
{
  const Dims &d = x->d;
  if (d.NZ - 2 < 1 || d.NY - 2 < 1) return hipSuccess;
  const int kp = x->kplanes < 1 ? 1 : x->kplanes;
  const dim3 gr = grid_xy((d.NX + 1) / 2, d.NY - 2, (d.NZ - 2 + kp - 1) / kp);
  if (path == QESMC_PATH_PLAIN) {
    for (int colour = 0; colour < 2; ++colour) {
      const float omega = colour ? omega_black : omega_red;
      hipLaunchKernelGGL(k_rb_plain, gr, kBlock, 0, x->stream, d, colour, kp, x->d_coef[0],
                         x->d_coef[1], x->d_coef[2], x->d_coef[3], x->d_coef[4], x->d_coef[5],
                         x->d_R, x->d_lam, omega, 1.0f - omega, tol, x->d_err, t, itermax, x->d_status);
    }
  } else if (path == QESMC_PATH_FUSED) {
    return qesmc_launch_fused_iteration(x, t, itermax, omega_red, omega_black, tol);
  } else {
    return qesmc_launch_split_iteration(x, masked, t, itermax, omega_red, omega_black, tol);
  }
  return hipGetLastError();
}

hipError_t qesmc_launch_finish(qesmc_ctx *x, bool mirror, int alpha1, int alpha2)
{
  const Dims &d = x->d;
  if (mirror && d.NZ >= 2)
    hipLaunchKernelGGL(k_mirror, grid_xy(d.NX, d.NY, 1), kBlock, 0, x->stream, d, x->d_lam);
  const float p1 = 1.0f / (2.0f * (float)pow((double)alpha1, 2.0));
  const float p2 = 1.0f / (2.0f * (float)pow((double)alpha2, 2.0));
  hipLaunchKernelGGL(k_correct, grid_xy(d.nx, d.ny, d.nz), kBlock, 0, x->stream, d, x->dx, x->dy,
                     x->d_dz, x->d_coef[1], x->d_coef[3], x->d_coef[5], x->d_u0, x->d_v0, x->d_w0,
                     x->d_lam, p1, p2, x->d_u, x->d_v, x->d_w);
  if (d.NZ - 1 >= 1)
    hipLaunchKernelGGL(k_zero_solid, grid_xy(d.NX, d.NY, d.NZ - 1), kBlock, 0, x->stream, d,
                       x->d_flag, x->d_u, x->d_v, x->d_w);
  return hipGetLastError();
}

hipError_t qesmc_launch_div_linf(qesmc_ctx *x, int physical)
{
  const Dims &d = x->d;
  hipError_t st = hipMemsetAsync(x->d_norm, 0, sizeof(double), x->stream);
  if (st != hipSuccess) return st;
  if (d.NZ - 2 < 1 || d.NX < 3 || d.NY < 3) return hipSuccess;
  hipLaunchKernelGGL(k_div_linf, dim3(kPlaneBlocks, d.NZ - 2), kBlock, 0, x->stream, d,
                     (double)x->dx, (double)x->dy, x->d_dz, x->d_coef[0], x->d_coef[1],
                     x->d_coef[2], x->d_coef[3], x->d_coef[4], x->d_coef[5], x->d_flag, x->d_u,
                     x->d_v, x->d_w, physical, (unsigned long long *)x->d_norm);
  return hipGetLastError();
}

hipError_t qesmc_launch_set_u32(qesmc_ctx *x, unsigned *p, unsigned v)
{
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, x->stream, p, v);
  return hipGetLastError();
}

hipError_t qesmc_launch_fill_u32(qesmc_ctx *x, unsigned *p, long long count, unsigned v)
{
  if (count <= 0) return hipSuccess;
  long long blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, x->stream, p, count, v);
  return hipGetLastError();
}

// Iteration count and last max|dlambda| of a run from its convergence rows,
// into d_status[1] (scan) and d_status[2..3] (count, last bits).
hipError_t qesmc_launch_rows_summary(qesmc_ctx *x, int launched, float tol)
{
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, x->stream, x->d_status + 1, (unsigned)launched);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (launched > 0) {
    int blocks = (launched + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_rows_first_stop, dim3(blocks), dim3(256), 0, x->stream, x->d_err, launched, tol,
                       x->d_status + 1);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_rows_summary, dim3(1), dim3(1), 0, x->stream, x->d_err, x->d_status + 1, x->d_status,
                     launched, x->d_status + 2);
  return hipGetLastError();
}
