// One wave tile of a red-black SOR colour pass on the colour-split layout:
// 64*VQ consecutive half-cells of row j (this lane: half-cells q..q+VQ-1)
// marched over planes [k0, k1).  Shared by the single-problem kernel and the
// batched kernel; its pointer parameters are __restrict__ so that, once
// inlined, all loads of a plane are issued before its store and the
// per-layer wall values are read through the scalar cache.
//
// Layout (both colours): cell (i,j,k) of colour c = (i+j+k)&1 lives at
// half-index q = i>>1 of row (j,k) of colour array c, x = (k*NY + j)*HP + q.
// A colour-c cell at q has i = 2q + s, s = (j+k+c)&1; its neighbours in the
// other colour's array sit at q-1+s, q+s (x-1, x+1) and at q in rows j+-1
// and planes k+-1.  Arithmetic is the reference's operand for operand
// (src/winds/Solver_CPU_RB.cpp:109-151, compiled without contraction).
#pragma once

#include "qesmc_internal.h"

namespace qesmc {

template <int VQ>
struct Vec;
template <>
struct Vec<1> {
  __device__ static inline void load(const float *p, float *v) { v[0] = p[0]; }
  __device__ static inline void store(float *p, const float *v) { p[0] = v[0]; }
  __device__ static inline void loadb(const unsigned char *p, unsigned *v) { v[0] = p[0]; }
};
template <>
struct Vec<2> {
  __device__ static inline void load(const float *p, float *v)
  {
    const float2 t = *reinterpret_cast<const float2 *>(p);
    v[0] = t.x;
    v[1] = t.y;
  }
  __device__ static inline void store(float *p, const float *v)
  {
    *reinterpret_cast<float2 *>(p) = make_float2(v[0], v[1]);
  }
  __device__ static inline void loadb(const unsigned char *p, unsigned *v)
  {
    const unsigned short t = *reinterpret_cast<const unsigned short *>(p);
    v[0] = t & 0xffu;
    v[1] = t >> 8;
  }
};
template <>
struct Vec<4> {
  __device__ static inline void load(const float *p, float *v)
  {
    const float4 t = *reinterpret_cast<const float4 *>(p);
    v[0] = t.x;
    v[1] = t.y;
    v[2] = t.z;
    v[3] = t.w;
  }
  __device__ static inline void store(float *p, const float *v)
  {
    *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
  }
  __device__ static inline void loadb(const unsigned char *p, unsigned *v)
  {
    const unsigned t = *reinterpret_cast<const unsigned *>(p);
    v[0] = t & 0xffu;
    v[1] = (t >> 8) & 0xffu;
    v[2] = (t >> 16) & 0xffu;
    v[3] = t >> 24;
  }
};

// The relaxation of one cell, operand for operand the reference's
// (Solver_CPU_RB.cpp:117-124): xp, xm, yp, ym, zp, zm are lambda at i+1, i-1,
// j+1, j-1, k+1, k-1.  Shared by every split-layout kernel so that all of
// them round identically.
__device__ __forceinline__ float relax(float ec, float fc, float gc, float hc, float mc, float nc, float xp,
                                       float xm, float yp, float ym, float zp, float zm, float r, float o,
                                       float omega, float one_m_omega)
{
  return (omega / (ec + fc + gc + hc + mc + nc)) * (ec * xp + fc * xm + gc * yp + hc * ym + mc * zp + nc * zm - r)
         + one_m_omega * o;
}

// CanonP: `const unsigned *` (a kernel argument) or a constant-address-space
// pointer (read from a table); either way the six wall values of a layer are
// wave-uniform scalar loads.
template <bool MASKED, int VQ, typename CanonP>
__device__ __forceinline__ float sor_tile(const Dims &d, int colour, int t, int j, int q, int k0, int k1,
                                          float *__restrict__ own, const float *__restrict__ oth,
                                          const float *__restrict__ Rc, const unsigned char *__restrict__ mask,
                                          CanonP __restrict__ canon, const float *__restrict__ ce,
                                          const float *__restrict__ cf, const float *__restrict__ cg,
                                          const float *__restrict__ ch, const float *__restrict__ cm,
                                          const float *__restrict__ cn, float omega, float one_m_omega,
                                          unsigned *status, float dl)
{
#pragma unroll 1
  for (int k = k0; k < k1; ++k) {
    const int s = (j + k + colour) & 1;
    const long long x = (long long)(k * d.NY + j) * d.HP + q;
    float o[VQ], r[VQ], c[VQ], yp[VQ], ym[VQ], zp[VQ], zm[VQ];
    Vec<VQ>::load(own + x, o);
    Vec<VQ>::load(Rc + x, r);
    Vec<VQ>::load(oth + x, c);
    Vec<VQ>::load(oth + x + d.HP, yp);
    Vec<VQ>::load(oth + x - d.HP, ym);
    Vec<VQ>::load(oth + x + d.splane, zp);
    Vec<VQ>::load(oth + x - d.splane, zm);
    const float left = oth[x - 1], right = oth[x + VQ];
    unsigned mk[VQ];
    // the layer's six wall values, read unconditionally (uniform addresses,
    // no per-lane branches around the loads)
    float cv[6];
    if (MASKED) {
      Vec<VQ>::loadb(mask + x, mk);
#pragma unroll
      for (int a = 0; a < 6; ++a) cv[a] = __uint_as_float(canon[a * d.NZ + k]);
    }
    float nvv[VQ];
    bool upd[VQ];
#pragma unroll
    for (int v = 0; v < VQ; ++v) {
      const int i = 2 * (q + v) + s;
      upd[v] = (i >= 1 && i <= d.NX - 2);
      const float xm = s ? c[v] : (v > 0 ? c[v - 1] : left);
      const float xp = s ? (v + 1 < VQ ? c[v + 1] : right) : c[v];
      float ec, fc, gc, hc, mc, nc;
      if (MASKED) {
        const unsigned m = mk[v];
        ec = (m & 1u) ? cv[0] : 0.f;
        fc = (m & 2u) ? cv[1] : 0.f;
        gc = (m & 4u) ? cv[2] : 0.f;
        hc = (m & 8u) ? cv[3] : 0.f;
        mc = (m & 16u) ? cv[4] : 0.f;
        nc = (m & 32u) ? cv[5] : 0.f;
      } else {
        ec = upd[v] ? ce[x + v] : 0.f;
        fc = upd[v] ? cf[x + v] : 0.f;
        gc = upd[v] ? cg[x + v] : 0.f;
        hc = upd[v] ? ch[x + v] : 0.f;
        mc = upd[v] ? cm[x + v] : 0.f;
        nc = upd[v] ? cn[x + v] : 0.f;
      }
      const float nv = relax(ec, fc, gc, hc, mc, nc, xp, xm, yp[v], ym[v], zp[v], zm[v], r[v], o[v], omega,
                             one_m_omega);
      nvv[v] = nv;
      if (upd[v]) {
        if (!isfinite(nv)) atomicOr(status, ST_NONFINITE);
        const float dd = fabsf(nv - o[v]);
        dl = (dd > dl) ? dd : dl;  // NaN never wins, as in the reference's max test
        o[v] = nv;
      }
    }
    // positions that are not updated are written back unchanged: every
    // half-cell of this row segment belongs to this lane in this pass
    Vec<VQ>::store(own + x, o);
    // k = 0 mirror (Solver_CPU_RB.cpp:154-162), kept in the other colour:
    // own plane 0 at this half-index is ghost (i0,j,0), whose only reader is
    // cell (i0,j,1) of the other colour, read in the *next* pass.  Its value
    // must be lambda(i0,j,1) at the end of the previous iteration, which
    // is oth plane 1 right now (that colour is not relaxed until the next
    // pass).  In the first red pass the ghosts keep their initial values,
    // as the reference mirrors only after an iteration.
    if (k == 1 && !(colour == 0 && t == 0)) {
#pragma unroll
      for (int v = 0; v < VQ; ++v) {
        const int i0 = 2 * (q + v) + (s ^ 1);
        if (i0 >= 1 && i0 <= d.NX - 2) own[x - d.splane + v] = c[v];
      }
    }
    (void)nvv;
  }
  return dl;
}

}  // namespace qesmc
