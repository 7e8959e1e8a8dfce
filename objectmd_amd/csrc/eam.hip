// Two-pass EAM on the full neighbour list (gather form, no atomics).
//
//   density pass  rho_i = sum_j f(r_ij) for r <= EDENS range, then
//                 F'(rho_i) and F(rho_i) once per atom in FP64
//                 (TEmbedding::IterationNode/PreCalculation/PostForce,
//                 src/gadget/team.cpp:81-145; the reference re-evaluates
//                 F'(rho) for every pair, team.cpp:230-231)
//   force pass    f_i = sum_j fr_ij (x_i - x_j) with
//                 fr = -((F'_i + F'_j) rho'(r) + phi'(r)) / r for
//                 r <= PAIR range (TForceEAM::ComputeHalf, team.cpp:224-237);
//                 per-atom energy 1/2 sum phi + F(rho) and full pair virial
//                 (ReturnForce, forcekernel.h:92-110) on energy steps only.
//
// Pair displacements come from FP32 cell-relative snapshots:
// d = q_i - q_j - o * h with o the stencil offset of the list entry, so the
// rounding is that of |q| ~ one cell, not of the absolute coordinate.
// Per-atom sums are Kahan-compensated FP32.
#include <hip/hip_runtime.h>

#include "omdg_internal.h"
#include "spline_dev.h"

namespace omdg {

namespace {

struct EamArgs {
  const float4* q;
  const uint32_t* nbr;
  const int* nnbr;
  float* rho;
  float* fp;
  float* emb;
  float* f;  // 3 planes
  float* pot;
  float* vir;
  int* err;
  int n;
  float cut_e2, cut_p2;
  float hx, hy, hz;
  TabView phi, edens, embed;
};

constexpr int kBlock = 256;

template <bool ENERGY>
__global__ void __launch_bounds__(kBlock) k_eam_density(EamArgs a) {
  __shared__ float4 offs[27];
  load_offsets(offs, a.hx, a.hy, a.hz);
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  const int n = a.n;
  const float4 qi = a.q[s];
  const int nn = a.nnbr[s];
  Kahan rho;
  for (int k = 0; k < nn; k++) {
    const uint32_t e = a.nbr[(size_t)k * n + s];
    const float4 qj = a.q[e & kSlotMask];
    const float4 o = offs[e >> kImgShift];
    const float dx = qi.x - qj.x - o.x;
    const float dy = qi.y - qj.y - o.y;
    const float dz = qi.z - qj.z - o.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    if (r2 <= a.cut_e2) {
      float v;
      spline_value32(a.edens, sqrtf(r2), v, a.err);
      rho.add(v);
    }
  }
  const float r = rho.s;
  a.rho[s] = r;
  // embedding function and derivative in FP64 (once per atom)
  double F, dF;
  spline_read64(a.embed, (double)r, F, dF, a.err);
  a.fp[s] = (float)dF;
  if (ENERGY) a.emb[s] = (float)F;
}

template <bool ENERGY>
__global__ void __launch_bounds__(kBlock) k_eam_force(EamArgs a) {
  __shared__ float4 offs[27];
  load_offsets(offs, a.hx, a.hy, a.hz);
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  const int n = a.n;
  const float4 qi = a.q[s];
  const float fpi = a.fp[s];
  const int nn = a.nnbr[s];
  Kahan fx, fy, fz, pot, vir;
  for (int k = 0; k < nn; k++) {
    const uint32_t e = a.nbr[(size_t)k * n + s];
    const uint32_t t = e & kSlotMask;
    const float4 qj = a.q[t];
    const float4 o = offs[e >> kImgShift];
    const float dx = qi.x - qj.x - o.x;
    const float dy = qi.y - qj.y - o.y;
    const float dz = qi.z - qj.z - o.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    if (r2 <= a.cut_p2) {
      const float r = sqrtf(r2);
      float p, dp, rv, drho = 0.f;
      spline_read32(a.phi, r, p, dp, a.err);
      if (r2 <= a.cut_e2) spline_read32(a.edens, r, rv, drho, a.err);
      const float fr = -((fpi + a.fp[t]) * drho + dp) / r;
      fx.add(fr * dx);
      fy.add(fr * dy);
      fz.add(fr * dz);
      if (ENERGY) {
        pot.add(0.5f * p);
        vir.add(fr * r2);
      }
    }
  }
  a.f[s] = fx.s;
  a.f[(size_t)n + s] = fy.s;
  a.f[2 * (size_t)n + s] = fz.s;
  if (ENERGY) {
    a.pot[s] = pot.s + a.emb[s];
    a.vir[s] = vir.s;
  }
}

}  // namespace

void eam_forces(Ctx& c, bool energy) {
  EamArgs a;
  a.q = c.q;
  a.nbr = c.nbr;
  a.nnbr = c.nnbr;
  a.rho = c.rho;
  a.fp = c.fp;
  a.emb = c.emb;
  a.f = c.f32;
  a.pot = c.pot32;
  a.vir = c.vir32;
  a.err = c.dstat + ST_ERR;
  a.n = c.n;
  a.phi = tab_view(c, c.tab_index[OMDG_TAB_PAIR][0][0]);
  a.edens = tab_view(c, c.tab_index[OMDG_TAB_EDENS][0][0]);
  a.embed = tab_view(c, c.tab_index[OMDG_TAB_EMBED][0][0]);
  // cut radii: table max ranges (team.cpp:78, team.cpp:212)
  const double ce = a.edens.r0d + (double)(a.edens.n - 1) * a.edens.drd;
  const double cp = a.phi.r0d + (double)(a.phi.n - 1) * a.phi.drd;
  a.cut_e2 = (float)(ce * ce);
  a.cut_p2 = (float)(cp * cp);
  a.hx = (float)c.grid.h[0];
  a.hy = (float)c.grid.h[1];
  a.hz = (float)c.grid.h[2];
  const int nb = grid_blocks(c.n, kBlock);
  {
    TimerScope ts(c, TG_PASS1);
    if (energy) k_eam_density<true><<<nb, kBlock, 0, c.stream>>>(a);
    else k_eam_density<false><<<nb, kBlock, 0, c.stream>>>(a);
  }
  {
    TimerScope ts(c, TG_FORCE);
    if (energy) k_eam_force<true><<<nb, kBlock, 0, c.stream>>>(a);
    else k_eam_force<false><<<nb, kBlock, 0, c.stream>>>(a);
  }
  OMDG_HIP(hipGetLastError());
}

}  // namespace omdg
