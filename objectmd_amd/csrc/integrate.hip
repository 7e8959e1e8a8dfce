// Velocity Verlet in FP64 (MDIntegrator::Integrate,
// src/omd/integrator.cpp:225-259) and the global reductions of
// MDSystem::MeasurePotential/MeasureKinetic (src/omd/system.cpp:1060-1094).
//
// kick_drift: v += (dt/2 f)/m; x += dt v; refresh the FP32 cell-relative
//             snapshot; track the largest displacement since the last list
//             build and the largest speed (list-validity check on the host).
// kick:       v += (dt/2 f)/m.
// Inactive atoms are not moved; fixed atoms (flag bit 8) keep x and v.
#include <hip/hip_runtime.h>

#include "omdg_internal.h"

namespace omdg {

namespace {

constexpr int kBlock = 256;

struct IntArgs {
  double* x;
  double* v;
  const float* f32;
  const double* f64;
  const int* type;
  const int* flag;
  const int* cellid;
  float4* q;
  const float4* qb;
  const double* xb;
  int* stat;
  float* part_max;  // per-block max |disp|^2 and |v|^2
  int n;
  double dt, hdt;
  double mass[kMaxTypes];
  double hdt_over_m[kMaxTypes];  // FP32-force mode: (dt/2)/m
  int M[3];
  double origin[3], h[3];
  float half_skin2;
};

__device__ inline float wave_max(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Block maxima go to a per-block slot with a plain store (thousands of
// same-address atomics would serialise in L2); k_status_max folds the slots
// into the status words when the host checks the list.
__device__ inline void block_max_store(float a, float b, float* part) {
  __shared__ float sa[kBlock / 64], sb[kBlock / 64];
  a = wave_max(a);
  b = wave_max(b);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l == 0) {
    sa[w] = a;
    sb[w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float ma = sa[0], mb = sb[0];
    for (int k = 1; k < kBlock / 64; k++) {
      ma = fmaxf(ma, sa[k]);
      mb = fmaxf(mb, sb[k]);
    }
    part[2 * blockIdx.x] = ma;
    part[2 * blockIdx.x + 1] = mb;
  }
}

__global__ void __launch_bounds__(kBlock) k_status_max(const float* part, int nb, int* stat) {
  float a = 0.f, b = 0.f;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    a = fmaxf(a, part[2 * k]);
    b = fmaxf(b, part[2 * k + 1]);
  }
  __shared__ float sa[kBlock / 64], sb[kBlock / 64];
  a = wave_max(a);
  b = wave_max(b);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l == 0) {
    sa[w] = a;
    sb[w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBlock / 64; k++) {
      a = fmaxf(a, sa[k]);
      b = fmaxf(b, sb[k]);
    }
    a = fmaxf(a, sa[0]);
    b = fmaxf(b, sb[0]);
    // non-negative floats order like their bit patterns
    stat[ST_DISP2] = __float_as_int(a);
    stat[ST_VMAX2] = __float_as_int(b);
  }
}

// FP64 updates are rounded as written (no fused multiply-add), like the
// reference's x86-64 build.
template <bool FP64F, bool FLAGS>
__global__ void __launch_bounds__(kBlock) k_kick_drift(IntArgs a) {
#pragma clang fp contract(off)
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  float disp2 = 0.f, v2max = 0.f;
  if (s < a.n) {
    const int n = a.n;
    bool move = true;
    if (FLAGS) {
      const int fl = a.flag[s];
      move = (fl & OMDG_FLAG_ACTIVE) && !(fl & OMDG_FLAG_FIXED);
    }
    float4 qb;
    int ty;
    if (FP64F) {
      ty = a.type[s];
    } else {
      qb = a.qb[s];
      ty = snap_type(qb.w);
    }
    const double m = a.mass[ty];
    const double hdt_m = a.hdt_over_m[ty];
    double xs[3];
    for (int d = 0; d < 3; d++) {
      const size_t k = (size_t)d * n + s;
      double x = a.x[k];
      if (move) {
        double v;
        if (FP64F) {
          v = a.v[k] + a.hdt * a.f64[k] / m;  // integrator.cpp:236 order
        } else {
          v = a.v[k] + (double)a.f32[k] * hdt_m;
        }
        x = x + a.dt * v;
        a.v[k] = v;
        a.x[k] = x;
        v2max += (float)(v * v);
      }
      xs[d] = x;
    }
    if (FP64F) {
      double d2 = 0;
      for (int d = 0; d < 3; d++) {
        const double dd = xs[d] - a.xb[(size_t)d * n + s];
        d2 += dd * dd;
      }
      disp2 = (float)d2;
    } else {
      int k0, k1, k2;
      snap_cell(qb.w, k0, k1, k2);
      float4 qn;
      qn.x = (float)(xs[0] - (a.origin[0] + (double)k0 * a.h[0]));
      qn.y = (float)(xs[1] - (a.origin[1] + (double)k1 * a.h[1]));
      qn.z = (float)(xs[2] - (a.origin[2] + (double)k2 * a.h[2]));
      qn.w = qb.w;
      a.q[s] = qn;
      const float ex = qn.x - qb.x, ey = qn.y - qb.y, ez = qn.z - qb.z;
      disp2 = ex * ex + ey * ey + ez * ez;
    }
    if (disp2 > a.half_skin2) atomicOr(a.stat + ST_ERR, ERR_SKIN);
  }
  block_max_store(disp2, v2max, a.part_max);
}

template <bool FP64F, bool FLAGS>
__global__ void __launch_bounds__(kBlock) k_kick(IntArgs a) {
#pragma clang fp contract(off)
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  if (FLAGS) {
    const int fl = a.flag[s];
    if (!(fl & OMDG_FLAG_ACTIVE) || (fl & OMDG_FLAG_FIXED)) return;
  }
  const int n = a.n;
  const int ty = a.type[s];
  for (int d = 0; d < 3; d++) {
    const size_t k = (size_t)d * n + s;
    if (FP64F) a.v[k] = a.v[k] + a.hdt * a.f64[k] / a.mass[ty];  // integrator.cpp:253
    else a.v[k] = a.v[k] + (double)a.f32[k] * a.hdt_over_m[ty];
  }
}

IntArgs make_args(Ctx& c, double dt) {
  IntArgs a;
  a.x = c.x[c.cur];
  a.v = c.v[c.cur];
  a.f32 = c.f32;
  a.f64 = c.f64;
  a.type = c.type[c.cur];
  a.flag = c.flag[c.cur];
  a.cellid = c.cellid;
  a.q = c.q;
  a.qb = c.qb;
  a.xb = c.xb;
  a.stat = c.dstat;
  a.part_max = c.dmax;
  a.n = c.n;
  a.dt = dt;
  a.hdt = 0.5 * dt;
  for (int t = 0; t < kMaxTypes; t++) {
    a.mass[t] = c.mass[t];
    a.hdt_over_m[t] = a.hdt / c.mass[t];
  }
  for (int d = 0; d < 3; d++) {
    a.M[d] = c.grid.M[d];
    a.origin[d] = c.grid.origin[d];
    a.h[d] = c.grid.h[d];
  }
  a.half_skin2 = (float)(0.25 * c.skin * c.skin);
  return a;
}

// ---- reductions -------------------------------------------------------------
struct RedArgs {
  const double* v;
  const int* type;
  const int* flag;
  const float* pot32;
  const float* vir32;
  const double* pot64;
  const double* vir64;
  double* part;
  int n;
  int fp64;
  int has_flags;
  int kinetic, potential;
  double mass[kMaxTypes];
};

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__global__ void __launch_bounds__(kBlock) k_energy_part(RedArgs a) {
  double ek = 0, ep = 0, vr = 0;
  const int n = a.n;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    if (a.has_flags && !(a.flag[s] & OMDG_FLAG_ACTIVE)) continue;
    if (a.kinetic) {
      const double vx = a.v[s], vy = a.v[(size_t)n + s], vz = a.v[2 * (size_t)n + s];
      ek += a.mass[a.type[s]] * (vx * vx + vy * vy + vz * vz);
    }
    if (a.potential) {
      if (a.fp64) {
        ep += a.pot64[s];
        vr += a.vir64[s];
      } else {
        ep += (double)a.pot32[s];
        vr += (double)a.vir32[s];
      }
    }
  }
  __shared__ double sh[3][kBlock / 64];
  ek = wave_sum(ek);
  ep = wave_sum(ep);
  vr = wave_sum(vr);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l == 0) {
    sh[0][w] = ek;
    sh[1][w] = ep;
    sh[2][w] = vr;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r[3] = {0, 0, 0};
    for (int k = 0; k < kBlock / 64; k++)
      for (int q = 0; q < 3; q++) r[q] += sh[q][k];
    for (int q = 0; q < 3; q++) a.part[blockIdx.x * 3 + q] = r[q];
  }
}

// one workgroup of kBlock threads sums the per-block partials
__global__ void __launch_bounds__(kBlock) k_energy_final(const double* part, int nb,
                                                         double* out) {
  double r[3] = {0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    for (int q = 0; q < 3; q++) r[q] += part[b * 3 + q];
  __shared__ double sh[3][kBlock / 64];
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  for (int q = 0; q < 3; q++) {
    r[q] = wave_sum(r[q]);
    if (l == 0) sh[q][w] = r[q];
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; q++) {
      double s = 0;
      for (int k = 0; k < kBlock / 64; k++) s += sh[q][k];
      out[q] = s;
    }
}

}  // namespace

void integrate_kick_drift(Ctx& c, double dt) {
  IntArgs a = make_args(c, dt);
  const int nb = grid_blocks(c.n, kBlock);
  if (c.fp64) {
    if (c.has_flags) k_kick_drift<true, true><<<nb, kBlock, 0, c.stream>>>(a);
    else k_kick_drift<true, false><<<nb, kBlock, 0, c.stream>>>(a);
  } else {
    if (c.has_flags) k_kick_drift<false, true><<<nb, kBlock, 0, c.stream>>>(a);
    else k_kick_drift<false, false><<<nb, kBlock, 0, c.stream>>>(a);
  }
  OMDG_HIP(hipGetLastError());
}

void integrate_kick(Ctx& c, double dt) {
  IntArgs a = make_args(c, dt);
  const int nb = grid_blocks(c.n, kBlock);
  if (c.fp64) {
    if (c.has_flags) k_kick<true, true><<<nb, kBlock, 0, c.stream>>>(a);
    else k_kick<true, false><<<nb, kBlock, 0, c.stream>>>(a);
  } else {
    if (c.has_flags) k_kick<false, true><<<nb, kBlock, 0, c.stream>>>(a);
    else k_kick<false, false><<<nb, kBlock, 0, c.stream>>>(a);
  }
  OMDG_HIP(hipGetLastError());
}

void integrate_status(Ctx& c) {
  k_status_max<<<1, kBlock, 0, c.stream>>>(c.dmax, grid_blocks(c.n, kBlock), c.dstat);
  OMDG_HIP(hipGetLastError());
}

void zero_forces(Ctx& c) {
  OMDG_HIP(hipMemsetAsync(c.f32, 0, sizeof(float) * 3 * (size_t)c.n, c.stream));
  OMDG_HIP(hipMemsetAsync(c.f64, 0, sizeof(double) * 3 * (size_t)c.n, c.stream));
}

void reduce_energy(Ctx& c, bool kinetic, bool potential) {
  RedArgs a;
  a.v = c.v[c.cur];
  a.type = c.type[c.cur];
  a.flag = c.flag[c.cur];
  a.pot32 = c.pot32;
  a.vir32 = c.vir32;
  a.pot64 = c.pot64;
  a.vir64 = c.vir64;
  a.part = c.dred;
  a.n = c.n;
  a.fp64 = c.fp64;
  a.has_flags = c.has_flags ? 1 : 0;
  a.kinetic = kinetic;
  a.potential = potential;
  for (int t = 0; t < kMaxTypes; t++) a.mass[t] = c.mass[t];
  const int nb = c.red_blocks;
  k_energy_part<<<nb, kBlock, 0, c.stream>>>(a);
  k_energy_final<<<1, kBlock, 0, c.stream>>>(c.dred, nb, c.dres);
  OMDG_HIP(hipGetLastError());
  OMDG_HIP(hipMemcpyAsync(c.hres, c.dres, sizeof(double) * 3, hipMemcpyDeviceToHost, c.stream));
  OMDG_HIP(hipStreamSynchronize(c.stream));
  if (kinetic) c.ekin = 0.5 * c.hres[0];
  if (potential) {
    c.epot = c.hres[1];
    c.virial = 0.5 * c.hres[2];
  }
}

}  // namespace omdg
