// Three-body potentials on the full neighbour list, gather form (each atom
// accumulates its own force; no atomics, deterministic).  FP32 arithmetic on
// the cell-relative snapshots with Kahan-compensated per-atom sums.
//
// Stillinger-Weber (src/gadget/sw.cpp), three passes:
//   pass 1    list scan: each atom's near list (its SW neighbours inside
//             a*sigma with their displacement vectors); pairs of other type
//             combinations with a PAIR table use the table (TForcePair,
//             tpair.h:62-71); then the two-body terms of the near list: ZBL
//             core below sx0, polynomial splice below sx1, SW beyond
//             (sw.cpp:118-189, 240-252)
//   pass 2    three-body h(j,i,k) for every centre i and pair {j,k} of its
//             near neighbours (sw.cpp:191-236, 254-273), evaluated per centre:
//             the centre's force, and per leg the force and energy share it
//             receives from this centre's triplets
//   pass 3    each atom adds the leg shares stored by its near neighbours
//             for it; every triplet gives a third of its energy to each of
//             its atoms, as the reference does
// Tersoff (PRB 37 6991 with the cutoff of include/potential/tersoff.h:247-257)
//   pass 1    per bond (i,j): zeta_ij, b_ij, the pair energy 1/2 V_ij and
//             P_ij = 1/2 fc fA db/dzeta stored per list slot
//   pass 2    atom l gathers the pair terms of bonds (l,j) and (j,l), the
//             zeta derivatives of its own bonds, of the bonds (j,l) and of the
//             bonds (j,k) in whose zeta it appears.
#include <hip/hip_runtime.h>

#include <cmath>

#include "omdg_internal.h"
#include "spline_dev.h"

namespace omdg {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxNear = 24;  // three-body neighbours inside the cut kept per atom

struct TBArgs {
  const float4* q;
  const uint32_t* nbr;
  const int* nnbr;
  float* f;
  float* pot;
  float* vir;
  float* bond_b;  // Tersoff: b_ij per list slot
  float* bond_p;  // Tersoff: P_ij per list slot
  float4* near;   // SW: near list, kMaxNear slot rows of n (x_j - x_l, entry bits)
  float4* gsum;   // SW: per near entry, force on that leg from the centre's triplets
  int* near_n;    // SW: near-list length per atom
  const int* only_if;  // Tersoff passes run only when *only_if != 0 (if set)
  int* err;
  int n;
  float hx, hy, hz;
  // independent replicas: each atom's replica and the cell edges of its grid
  const int* rep;        // null for one system
  const float4* rep_h;
  float rc, rc2;
  SwParams sw;
  TersoffParams ts;
  // SW constants (sw.cpp:62-93)
  float c1, c2, c3, c4, c5, c6, c7, c8, c9, c10;
  float zc, ze[4];
  int sw_type;  // atom type the SW parameters apply to (-1: all)
  int pi, qi;   // SW exponents p, q when small integers (-1 otherwise)
  signed char tab[kMaxTypes * kMaxTypes];
  float tcut2[kMaxTypes * kMaxTypes];
  TabView tabs[kMaxTypes * kMaxTypes];
};

// a neighbour inside the cut (Tersoff passes)
struct Near {
  float dx, dy, dz, r;
  uint32_t e;
};

__device__ inline float3 disp(const float4& a, const float4& b, const float4& o) {
  // x_b(image) - x_a for a list entry of a with offset o: q_b - q_a + o
  return make_float3(b.x - a.x + o.x, b.y - a.y + o.y, b.z - a.z + o.z);
}

// ---------------------------------------------------------------- SW ------
struct SwTerm {
  float gj[3];  // force on leg j
  float gk[3];  // force on leg k
  float e;      // triplet energy
};

// ThreeBodyTerm of sw.cpp:191-236 with d1 = x_j - x_i, d2 = x_k - x_i.
__device__ inline void sw_three(const TBArgs& a, const float* d1, float r1, const float* d2,
                                float r2, SwTerm& t) {
  const float ira0 = 1.0f / (r1 - a.rc), ira1 = 1.0f / (r2 - a.rc);
  const float fi0 = a.c7 * ira0, fi1 = a.c7 * ira1;
  const float ir12 = 1.0f / (r1 * r2);
  const float fex = expf(fi0) * expf(fi1);
  const float cs = (d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2]) * ir12;
  const float dc = cs - a.c10;
  const float ftheta = a.c9 * fex * dc;
  const float ft12 = ir12 * ftheta;
  const float fp = a.c8 * fex * dc * dc;
  const float fcos = cs * ftheta;
  const float fr0 = fp * fi0 * ira0 / r1, fr1 = fp * fi1 * ira1 / r2;
  const float ft0 = fcos / (r1 * r1), ft1 = fcos / (r2 * r2);
  for (int m = 0; m < 3; m++) {
    t.gj[m] = d1[m] * (fr0 + ft0) - d2[m] * ft12;
    t.gk[m] = d2[m] * (fr1 + ft1) - d1[m] * ft12;
  }
  t.e = fp;
}

// Two-body SW with the ZBL core and splice; returns fr (force on the first
// atom = fr * (x_first - x_second)) and the half energy.
__device__ inline void sw_two(const TBArgs& a, float r, float& fr, float& ehalf) {
  const SwParams& p = a.sw;
  if (p.usezbl && r < p.sx0) {
    float ex[4], s = 0.f, sf = 0.f;
    const float w[4] = {0.1818f, 0.5099f, 0.2802f, 0.02817f};
    const float ir = 1.0f / r;
    for (int m = 0; m < 4; m++) {
      ex[m] = w[m] * expf(-a.ze[m] * r);
      s += ex[m];
      sf += (ir + a.ze[m]) * ex[m];
    }
    fr = sf * a.zc / (r * r);
    ehalf = 0.5f * s * a.zc / r;
  } else if (p.usezbl && r < p.sx1) {
    // sw.cpp:148-164 as written: force coefficient sa1 + sa2 r + sa3 r^2
    const float g = p.sa1 + p.sa2 * r + p.sa3 * r * r;
    ehalf = 0.5f * (p.sa0 + g * r);
    fr = -g;
  } else {
    const float ira = 1.0f / (r - a.rc);
    const float rp = powf(r, -p.p), rq = powf(r, -p.q);
    const float fex = expf(p.sigma * ira);
    fr = (a.c1 * rp - a.c2 * rq + (a.c3 * rp - a.c4 * rq) * ira * ira * r) * fex / (r * r);
    ehalf = 0.5f * (a.c5 * rp - a.c6 * rq) * fex;
  }
}

// stencil offset o * h of image code c (c = (ox+1) + 3 (oy+1) + 9 (oz+1))
__device__ inline float4 off_of(uint32_t code, const float4& h) {
  const int c = (int)code;
  return make_float4((float)(c % 3 - 1) * h.x, (float)((c / 3) % 3 - 1) * h.y,
                     (float)(c / 9 - 1) * h.z, 0.f);
}

__device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// One atom's entries of a slot-major per-atom array: entry k of atom i at
// k * n + i, so a wavefront reading entry k of consecutive atoms reads one
// contiguous run.
struct Col {
  float4* p;
  size_t n;
  __device__ float4& operator[](int k) const { return p[(size_t)k * n]; }
};

// r^-p for the SW exponents: repeated products when p is a small integer
__device__ inline float pow_neg(float r, float ir, float p, int pi) {
  if (pi >= 0) {
    float v = 1.0f;
    for (int k = 0; k < pi; k++) v *= ir;
    return v;
  }
  return powf(r, -p);
}

// Two-body SW with the ZBL core and splice; returns fr (force on the first
// atom = fr * (x_first - x_second)) and the half energy.
__device__ inline void sw_two_fast(const TBArgs& a, float r, float& fr, float& ehalf) {
  const SwParams& p = a.sw;
  if (p.usezbl && r < p.sx1) {
    sw_two(a, r, fr, ehalf);  // ZBL core and splice (close collisions)
    return;
  }
  const float ir = rcp(r);
  const float ira = rcp(r - a.rc);
  const float rp = pow_neg(r, ir, p.p, a.pi), rq = pow_neg(r, ir, p.q, a.qi);
  const float fex = expf(p.sigma * ira);
  fr = (a.c1 * rp - a.c2 * rq + (a.c3 * rp - a.c4 * rq) * ira * ira * r) * fex * ir * ir;
  ehalf = 0.5f * (a.c5 * rp - a.c6 * rq) * fex;
}

// Pass 1: the list scan and the two-body terms.  Phase A walks the full
// list: SW neighbours inside a*sigma go to the atom's near list as
// {x_j - x_l, entry} (slot-major, kMaxNear slots), pairs of other type combinations with a
// PAIR table take the table (TForcePair, tpair.h:62-71).  Phase B evaluates
// the SW two-body term of each near neighbour (no divergence from the list
// scan).
template <bool ENERGY>
__global__ void __launch_bounds__(kBlock) k_sw_pair(TBArgs a) {
  const int l = xcd_block() * blockDim.x + threadIdx.x;
  if (l >= a.n) return;
  const int n = a.n;
  // neighbours belong to the atom's replica
  const float4 hh = a.rep ? a.rep_h[a.rep[l]] : make_float4(a.hx, a.hy, a.hz, 0.f);
  const float4 ql = a.q[l];
  const int tl = snap_type(ql.w);
  const bool l_sw = a.sw_type < 0 || tl == a.sw_type;
  const Col nl{a.near + l, (size_t)n};
  Kahan fx, fy, fz;
  float pot = 0.f, vir = 0.f;
  int nn = 0;
  const int cnt = a.nnbr[l];
  for (int k = 0; k < cnt; k++) {
    const uint32_t e = a.nbr[nidx(l, k, n)];
    const float4 qj = a.q[e & kSlotMask];
    const float3 d = disp(ql, qj, off_of(e >> kImgShift, hh));  // x_j - x_l
    const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    if (r2 < 1e-6f) {
      atomicOr(a.err, ERR_OVERLAP);
      continue;
    }
    const int tj = snap_type(qj.w);
    const bool pair_sw = l_sw && (a.sw_type < 0 || tj == a.sw_type);
    if (pair_sw) {
      if (r2 < a.rc2) {
        if (nn < kMaxNear) nl[nn] = make_float4(d.x, d.y, d.z, __uint_as_float(e));
        else atomicOr(a.err, ERR_NEAR);
        nn++;
      }
    } else {
      const int slot = tl * kMaxTypes + tj;
      const int tb = a.tab[slot];
      if (tb >= 0 && r2 < a.tcut2[slot]) {
        const float r = sqrtf(r2);
        float p, dp;
        spline_read32(a.tabs[tb], r, p, dp, a.err);
        const float ff = -dp / r;  // force on l = ff * (x_l - x_j)
        fx.add(-ff * d.x);
        fy.add(-ff * d.y);
        fz.add(-ff * d.z);
        if (ENERGY) {
          pot += 0.5f * p;
          vir += ff * r2;
        }
      }
    }
  }
  if (nn > kMaxNear) nn = kMaxNear;
  for (int u = 0; u < nn; u++) {
    const float4 d = nl[u];
    const float r = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
    float fr, eh;
    sw_two_fast(a, r, fr, eh);
    // force on l = fr * (x_l - x_j)
    fx.add(-fr * d.x);
    fy.add(-fr * d.y);
    fz.add(-fr * d.z);
    if (ENERGY) pot += eh;
  }
  a.near_n[l] = nn;
  a.f[l] = fx.s;
  a.f[(size_t)n + l] = fy.s;
  a.f[2 * (size_t)n + l] = fz.s;
  if (ENERGY) {
    a.pot[l] = pot;
    a.vir[l] = vir;
  }
}

// Leg data {r, 1/r, 1/(r - a sigma), exp(gamma sigma / (r - a sigma))} of a
// near entry (recomputed where needed: cheaper than storing it).
__device__ inline float4 leg_data(const TBArgs& a, const float4& d) {
  const float r = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
  const float ira = rcp(r - a.rc);
  return make_float4(r, rcp(r), ira, expf(a.c7 * ira));
}

// Force on leg j of the SW triplet h(j, i, k) (sw.cpp:191-236) and the
// triplet energy, from the legs' displacements d1 = x_j - x_i, d2 = x_k - x_i
// and their leg data {r, 1/r, 1/(r - a sigma), exp(gamma sigma/(r - a sigma))}.
__device__ inline float sw_leg(const TBArgs& a, const float4& d1, const float4& x1,
                               const float4& d2, const float4& x2, float g[3]) {
  const float ir12 = x1.y * x2.y;
  const float cs = (d1.x * d2.x + d1.y * d2.y + d1.z * d2.z) * ir12;
  const float dc = cs - a.c10;
  const float fex = x1.w * x2.w;
  const float ftheta = a.c9 * fex * dc;
  const float fp = a.c8 * fex * dc * dc;
  const float ft12 = ir12 * ftheta;
  // fp * fi0 * ira0 / r1 with fi0 = gamma sigma ira0, and cos theta ftheta / r1^2
  const float f1 = fp * a.c7 * x1.z * x1.z * x1.y + cs * ftheta * x1.y * x1.y;
  g[0] = d1.x * f1 - d2.x * ft12;
  g[1] = d1.y * f1 - d2.y * ft12;
  g[2] = d1.z * f1 - d2.z * ft12;
  return fp;
}

// Pass 2: the triplets centred at each atom i, once per ordered leg pair
// (u, w): leg u's force summed over w goes to gsum[i][u] with a third of
// the triplet energies (the leg's share); the centre takes minus the sum of
// its legs' forces and a third of each triplet's energy.
template <bool ENERGY>
__global__ void __launch_bounds__(kBlock) k_sw_centre(TBArgs a) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int n = a.n;
  const int nn = a.near_n[i];
  if (nn < 2) return;
  const Col ni{a.near + i, (size_t)n}, gi{a.gsum + i, (size_t)n};
  Kahan fx, fy, fz;
  fx.s = a.f[i];
  fy.s = a.f[(size_t)n + i];
  fz.s = a.f[2 * (size_t)n + i];
  float esum = 0.f;
  for (int u = 0; u < nn; u++) {
    const float4 du = ni[u], xu = leg_data(a, du);
    float gx = 0.f, gy = 0.f, gz = 0.f, e = 0.f;
    for (int w = 0; w < nn; w++) {
      if (w == u) continue;
      float g[3];
      const float4 dw = ni[w];
      e += sw_leg(a, du, xu, dw, leg_data(a, dw), g);
      gx += g[0];
      gy += g[1];
      gz += g[2];
    }
    gi[u] = make_float4(gx, gy, gz, e * (1.0f / 3.0f));
    fx.add(-gx);
    fy.add(-gy);
    fz.add(-gz);
    esum += e;
  }
  a.f[i] = fx.s;
  a.f[(size_t)n + i] = fy.s;
  a.f[2 * (size_t)n + i] = fz.s;
  // every unordered triplet was summed twice: a third of it to the centre
  if (ENERGY) a.pot[i] += esum * (1.0f / 6.0f);
}

// Pass 3: atom l collects, from each near neighbour c, the force and energy
// share of the triplets centred at c with l as a leg (gsum[c][slot of l]).
template <bool ENERGY>
__global__ void __launch_bounds__(kBlock) k_sw_legs(TBArgs a) {
  const int l = xcd_block() * blockDim.x + threadIdx.x;
  if (l >= a.n) return;
  const int n = a.n;
  const int nn = a.near_n[l];
  if (nn == 0) return;
  const Col nl{a.near + l, (size_t)n};
  Kahan fx, fy, fz;
  fx.s = a.f[l];
  fy.s = a.f[(size_t)n + l];
  fz.s = a.f[2 * (size_t)n + l];
  float pot = 0.f;
  for (int u = 0; u < nn; u++) {
    const float4 du = nl[u];
    const int c = (int)(__float_as_uint(du.w) & kSlotMask);
    const int nc = a.near_n[c];
    if (nc < 2) continue;
    const Col ncl{a.near + c, (size_t)n};
    for (int s = 0; s < nc; s++) {
      const float4 dk = ncl[s];  // x_k - x_c
      if ((int)(__float_as_uint(dk.w) & kSlotMask) != l) continue;
      const float ex = dk.x + du.x, ey = dk.y + du.y, ez = dk.z + du.z;
      if (ex * ex + ey * ey + ez * ez >= 1e-4f) continue;  // another image of l
      const float4 g = a.gsum[(size_t)s * n + c];
      fx.add(g.x);
      fy.add(g.y);
      fz.add(g.z);
      if (ENERGY) pot += g.w;
      break;
    }
  }
  a.f[l] = fx.s;
  a.f[(size_t)n + l] = fy.s;
  a.f[2 * (size_t)n + l] = fz.s;
  if (ENERGY) a.pot[l] += pot;
}

// ----------------------------------------------------------- Tersoff ------
__device__ inline void ters_fc(const TersoffParams& p, float r, float& fc, float& dfc) {
  if (r < p.R - p.D) {
    fc = 1.f;
    dfc = 0.f;
  } else if (r < p.R + p.D) {
    const float arg = 1.5707963267948966f * (r - p.R) / p.D;
    fc = 0.5f - 0.5f * sinf(arg);
    dfc = -0.7853981633974483f / p.D * cosf(arg);
  } else {
    fc = 0.f;
    dfc = 0.f;
  }
}

// zeta term T = fc(r_ik) g(theta) exp((l3 (r_ij - r_ik))^3) and its
// gradients with respect to x_j (gj) and x_k (gk); x_i gets -(gj + gk).
__device__ inline float ters_zeta_term(const TersoffParams& p, float l3m, const float* dij,
                                       float rij, const float* dik, float rik, float* gj,
                                       float* gk) {
  float fck, dfck;
  ters_fc(p, rik, fck, dfck);
  const float c2 = p.c * p.c, d2 = p.d * p.d;
  const float cs = (dij[0] * dik[0] + dij[1] * dik[1] + dij[2] * dik[2]) / (rij * rik);
  const float hc = p.h - cs;
  const float den = d2 + hc * hc;
  const float g = 1.f + c2 / d2 - c2 / den;
  const float dg = -2.f * c2 * hc / (den * den);
  const float dl = rij - rik;
  const float ex = expf(l3m * dl * dl * dl);
  const float dex = ex * l3m * 3.f * dl * dl;
  if (gj) {
    for (int q = 0; q < 3; q++) {
      const float uj = dij[q] / rij, uk = dik[q] / rik;
      const float dcj = (uk - cs * uj) / rij;
      const float dck = (uj - cs * uk) / rik;
      gj[q] = fck * dg * dcj * ex + fck * g * dex * uj;
      gk[q] = dfck * uk * g * ex + fck * dg * dck * ex - fck * g * dex * uk;
    }
  }
  return fck * g * ex;
}

template <bool ENERGY>
__global__ void __launch_bounds__(kBlock) k_tersoff_bond(TBArgs a) {
  if (a.only_if && !*a.only_if) return;  // uniform: a step of the surrogate
  __shared__ float4 offs[27];
  load_offsets(offs, a.hx, a.hy, a.hz);
  __syncthreads();
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int n = a.n;
  const TersoffParams& p = a.ts;
  const float l3m = p.lambda3 * p.lambda3 * p.lambda3;
  const float4 qi = a.q[i];
  const int cnt = a.nnbr[i];
  Near nr[kMaxNear];
  int slot[kMaxNear];
  int nn = 0;
  for (int k = 0; k < cnt; k++) {
    const uint32_t e = a.nbr[nidx(i, k, n)];
    const float3 d = disp(qi, a.q[e & kSlotMask], offs[e >> kImgShift]);
    const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    if (r2 < 1e-6f) atomicOr(a.err, ERR_OVERLAP);
    if (r2 < a.rc2 && r2 >= 1e-6f) {
      if (nn < kMaxNear) {
        nr[nn] = {d.x, d.y, d.z, sqrtf(r2), e};
        slot[nn] = k;
      } else {
        atomicOr(a.err, ERR_NEAR);
      }
      nn++;
    } else {
      a.bond_b[(size_t)k * n + i] = 0.f;
      a.bond_p[(size_t)k * n + i] = 0.f;
    }
  }
  if (nn > kMaxNear) nn = kMaxNear;
  float pot = 0.f;
  for (int u = 0; u < nn; u++) {
    const float dij[3] = {nr[u].dx, nr[u].dy, nr[u].dz};
    const float rij = nr[u].r;
    float zeta = 0.f;
    for (int w = 0; w < nn; w++) {
      if (w == u) continue;
      const float dik[3] = {nr[w].dx, nr[w].dy, nr[w].dz};
      zeta += ters_zeta_term(p, l3m, dij, rij, dik, nr[w].r, nullptr, nullptr);
    }
    float fc, dfc;
    ters_fc(p, rij, fc, dfc);
    const float fA = -p.B * expf(-p.lambda2 * rij);
    float b = 1.f, db = 0.f;
    if (zeta > 0.f) {
      // x = (beta zeta)^n overflows FP32 for close pairs: work with
      // lx = log x.  b = (1+x)^(-1/2n) = exp(-log1p(x) / 2n) and
      // db/dzeta = -1/2 (1+x)^(-1/2n - 1) beta^n zeta^(n-1)
      //          = -b x / (2 zeta (1+x)).
      const float lx = p.n * logf(p.beta * zeta);
      float l1p, t;
      if (lx > 30.f) {
        l1p = lx;  // log1p(x) = lx + log1p(1/x), 1/x < 1e-13
        t = 1.f;
      } else {
        const float x = expf(lx);
        l1p = log1pf(x);
        t = x / (1.f + x);
      }
      b = expf(-0.5f / p.n * l1p);
      db = -0.5f * b * t / zeta;
    }
    a.bond_b[(size_t)slot[u] * n + i] = b;
    a.bond_p[(size_t)slot[u] * n + i] = 0.5f * fc * fA * db;
    if (ENERGY) {
      const float fR = p.A * expf(-p.lambda1 * rij);
      pot += 0.5f * fc * (fR + b * fA);
    }
  }
  if (ENERGY) {
    a.pot[i] = pot;
    a.vir[i] = 0.f;
  }
}

__global__ void __launch_bounds__(kBlock) k_tersoff_force(TBArgs a) {
  if (a.only_if && !*a.only_if) return;
  __shared__ float4 offs[27];
  load_offsets(offs, a.hx, a.hy, a.hz);
  __syncthreads();
  const int l = xcd_block() * blockDim.x + threadIdx.x;
  if (l >= a.n) return;
  const int n = a.n;
  const TersoffParams& p = a.ts;
  const float l3m = p.lambda3 * p.lambda3 * p.lambda3;
  const float4 ql = a.q[l];
  const int cnt = a.nnbr[l];
  Near nr[kMaxNear];
  int slot[kMaxNear];
  int nn = 0;
  for (int k = 0; k < cnt && nn < kMaxNear; k++) {
    const uint32_t e = a.nbr[nidx(l, k, n)];
    const float3 d = disp(ql, a.q[e & kSlotMask], offs[e >> kImgShift]);
    const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    if (r2 < a.rc2 && r2 >= 1e-6f) {
      nr[nn] = {d.x, d.y, d.z, sqrtf(r2), e};
      slot[nn] = k;
      nn++;
    }
  }
  Kahan fx, fy, fz;
  auto addf = [&](float gx, float gy, float gz) {
    fx.add(gx);
    fy.add(gy);
    fz.add(gz);
  };
  // (1) bonds centred at l: pair part and zeta derivative w.r.t. the centre
  for (int u = 0; u < nn; u++) {
    const float dij[3] = {nr[u].dx, nr[u].dy, nr[u].dz};
    const float rij = nr[u].r;
    const float b = a.bond_b[(size_t)slot[u] * n + l];
    const float P = a.bond_p[(size_t)slot[u] * n + l];
    float fc, dfc;
    ters_fc(p, rij, fc, dfc);
    const float fR = p.A * expf(-p.lambda1 * rij);
    const float fA = -p.B * expf(-p.lambda2 * rij);
    const float dV = dfc * (fR + b * fA) + fc * (-p.lambda1 * fR - p.lambda2 * b * fA);
    const float pref = 0.5f * dV / rij;  // force on l: +pref * dij
    addf(pref * dij[0], pref * dij[1], pref * dij[2]);
    if (P != 0.f) {
      float sx = 0.f, sy = 0.f, sz = 0.f;
      for (int w = 0; w < nn; w++) {
        if (w == u) continue;
        const float dik[3] = {nr[w].dx, nr[w].dy, nr[w].dz};
        float gj[3], gk[3];
        ters_zeta_term(p, l3m, dij, rij, dik, nr[w].r, gj, gk);
        sx += gj[0] + gk[0];
        sy += gj[1] + gk[1];
        sz += gj[2] + gk[2];
      }
      addf(P * sx, P * sy, P * sz);
    }
  }
  // (2) bonds centred at a neighbour j: l as the j-atom (pair part of bond
  //     (j,l) and its zeta derivative) and l as a k-atom of bonds (j,k)
  for (int u = 0; u < nn; u++) {
    const uint32_t ej = nr[u].e;
    const int j = (int)(ej & kSlotMask);
    const float4 qj = a.q[j];
    const float djl[3] = {-nr[u].dx, -nr[u].dy, -nr[u].dz};  // x_l - x_j
    const float rjl = nr[u].r;
    const int cj = a.nnbr[j];
    // neighbours of j inside the cut (the bond (j,l) is among them)
    Near mj[kMaxNear];
    int ms[kMaxNear];
    int mm = 0, self = -1;
    for (int k = 0; k < cj && mm < kMaxNear; k++) {
      const uint32_t e = a.nbr[nidx(j, k, n)];
      const int t = (int)(e & kSlotMask);
      const float3 d = disp(qj, a.q[t], offs[e >> kImgShift]);  // x_t - x_j
      const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
      if (!(r2 < a.rc2) || r2 < 1e-6f) continue;
      if (t == l && self < 0) {
        const float ex = d.x - djl[0], ey = d.y - djl[1], ez = d.z - djl[2];
        if (ex * ex + ey * ey + ez * ez < 1e-4f) self = mm;
      }
      mj[mm] = {d.x, d.y, d.z, sqrtf(r2), e};
      ms[mm] = k;
      mm++;
    }
    if (self < 0) continue;  // cannot happen for a symmetric list
    // pair part of bond (j,l): force on l = -pref_jl * d_jl(x_l - x_j)
    {
      const float b = a.bond_b[(size_t)ms[self] * n + j];
      float fc, dfc;
      ters_fc(p, rjl, fc, dfc);
      const float fR = p.A * expf(-p.lambda1 * rjl);
      const float fA = -p.B * expf(-p.lambda2 * rjl);
      const float dV = dfc * (fR + b * fA) + fc * (-p.lambda1 * fR - p.lambda2 * b * fA);
      const float pref = 0.5f * dV / rjl;
      addf(-pref * djl[0], -pref * djl[1], -pref * djl[2]);
    }
    const float Pjl = a.bond_p[(size_t)ms[self] * n + j];
    float gsum[3] = {0.f, 0.f, 0.f};
    for (int w = 0; w < mm; w++) {
      if (w == self) continue;
      const float djk[3] = {mj[w].dx, mj[w].dy, mj[w].dz};
      float gj[3], gk[3];
      // bond (j,l) with k = w: gradient w.r.t. l as the j-atom
      ters_zeta_term(p, l3m, djl, rjl, djk, mj[w].r, gj, gk);
      gsum[0] += gj[0];
      gsum[1] += gj[1];
      gsum[2] += gj[2];
      // bond (j,w) with l as the k-atom
      const float Pjk = a.bond_p[(size_t)ms[w] * n + j];
      if (Pjk != 0.f) {
        ters_zeta_term(p, l3m, djk, mj[w].r, djl, rjl, gj, gk);
        addf(-Pjk * gk[0], -Pjk * gk[1], -Pjk * gk[2]);
      }
    }
    addf(-Pjl * gsum[0], -Pjl * gsum[1], -Pjl * gsum[2]);
  }
  a.f[l] = fx.s;
  a.f[(size_t)n + l] = fy.s;
  a.f[2 * (size_t)n + l] = fz.s;
}

TBArgs base_args(Ctx& c) {
  TBArgs a{};
  a.q = c.q;
  a.nbr = c.nbr;
  a.nnbr = c.nnbr;
  a.f = c.f32;
  a.pot = c.pot32;
  a.vir = c.vir32;
  a.err = c.dstat + ST_ERR;
  a.n = c.n;
  a.hx = (float)c.grid.h[0];
  a.hy = (float)c.grid.h[1];
  a.hz = (float)c.grid.h[2];
  if (c.nrep > 0) {
    a.rep = c.rep;
    a.rep_h = c.rep_h;
  }
  return a;
}

}  // namespace

// Near-list storage shared by the SW and surrogate passes: k float4
// entries per atom (slot-major), a per-entry float4 of leg results, and the
// per-atom entry count.
void near_arrays(Ctx& c, int k) {
  if ((size_t)c.n * k <= c.cap_near) return;
  for (float4** pp : {&c.near, &c.gsum}) {
    if (*pp) (void)hipFree(*pp);
    *pp = nullptr;
  }
  if (c.near_n) (void)hipFree(c.near_n);
  c.near_n = nullptr;
  c.cap_near = 0;
  const size_t bytes = sizeof(float4) * k * (size_t)c.n;
  OMDG_HIP(hipMalloc(reinterpret_cast<void**>(&c.near), bytes));
  OMDG_HIP(hipMalloc(reinterpret_cast<void**>(&c.gsum), bytes));
  OMDG_HIP(hipMalloc(reinterpret_cast<void**>(&c.near_n), sizeof(int) * (size_t)c.n));
  c.cap_near = (size_t)c.n * k;
}

void sw_forces(Ctx& c, bool energy) {
  TBArgs a = base_args(c);
  const SwParams& p = c.sw;
  a.sw = p;
  const double eps = p.eps, A = p.A, B = p.B, pp = p.p, qq = p.q, sg = p.sigma;
  a.rc = (float)((double)p.a * sg);
  a.rc2 = a.rc * a.rc;
  // constants of sw.cpp:69-80, formed in FP64
  a.c1 = (float)(A * eps * pp * B * pow(sg, pp));
  a.c2 = (float)(A * eps * qq * pow(sg, qq));
  a.c3 = (float)(A * eps * B * pow(sg, pp + 1.0));
  a.c4 = (float)(A * eps * pow(sg, qq + 1.0));
  a.c5 = (float)(A * eps * B * pow(sg, pp));
  a.c6 = (float)(A * eps * pow(sg, qq));
  a.c7 = (float)(sg * (double)p.gamma);
  a.c8 = (float)((double)p.lambda * eps);
  a.c9 = (float)(2.0 * (double)p.lambda * eps);
  a.c10 = (float)(1.0 / (double)p.icos0);
  const double znor = 0.468521928 / (pow((double)p.z1, 0.23) + pow((double)p.z2, 0.23));
  a.zc = (float)(14.39964415 * (double)p.z1 * (double)p.z2);
  a.ze[0] = (float)(3.2000 / znor);
  a.ze[1] = (float)(0.9423 / znor);
  a.ze[2] = (float)(0.4029 / znor);
  a.ze[3] = (float)(0.2016 / znor);
  // SW acts between atoms of type 0 when other type pairs have tables
  bool any_tab = false;
  int used = 0;
  for (int i = 0; i < kMaxTypes; i++)
    for (int j = 0; j < kMaxTypes; j++) {
      const int s = i * kMaxTypes + j;
      const int tid = c.tab_index[OMDG_TAB_PAIR][i][j];
      a.tab[s] = -1;
      a.tcut2[s] = 0.f;
      if (tid < 0) continue;
      any_tab = true;
      int found = -1;
      for (int u = 0; u < used; u++)
        if (a.tabs[u].c32 == tab_view(c, tid).c32) found = u;
      if (found < 0) {
        a.tabs[used] = tab_view(c, tid);
        found = used++;
      }
      a.tab[s] = (signed char)found;
      const double cut = a.tabs[found].rcut;
      a.tcut2[s] = (float)(cut * cut);
    }
  a.sw_type = any_tab ? 0 : -1;
  const double pd = p.p, qd = p.q;
  a.pi = (pd == std::floor(pd) && pd >= 0 && pd <= 16) ? (int)pd : -1;
  a.qi = (qd == std::floor(qd) && qd >= 0 && qd <= 16) ? (int)qd : -1;
  near_arrays(c, kMaxNear);
  a.near = c.near;
  a.gsum = c.gsum;
  a.near_n = c.near_n;
  const int nb = grid_blocks(c.n, kBlock);
  {
    TimerScope ts(c, TG_PASS1);
    if (energy) k_sw_pair<true><<<nb, kBlock, 0, c.stream>>>(a);
    else k_sw_pair<false><<<nb, kBlock, 0, c.stream>>>(a);
  }
  {
    TimerScope ts(c, TG_FORCE);
    if (energy) {
      k_sw_centre<true><<<nb, kBlock, 0, c.stream>>>(a);
      k_sw_legs<true><<<nb, kBlock, 0, c.stream>>>(a);
    } else {
      k_sw_centre<false><<<nb, kBlock, 0, c.stream>>>(a);
      k_sw_legs<false><<<nb, kBlock, 0, c.stream>>>(a);
    }
  }
  OMDG_HIP(hipGetLastError());
}

void tersoff_forces_if(Ctx& c, bool energy, const int* only_if) {
  TBArgs a = base_args(c);
  a.only_if = only_if;
  a.ts = c.ters;
  a.rc = c.ters.R + c.ters.D;
  a.rc2 = a.rc * a.rc;
  const size_t need = (size_t)c.maxnbr * (size_t)c.n;
  if (need > c.cap_bond) {
    if (c.bond_b) (void)hipFree(c.bond_b);
    if (c.bond_p) (void)hipFree(c.bond_p);
    c.bond_b = c.bond_p = nullptr;
    OMDG_HIP(hipMalloc(reinterpret_cast<void**>(&c.bond_b), sizeof(float) * need));
    OMDG_HIP(hipMalloc(reinterpret_cast<void**>(&c.bond_p), sizeof(float) * need));
    c.cap_bond = need;
  }
  a.bond_b = c.bond_b;
  a.bond_p = c.bond_p;
  const int nb = grid_blocks(c.n, kBlock);
  {
    TimerScope ts(c, TG_PASS1);
    if (energy) k_tersoff_bond<true><<<nb, kBlock, 0, c.stream>>>(a);
    else k_tersoff_bond<false><<<nb, kBlock, 0, c.stream>>>(a);
  }
  {
    TimerScope ts(c, TG_FORCE);
    k_tersoff_force<<<nb, kBlock, 0, c.stream>>>(a);
  }
  OMDG_HIP(hipGetLastError());
}

void tersoff_forces(Ctx& c, bool energy) { tersoff_forces_if(c, energy, nullptr); }

}  // namespace omdg
