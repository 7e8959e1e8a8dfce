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
//
// LANES lanes serve one atom: lane j walks list entries j, j+LANES, ... with
// an FP32 partial sum (Kahan-compensated for the forces, plain for the
// densities, whose terms are all positive), and the partials are added in
// FP64 across the lanes.  Each lane first requests all its list words of a
// stretch (8-16), then the snapshots and table records of CHUNK entries
// before using any of them, so a wavefront keeps many independent loads in
// flight.  Every gather is unconditional (an invalid entry re-reads a valid
// slot and is masked afterwards): a conditional load would compile to a
// branch and a separate load per 32-bit component.  The force pass reads the
// snapshot and F'(rho) of a neighbour in one 16-B gather ({q.xyz, F'} written
// by the density pass).  The segment lookup is branch-free and the table
// range is checked once per lane from the extremes of its lookups.  When
// PAIR and EDENS share one grid their coefficients are interleaved per
// interval (32 B, one cache line per pair lookup).  (LANES, CHUNK) is picked
// per context by timing the candidates on the first evaluation.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "omdg_internal.h"
#include "spline_dev.h"

namespace omdg {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kNone = 0xFFFFFFFFu;

struct EamArgs {
  const float4* q;
  const uint32_t* nbr;
  const int* nnbr;
  float* rho;
  float4* qf;  // density pass: {q.xyz, F'(rho)} per atom for the force pass
  float* emb;
  float* f;  // 3 planes
  float* pot;
  float* vir;
  int* err;
  int n;
  float cut_e2, cut_p2;
  float hx, hy, hz;
  TabView phi, edens, embed;
  const float4* pack;  // {phi a,b,c,d}, {rho a,b,c,d} per interval, or null
};

// Per-lane partial in FP64, summed over the LANES lanes of one atom.
template <int LANES>
__device__ inline double lanes_sum(double v) {
#pragma unroll
  for (int o = 1; o < LANES; o <<= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Compensated partial -> FP64, summed over the LANES lanes of one atom.
template <int LANES>
__device__ inline double lanes_sum(const Kahan& k) {
  return lanes_sum<LANES>((double)k.s - (double)k.c);
}

// list words one lane requests before using any: a stretch of LANES * kPre
// entries costs one memory latency instead of one per CHUNK
template <int LANES, int CHUNK>
constexpr int pre_words() {
  constexpr int w = 64 / LANES > 16 ? 16 : 64 / LANES;
  return w < CHUNK ? CHUNK : w;
}

template <bool ENERGY, int LANES, int CHUNK>
__global__ void __launch_bounds__(kBlock) k_eam_density(EamArgs a) {
  constexpr int kPre = pre_words<LANES, CHUNK>();
  __shared__ float4 offs[32];  // codes 27..31 (the empty entry's 31) read zero
  load_offsets(offs, a.hx, a.hy, a.hz);
  if (threadIdx.x >= 27 && threadIdx.x < 32) offs[threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
  __syncthreads();
  const int g = xcd_block() * blockDim.x + threadIdx.x;
  const int s = g / LANES, j = g % LANES;
  const int n = a.n;
  const bool live = s < n;  // no early exit: all lanes take part in the shuffles
  const int sl = live ? s : 0;  // a valid slot for the unconditional loads
  const float4 qi = a.q[sl];
  const int nn = live ? a.nnbr[s] : 0;
  int err = 0;
  QRange qe;
  float rho = 0.f;  // positive terms: a plain FP32 partial per lane
  for (int b0 = j; b0 < nn; b0 += LANES * kPre) {
    // all list words of this stretch in flight at once
    uint32_t ew[kPre];
#pragma unroll
    for (int t = 0; t < kPre; t++) {
      const int k = b0 + t * LANES;
      // unconditional load (a repeat of the last entry past the end)
      const uint32_t w = a.nbr[nidx(sl, k < nn ? k : (nn > 0 ? nn - 1 : 0), n)];
      ew[t] = (k < nn) ? w : kNone;
    }
#pragma unroll
    for (int t0 = 0; t0 < kPre; t0 += CHUNK) {
      if (b0 + t0 * LANES >= nn) break;
      uint32_t e[CHUNK];
#pragma unroll
      for (int u = 0; u < CHUNK; u++) e[u] = ew[t0 + u];
      float4 qj[CHUNK];
#pragma unroll
      for (int u = 0; u < CHUNK; u++) qj[u] = a.q[e[u] != kNone ? (e[u] & kSlotMask) : sl];
      float r[CHUNK], dx[CHUNK];
      bool in[CHUNK];
      int seg[CHUNK];
#pragma unroll
      for (int u = 0; u < CHUNK; u++) {
        const float4 o = offs[e[u] >> kImgShift];
        const float ex = qi.x - qj[u].x - o.x;
        const float ey = qi.y - qj[u].y - o.y;
        const float ez = qi.z - qj[u].z - o.z;
        const float r2 = ex * ex + ey * ey + ez * ez;
        in[u] = (e[u] != kNone) && (r2 <= a.cut_e2);
        r[u] = in[u] ? r2 * __builtin_amdgcn_rsqf(r2) : a.edens.r0;
        seg[u] = spline_locate32_q(a.edens, r[u], dx[u], qe);
      }
      float4 c[CHUNK];
#pragma unroll
      for (int u = 0; u < CHUNK; u++) c[u] = a.edens.c32[seg[u]];
#pragma unroll
      for (int u = 0; u < CHUNK; u++) {
        float v = c[u].w + dx[u] * (c[u].z + dx[u] * (c[u].y + dx[u] * c[u].x));
        if (a.edens.rmult) v /= r[u];
        rho += in[u] ? v : 0.f;
      }
    }
  }
  err |= range_flags(a.edens, qe);
  if (err) atomicOr(a.err, err);
  const double rs = lanes_sum<LANES>((double)rho);
  if (live && j == 0) {
    a.rho[s] = (float)rs;
    // embedding function and derivative in FP64, once per atom
    double F, dF;
    spline_read64(a.embed, (double)(float)rs, F, dF, a.err);
    a.qf[s] = make_float4(qi.x, qi.y, qi.z, (float)dF);
    if (ENERGY) a.emb[s] = (float)F;
  }
}

template <bool ENERGY, bool PACKED, int LANES, int CHUNK>
__global__ void __launch_bounds__(kBlock) k_eam_force(EamArgs a) {
  constexpr int kPre = pre_words<LANES, CHUNK>();
  __shared__ float4 offs[32];  // codes 27..31 (the empty entry's 31) read zero
  load_offsets(offs, a.hx, a.hy, a.hz);
  if (threadIdx.x >= 27 && threadIdx.x < 32) offs[threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
  __syncthreads();
  const int g = xcd_block() * blockDim.x + threadIdx.x;
  const int s = g / LANES, j = g % LANES;
  const int n = a.n;
  const bool live = s < n;
  const int sl = live ? s : 0;  // a valid slot for the unconditional loads
  const float4 qi = a.qf[sl];
  const float fpi = qi.w;
  const int nn = live ? a.nnbr[s] : 0;
  int err = 0;
  QRange qp, qe;
  Kahan fx, fy, fz, pot, vir;
  for (int b0 = j; b0 < nn; b0 += LANES * kPre) {
    // all list words of this stretch in flight at once
    uint32_t ew[kPre];
#pragma unroll
    for (int t = 0; t < kPre; t++) {
      const int k = b0 + t * LANES;
      // unconditional load (a repeat of the last entry past the end)
      const uint32_t w = a.nbr[nidx(sl, k < nn ? k : (nn > 0 ? nn - 1 : 0), n)];
      ew[t] = (k < nn) ? w : kNone;
    }
#pragma unroll
    for (int t0 = 0; t0 < kPre; t0 += CHUNK) {
      if (b0 + t0 * LANES >= nn) break;
      uint32_t e[CHUNK];
#pragma unroll
      for (int u = 0; u < CHUNK; u++) e[u] = ew[t0 + u];
      float4 qj[CHUNK];
#pragma unroll
      for (int u = 0; u < CHUNK; u++) qj[u] = a.qf[e[u] != kNone ? (e[u] & kSlotMask) : sl];
      float d[CHUNK][3], r2[CHUNK], rinv[CHUNK], dxp[CHUNK], dxe[CHUNK];
      bool in[CHUNK], ine[CHUNK];
      int sp[CHUNK], se[CHUNK];
#pragma unroll
      for (int u = 0; u < CHUNK; u++) {
        const float4 o = offs[e[u] >> kImgShift];
        d[u][0] = qi.x - qj[u].x - o.x;
        d[u][1] = qi.y - qj[u].y - o.y;
        d[u][2] = qi.z - qj[u].z - o.z;
        r2[u] = d[u][0] * d[u][0] + d[u][1] * d[u][1] + d[u][2] * d[u][2];
        in[u] = (e[u] != kNone) && (r2[u] <= a.cut_p2);
        ine[u] = in[u] && (r2[u] <= a.cut_e2);
        rinv[u] = in[u] ? __builtin_amdgcn_rsqf(r2[u]) : 1.f;
        const float r = in[u] ? r2[u] * rinv[u] : a.phi.r0;
        sp[u] = spline_locate32_q(a.phi, r, dxp[u], qp);
        if (PACKED) {
          se[u] = sp[u];
          dxe[u] = dxp[u];
        } else {
          se[u] = spline_locate32_q(a.edens, ine[u] ? r : a.edens.r0, dxe[u], qe);
        }
      }
      float4 cp[CHUNK], ce[CHUNK];
#pragma unroll
      for (int u = 0; u < CHUNK; u++) {
        if (PACKED) {
          cp[u] = a.pack[2 * sp[u]];
          ce[u] = a.pack[2 * sp[u] + 1];
        } else {
          cp[u] = a.phi.c32[sp[u]];
          ce[u] = a.edens.c32[se[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < CHUNK; u++) {
        const float x = dxp[u];
        float p = cp[u].w + x * (cp[u].z + x * (cp[u].y + x * cp[u].x));
        float dp = cp[u].z + x * (2.0f * cp[u].y + 3.0f * x * cp[u].x);
        if (a.phi.rmult) {
          p *= rinv[u];
          dp = (dp - p) * rinv[u];
        }
        const float y = dxe[u];
        float drho = ce[u].z + y * (2.0f * ce[u].y + 3.0f * y * ce[u].x);
        if (a.edens.rmult) {
          const float v = (ce[u].w + y * (ce[u].z + y * (ce[u].y + y * ce[u].x))) * rinv[u];
          drho = (drho - v) * rinv[u];
        }
        drho = ine[u] ? drho : 0.f;  // GetRhoDeriv beyond the EDENS range (team.cpp:130)
        const float fr = in[u] ? -((fpi + qj[u].w) * drho + dp) * rinv[u] : 0.f;
        fx.add(fr * d[u][0]);
        fy.add(fr * d[u][1]);
        fz.add(fr * d[u][2]);
        if (ENERGY) {
          pot.add(in[u] ? 0.5f * p : 0.f);
          vir.add(fr * r2[u]);
        }
      }
    }
  }
  err |= range_flags(a.phi, qp);
  if (!PACKED) err |= range_flags(a.edens, qe);
  if (err) atomicOr(a.err, err);
  const double sx = lanes_sum<LANES>(fx), sy = lanes_sum<LANES>(fy), sz = lanes_sum<LANES>(fz);
  double sp = 0.0, sv = 0.0;
  if (ENERGY) {
    sp = lanes_sum<LANES>(pot);
    sv = lanes_sum<LANES>(vir);
  }
  if (live && j == 0) {
    a.f[s] = (float)sx;
    a.f[(size_t)n + s] = (float)sy;
    a.f[2 * (size_t)n + s] = (float)sz;
    if (ENERGY) {
      a.pot[s] = (float)(sp + (double)a.emb[s]);
      a.vir[s] = (float)sv;
    }
  }
}

// Interleave the PAIR and EDENS coefficients when they share one grid.
void build_pack(Ctx& c, int tp, int te) {
  c.eam_pack_ok = true;
  const DevTable& P = c.tables[tp];
  const DevTable& E = c.tables[te];
  if (c.eam_pack) (void)hipFree(c.eam_pack);
  c.eam_pack = nullptr;
  if (P.n != E.n || P.dr != E.dr || P.r0 != E.r0) return;
  std::vector<float4> pk(2 * (size_t)P.n);
  for (int i = 0; i < P.n; i++) {
    const double* p = &c.host_c64[4 * ((size_t)P.off + i)];
    const double* e = &c.host_c64[4 * ((size_t)E.off + i)];
    pk[2 * i] = make_float4((float)p[0], (float)p[1], (float)p[2], (float)p[3]);
    pk[2 * i + 1] = make_float4((float)e[0], (float)e[1], (float)e[2], (float)e[3]);
  }
  OMDG_HIP(hipMalloc(reinterpret_cast<void**>(&c.eam_pack), sizeof(float4) * pk.size()));
  OMDG_HIP(hipMemcpy(c.eam_pack, pk.data(), sizeof(float4) * pk.size(), hipMemcpyHostToDevice));
}

// Candidate (LANES, CHUNK) shapes of the two passes.
struct Shape {
  int lanes, chunk;
};
constexpr Shape kShapes[] = {{4, 2}, {4, 4}, {8, 1}, {8, 2}, {2, 4}, {16, 1}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

template <int L, int C>
void launch_shape(Ctx& c, const EamArgs& a, bool energy, bool pass1, bool pass2) {
  const int nb = grid_blocks((long)c.n * L, kBlock);
  if (pass1) {
    TimerScope ts(c, TG_PASS1);
    if (energy) k_eam_density<true, L, C><<<nb, kBlock, 0, c.stream>>>(a);
    else k_eam_density<false, L, C><<<nb, kBlock, 0, c.stream>>>(a);
  }
  if (pass2) {
    TimerScope ts(c, TG_FORCE);
    if (a.pack) {
      if (energy) k_eam_force<true, true, L, C><<<nb, kBlock, 0, c.stream>>>(a);
      else k_eam_force<false, true, L, C><<<nb, kBlock, 0, c.stream>>>(a);
    } else {
      if (energy) k_eam_force<true, false, L, C><<<nb, kBlock, 0, c.stream>>>(a);
      else k_eam_force<false, false, L, C><<<nb, kBlock, 0, c.stream>>>(a);
    }
  }
}

void launch(Ctx& c, const EamArgs& a, bool energy, int shape, bool pass1 = true,
            bool pass2 = true) {
  switch (shape) {
    case 0: launch_shape<4, 2>(c, a, energy, pass1, pass2); break;
    case 1: launch_shape<4, 4>(c, a, energy, pass1, pass2); break;
    case 2: launch_shape<8, 1>(c, a, energy, pass1, pass2); break;
    case 3: launch_shape<8, 2>(c, a, energy, pass1, pass2); break;
    case 4: launch_shape<2, 4>(c, a, energy, pass1, pass2); break;
    default: launch_shape<16, 1>(c, a, energy, pass1, pass2); break;
  }
}

// Time every shape of one pass (three launches each, after one warm-up) on
// the current list and keep the fastest.  Both passes recompute their
// outputs from their inputs, so the extra launches change nothing.
int tune_pass(Ctx& c, const EamArgs& a, bool pass1) {
  hipEvent_t e0, e1;
  OMDG_HIP(hipEventCreate(&e0));
  OMDG_HIP(hipEventCreate(&e1));
  const bool was_timing = c.timing;
  c.timing = false;
  int best = 0;
  float best_ms = 1e30f;
  for (int v = 0; v < kNumShapes; v++) {
    launch(c, a, false, v, pass1, !pass1);
    OMDG_HIP(hipEventRecord(e0, c.stream));
    for (int r = 0; r < 3; r++) launch(c, a, false, v, pass1, !pass1);
    OMDG_HIP(hipEventRecord(e1, c.stream));
    OMDG_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    OMDG_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) {
      best_ms = ms;
      best = v;
    }
  }
  c.timing = was_timing;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

}  // namespace

void eam_forces(Ctx& c, bool energy) {
  EamArgs a;
  a.q = c.q;
  a.nbr = c.nbr;
  a.nnbr = c.nnbr;
  a.rho = c.rho;
  a.qf = c.qf;
  a.emb = c.emb;
  a.f = c.f32;
  a.pot = c.pot32;
  a.vir = c.vir32;
  a.err = c.dstat + ST_ERR;
  a.n = c.n;
  const int tp = c.tab_index[OMDG_TAB_PAIR][0][0];
  const int te = c.tab_index[OMDG_TAB_EDENS][0][0];
  a.phi = tab_view(c, tp);
  a.edens = tab_view(c, te);
  a.embed = tab_view(c, c.tab_index[OMDG_TAB_EMBED][0][0]);
  if (!c.eam_pack_ok) build_pack(c, tp, te);
  a.pack = c.eam_pack;
  // cut radii: table ranges (team.cpp:78, team.cpp:212) up to the end of
  // their non-zero part
  const double ce = a.edens.rcut;
  const double cp = a.phi.rcut;
  a.cut_e2 = (float)(ce * ce);
  a.cut_p2 = (float)(cp * cp);
  a.hx = (float)c.grid.h[0];
  a.hy = (float)c.grid.h[1];
  a.hz = (float)c.grid.h[2];
  if (c.eam_shape[0] < 0) {
    // small systems keep the default; large ones time the candidates
    if (c.n >= 16384) {
      launch(c, a, false, 0, true, false);  // F'(rho) for the force-pass tuning
      c.eam_shape[0] = tune_pass(c, a, true);
      c.eam_shape[1] = tune_pass(c, a, false);
      if (std::getenv("OMDG_VERBOSE"))
        std::fprintf(stderr, "[omdg] EAM shapes: density lanes=%d chunk=%d, force lanes=%d chunk=%d\n",
                     kShapes[c.eam_shape[0]].lanes, kShapes[c.eam_shape[0]].chunk,
                     kShapes[c.eam_shape[1]].lanes, kShapes[c.eam_shape[1]].chunk);
    } else {
      c.eam_shape[0] = c.eam_shape[1] = 0;
    }
  }
  launch(c, a, energy, c.eam_shape[0], true, false);
  launch(c, a, energy, c.eam_shape[1], false, true);
  OMDG_HIP(hipGetLastError());
}

}  // namespace omdg
