// Reverse pass of one WCSPH substep on gfx950: the hand-written adjoint of
// the density sum, the Tait pressure, the pressure and artificial-viscosity
// pair forces, the wall pressure and wall viscosity, and the symplectic
// Euler advance with its container clamp (src/fluid/step.rs:144-224,
// 262-339; src/fluid/mod.rs:621-675).
//
// The forward substep, for particle i in the substep's sorted order:
//   rho_i = m (sum_j W(|o_ij|) + sum_w W(|o_iw|)),  o = x_i - x_j
//   p_i = max(B (min(rho_i / rho0, 1.1)^7 - 1), 0),  s_i = p_i / rho_i^2
//   a_i = sum_{j != i} -g_ij m (s_i + s_j + D_ij)
//         + sum_w -g_iw m (s_i + p_i / rho0^2) + wall viscosity + gravity
//   g = o f(d), f = W'(d) / d = -20 sigma (1 - q)^3 / r^2
//   D_ij = c_ij min(u_ij, 0), c_ij = -nu 2 / ((d^2 + eps)(rho_i + rho_j)),
//   u_ij = (v_i - v_j) . o_ij
//   v' = v + a dt, x' = x + v' dt, clamped components pinned (zero adjoint).
// Pair terms are antisymmetric (F_ji = -F_ij with the same K), so each
// particle gathers its whole adjoint from its own neighbourhood using the
// acceleration adjoints of both ends of each pair: no atomics, and the
// neighbourhood is the forward's (same cells, same |o|^2 < r^2 test).
//
// The arithmetic is FP32 like the forward; the viscosity-coefficient
// gradient is summed per block in FP64.
#include "adjoint.hpp"

namespace sfl {

// The walls near `at`: through the searched table when KEYED, else the
// dense one given as its parts (see visit_dense).
template <bool KEYED, class F>
__device__ __forceinline__ void adj_walls(const AdjointArgs& a, float4 at, F&& f) {
  if constexpr (KEYED) {
    const GridView g{{a.wall_box.lo[0], a.wall_box.lo[1], a.wall_box.lo[2]},
                     {a.wall_box.span[0], a.wall_box.span[1], a.wall_box.span[2]},
                     a.wall_start,
                     a.wall_keys,
                     a.wall_keyed};
    visit_near(a.walls, g, at, a.radius, a.reach, f);
  } else {
    visit_dense(a.walls, a.wall_start, a.wall_box.lo, a.wall_box.span, at, a.radius, a.reach, f);
  }
}

namespace {

struct Kern {
  float f;    // W'(d) / d
  float fpd;  // (d/dd (W'(d) / d)) / d
};

// f and f'/d of the Wendland C2 kernel at distance d > 0 (q = d / r < 1).
__device__ __forceinline__ Kern kern(float d, float radius, float scale) {
  float q = d / radius;
  Kern k{0.0f, 0.0f};
  if (q >= 1.0f) return k;
  float t = 1.0f - q;
  float r2 = radius * radius;
  k.f = -20.0f * scale * (t * t * t) / r2;
  k.fpd = 60.0f * scale * (t * t) / (r2 * radius * d);
  return k;
}

__device__ __forceinline__ float3 f3(float4 a) { return make_float3(a.x, a.y, a.z); }
__device__ __forceinline__ float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 operator-(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 operator*(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }

// Tait pressure and its slope dp/drho (0 where the cap or the max(.., 0)
// holds).
__device__ __forceinline__ void tait(float rho, float rest, float stiffness, float& p, float& dp) {
  float ratio = fminf(rho / rest, 1.1f);
  float r2 = ratio * ratio;
  float r6 = r2 * r2 * r2;
  p = fmaxf(stiffness * (ratio * r6 - 1.0f), 0.0f);
  dp = (p > 0.0f && rho / rest < 1.1f) ? 7.0f * stiffness * r6 / rest : 0.0f;
}

}  // namespace

// Adjoints of the final state, given per particle id, into the world's
// current (sorted) order.
__global__ __launch_bounds__(256) void k_adj_load(uint32_t n, const float4* __restrict__ P,
                                                  const float* __restrict__ xbar, const float* __restrict__ vbar,
                                                  float4* __restrict__ Xb, float4* __restrict__ Vb) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t id = id_of(P[i]);
  Xb[i] = make_float4(xbar[3 * id], xbar[3 * id + 1], xbar[3 * id + 2], 0.0f);
  Vb[i] = make_float4(vbar[3 * id], vbar[3 * id + 1], vbar[3 * id + 2], 0.0f);
}

// Adjoints of the recorded initial state back to id order.
__global__ __launch_bounds__(256) void k_adj_store(uint32_t n, const float4* __restrict__ P,
                                                   const float4* __restrict__ Xb, const float4* __restrict__ Vb,
                                                   float* __restrict__ xbar, float* __restrict__ vbar) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t id = id_of(P[i]);
  float4 x = Xb[i], v = Vb[i];
  xbar[3 * id] = x.x;
  xbar[3 * id + 1] = x.y;
  xbar[3 * id + 2] = x.z;
  vbar[3 * id] = v.x;
  vbar[3 * id + 1] = v.y;
  vbar[3 * id + 2] = v.z;
}

// R1: advance and velocity update. From the adjoints of the substep's
// output (x_out, v_out), the adjoints of x, v through x' = x + v' dt,
// v' = v + a dt, and the acceleration adjoint a_bar = v'_bar dt; clamped
// components are constants of the output.
__global__ __launch_bounds__(256) void k_adj_advance(AdjointArgs a, const float4* __restrict__ XbOut,
                                                     const float4* __restrict__ VbOut, float4* __restrict__ Xb,
                                                     float4* __restrict__ Vb, float4* __restrict__ Ab) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  uint32_t c = a.clamps[i];
  float4 xo = XbOut[i], vo = VbOut[i];
  float xb[3] = {xo.x, xo.y, xo.z};
  float vb[3] = {vo.x, vo.y, vo.z};
  for (int k = 0; k < 3; ++k) {
    if (c & (1u << k)) {
      xb[k] = 0.0f;
      vb[k] = 0.0f;
    } else {
      vb[k] = vb[k] + xb[k] * a.dt;
    }
  }
  Xb[i] = make_float4(xb[0], xb[1], xb[2], 0.0f);
  Vb[i] = make_float4(vb[0], vb[1], vb[2], 0.0f);
  Ab[i] = make_float4(vb[0] * a.dt, vb[1] * a.dt, vb[2] * a.dt, 0.0f);
}

// R2: the force loop. Adds to x_bar, v_bar the pair, wall-pressure and
// wall-viscosity terms, writes rho_bar (through s = p / rho^2 and the
// wall's p / rho0^2 as well), and folds the viscosity coefficient's
// adjoint times d nu / d alpha into nu_acc[block] (FP64).
template <bool KEYED>
__global__ __launch_bounds__(256) void k_adj_force(AdjointArgs a, const float4* __restrict__ Ab,
                                                   float4* __restrict__ Xb, float4* __restrict__ Vb,
                                                   float* __restrict__ Rb, double* __restrict__ nu_acc) {
  __shared__ double part[4];
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double nub = 0.0;
  if (i < a.n) {
    const float m = a.mass, r = a.radius, sc = a.scale;
    CellBox box = *a.box;
    float4 pi4 = a.P[i], vi4 = a.V[i];
    float3 vi = f3(vi4);
    float rho = vi4.w;
    float si = a.scaled[i];
    float3 ai = f3(Ab[i]);
    float3 xb = f3(Xb[i]), vb = f3(Vb[i]);
    float sbar = 0.0f, pbar = 0.0f, rbar = 0.0f, nsum = 0.0f;
    visit_dense(a.P, a.cell_start, box.lo, box.span, pi4, r, a.reach,
                [&](uint32_t j, float4 pj, float ox, float oy, float oz, float sq) {
                  if (sq <= 0.0f) return;
                  float d = sqrtf(sq);
                  Kern k = kern(d, r, sc);
                  float3 o = make_float3(ox, oy, oz);
                  float3 g = o * k.f;
                  float4 vj4 = a.V[j];
                  float3 dv = vi - f3(vj4);
                  float u = dot(dv, o);
                  float t = fminf(u, 0.0f);
                  float den = sq + a.soften;
                  float rsum = rho + vj4.w;
                  float cd = -a.nu * 2.0f / (den * rsum);
                  float D = cd * t;
                  float K = m * (si + a.scaled[j] + D);
                  float3 delta = f3(Ab[j]) - ai;
                  float G = dot(g, delta);
                  float od = dot(o, delta);
                  sbar += m * G;
                  rbar += m * G * (-D / rsum);
                  float3 dx = o * (m * G * (-2.0f * D / den)) + (delta * k.f + o * (k.fpd * od)) * K;
                  if (u < 0.0f) {
                    vb = vb + o * (m * cd * G);
                    dx = dx + dv * (m * G * cd);
                  }
                  xb = xb + dx;
                  nsum += m * dot(g, ai) * (2.0f * t / (den * rsum));
                });
    if (a.wall_cells) {
      float p, dp;
      tait(rho, a.rest, a.stiffness, p, dp);
      float S = si + p / (a.rest * a.rest);
      auto walls = [&](auto&& f) { adj_walls<KEYED>(a, pi4, f); };
      // forward: away, normal, the velocity into the wall
      float3 away = make_float3(0.0f, 0.0f, 0.0f);
      walls([&](uint32_t, float4, float ox, float oy, float oz, float sq) {
        if (sq <= 0.0f) return;
        Kern k = kern(sqrtf(sq), r, sc);
        away = away - make_float3(ox, oy, oz) * k.f;
      });
      float len = sqrtf(dot(away, away));
      float rcp = 1.0f / len;
      bool valid = isfinite(rcp) && rcp > 0.0f;
      float3 nrm = valid ? away * rcp : make_float3(0.0f, 0.0f, 0.0f);
      float vn = dot(vi, nrm);
      float tau = fminf(vn, 0.0f);
      float3 U = nrm * tau;
      float3 ub = make_float3(0.0f, 0.0f, 0.0f);
      walls([&](uint32_t, float4, float ox, float oy, float oz, float sq) {
        if (sq <= 0.0f) return;
        float d = sqrtf(sq);
        Kern k = kern(d, r, sc);
        float3 o = make_float3(ox, oy, oz);
        float3 g = o * k.f;
        float ga = dot(g, ai), oa = dot(o, ai);
        float3 hq = ai * k.f + o * (k.fpd * oa);  // (dg/do)^T a_bar
        sbar -= m * ga;
        pbar -= m * ga / (a.rest * a.rest);
        xb = xb - hq * (m * S);
        float omega = dot(U, o);
        if (omega < 0.0f) {
          float denw = sq + a.soften;
          float kappa = -a.wall_nu / (denw * rho);
          rbar += m * kappa * omega * ga / rho;
          ub = ub - o * (m * kappa * ga);
          xb = xb - (o * (-2.0f * kappa * omega * ga / denw) + U * (kappa * ga) + hq * (kappa * omega)) * m;
        }
      });
      if (valid) {
        float taub = dot(ub, nrm);
        float3 nb = ub * tau;
        if (vn < 0.0f) {
          nb = nb + vi * taub;
          vb = vb + nrm * taub;
        }
        float3 awb = (nb - nrm * dot(nrm, nb)) * rcp;
        walls([&](uint32_t, float4, float ox, float oy, float oz, float sq) {
          if (sq <= 0.0f) return;
          Kern k = kern(sqrtf(sq), r, sc);
          float3 o = make_float3(ox, oy, oz);
          xb = xb - (awb * k.f + o * (k.fpd * dot(o, awb)));
        });
      }
    }
    float p, dp;
    tait(rho, a.rest, a.stiffness, p, dp);
    rbar += sbar * (dp / (rho * rho) - 2.0f * p / (rho * rho * rho)) + pbar * dp;
    Xb[i] = make_float4(xb.x, xb.y, xb.z, 0.0f);
    Vb[i] = make_float4(vb.x, vb.y, vb.z, 0.0f);
    Rb[i] = rbar;
    nub = (double)nsum;
  }
  for (int off = 32; off >= 1; off >>= 1) nub += __shfl_xor(nub, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = nub;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) s += part[w];
    nu_acc[blockIdx.x] += s * (double)a.dnu_dalpha;
  }
}

// R3: the density sum. x_bar_i += m sum_j g_ij (rho_bar_i + rho_bar_j) +
// m sum_w g_iw rho_bar_i, then both adjoints go back through the sort to
// the previous substep's order.
template <bool KEYED>
__global__ __launch_bounds__(256) void k_adj_density(AdjointArgs a, const float* __restrict__ Rb,
                                                     const float4* __restrict__ Xb, const float4* __restrict__ Vb,
                                                     float4* __restrict__ XbPrev, float4* __restrict__ VbPrev) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const float m = a.mass, r = a.radius, sc = a.scale;
  CellBox box = *a.box;
  float4 pi4 = a.P[i];
  float rbi = Rb[i];
  float3 xb = f3(Xb[i]);
  visit_dense(a.P, a.cell_start, box.lo, box.span, pi4, r, a.reach,
              [&](uint32_t j, float4, float ox, float oy, float oz, float sq) {
                if (sq <= 0.0f) return;
                Kern k = kern(sqrtf(sq), r, sc);
                xb = xb + make_float3(ox, oy, oz) * (m * k.f * (rbi + Rb[j]));
              });
  if (a.wall_cells)
    adj_walls<KEYED>(a, pi4, [&](uint32_t, float4, float ox, float oy, float oz, float sq) {
      if (sq <= 0.0f) return;
      Kern k = kern(sqrtf(sq), r, sc);
      xb = xb + make_float3(ox, oy, oz) * (m * k.f * rbi);
    });
  uint32_t dst = a.src[i];
  XbPrev[dst] = make_float4(xb.x, xb.y, xb.z, 0.0f);
  VbPrev[dst] = Vb[i];
}

// Counts particles whose replayed (x, id, v, rho) differ in any bit from
// the recorded state.
__global__ __launch_bounds__(256) void k_adj_compare(uint32_t n, const float4* __restrict__ P,
                                                     const float4* __restrict__ V, const float4* __restrict__ P2,
                                                     const float4* __restrict__ V2, uint32_t* __restrict__ count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool differs = false;
  if (i < n) {
    uint4 a = *reinterpret_cast<const uint4*>(P + i), b = *reinterpret_cast<const uint4*>(P2 + i);
    uint4 c = *reinterpret_cast<const uint4*>(V + i), d = *reinterpret_cast<const uint4*>(V2 + i);
    differs = a.x != b.x || a.y != b.y || a.z != b.z || a.w != b.w || c.x != d.x || c.y != d.y || c.z != d.z ||
              c.w != d.w;
  }
  wave_sum_to(count, differs ? 1u : 0u);
}

// Copies the cell box the grid setup chose into the tape.
__global__ void k_adj_keep_box(const Control* ctl, CellBox* box) {
  if (threadIdx.x != 0) return;
  for (int k = 0; k < 3; ++k) {
    box->lo[k] = ctl->lo[k];
    box->span[k] = ctl->span[k];
  }
}

static inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

void launch_adj_load(hipStream_t s, uint32_t n, const float4* P, const float* xbar, const float* vbar, float4* Xb,
                     float4* Vb) {
  if (n) hipLaunchKernelGGL(k_adj_load, blocks_for(n, 256), 256, 0, s, n, P, xbar, vbar, Xb, Vb);
}
void launch_adj_store(hipStream_t s, uint32_t n, const float4* P, const float4* Xb, const float4* Vb, float* xbar,
                      float* vbar) {
  if (n) hipLaunchKernelGGL(k_adj_store, blocks_for(n, 256), 256, 0, s, n, P, Xb, Vb, xbar, vbar);
}
void launch_adj_advance(hipStream_t s, const AdjointArgs& a, const float4* XbOut, const float4* VbOut, float4* Xb,
                        float4* Vb, float4* Ab) {
  if (a.n) hipLaunchKernelGGL(k_adj_advance, blocks_for(a.n, 256), 256, 0, s, a, XbOut, VbOut, Xb, Vb, Ab);
}
uint32_t adj_force_blocks(uint32_t n) { return blocks_for(n, 256); }
void launch_adj_force(hipStream_t s, const AdjointArgs& a, const float4* Ab, float4* Xb, float4* Vb, float* Rb,
                      double* nu_acc) {
  if (!a.n) return;
  if (a.wall_keys)
    hipLaunchKernelGGL(k_adj_force<true>, blocks_for(a.n, 256), 256, 0, s, a, Ab, Xb, Vb, Rb, nu_acc);
  else
    hipLaunchKernelGGL(k_adj_force<false>, blocks_for(a.n, 256), 256, 0, s, a, Ab, Xb, Vb, Rb, nu_acc);
}
void launch_adj_density(hipStream_t s, const AdjointArgs& a, const float* Rb, const float4* Xb, const float4* Vb,
                        float4* XbPrev, float4* VbPrev) {
  if (!a.n) return;
  if (a.wall_keys)
    hipLaunchKernelGGL(k_adj_density<true>, blocks_for(a.n, 256), 256, 0, s, a, Rb, Xb, Vb, XbPrev, VbPrev);
  else
    hipLaunchKernelGGL(k_adj_density<false>, blocks_for(a.n, 256), 256, 0, s, a, Rb, Xb, Vb, XbPrev, VbPrev);
}
void launch_adj_compare(hipStream_t s, uint32_t n, const float4* P, const float4* V, const float4* P2,
                        const float4* V2, uint32_t* count) {
  if (n) hipLaunchKernelGGL(k_adj_compare, blocks_for(n, 256), 256, 0, s, n, P, V, P2, V2, count);
}
void launch_adj_keep_box(hipStream_t s, const Control* ctl, CellBox* box) {
  hipLaunchKernelGGL(k_adj_keep_box, 1, 64, 0, s, ctl, box);
}

}  // namespace sfl
