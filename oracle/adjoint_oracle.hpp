// TEST INFRASTRUCTURE, NOT PRODUCT CODE: the reverse pass of the oracle's
// WCSPH rollout. AdjointTape<R, Q> records a World<R> and runs the reverse
// pass in precision Q: <float, double> is the FP64 gradient of the
// reference's f32 rollout (the device's FP32 trajectory, bit for bit), the
// reference the device adjoint is checked against; <double, double> the
// gradient of the FP64 rollout, which central differences pin.
//
// Recording hooks World<R>::wcsph around every wcsph_substep
// (src/fluid/step.rs:144-224): before it, the state in the previous order;
// after it, the substep's sort (ids), neighbour lists, densities, pushes and
// output. From those the tape keeps, per substep, the permutation of the
// sort, the sorted input state, the densities, the lists and which
// components the container clamp pinned (advance, step.rs:305-339).
//
// The reverse pass is the adjoint of, for particle i of a substep,
//   rho_i = m (sum_j W(|o_ij|) + sum_w W(|o_iw|)),  o = x_i - x_j
//   p_i = max(B (min(rho_i / rho0, 1.1)^7 - 1), 0),  s_i = p_i / rho_i^2
//   a_i = sum_{j != i} -g_ij m (s_i + s_j + D_ij) + sum_w -g_iw m (s_i + p_i / rho0^2)
//         + wall viscosity (step.rs:262-301) + gravity
//   g = o W'(d) / d,  D_ij = -nu min((v_i - v_j) . o, 0) / (d^2 + eps) 2 / (rho_i + rho_j)
//   v' = v + a dt,  x' = x + v' dt,  clamped components pinned.
// Each particle gathers its adjoint over its own lists (pair terms are
// antisymmetric), as the device kernels do; it is written independently of
// them, from the same derivation, and is pinned by central differences of
// the forward (tests/test_oracle_adjoint.py).
//
// Planted mistakes (for the tests that show the difference checks bite):
//   4: drop the viscosity term's density dependence
//   5: drop the wall-viscosity normal's dependence on position
#pragma once

#include "fluid_oracle.hpp"

namespace oracle {

template <class R, class Q = R>
class AdjointTape {
 public:
  using V = Vec3<Q>;
  using VR = Vec3<R>;
  static V up(VR a) { return {Q(a.x), Q(a.y), Q(a.z)}; }

  struct Substep {
    std::vector<uint32_t> src;  // sorted index -> index before the sort
    std::vector<V> x, v;        // sorted input state
    std::vector<Q> rho;
    Lists fluid, walls;
    std::vector<uint8_t> clamps;
    Q dt, nu, wall_nu, dnu;
  };

  World<R>& world;
  std::vector<Substep> substeps;
  std::vector<uint32_t> ids0;  // order at recording start
  std::vector<VR> pre_x, pre_v;
  std::vector<uint32_t> pre_ids;

  explicit AdjointTape(World<R>& w) : world(w) {}

  // `steps` WCSPH steps of dt, recorded.
  int32_t record(float dt, uint32_t steps) {
    substeps.clear();
    ids0 = world.ids;
    world.substep_hook = [this](bool after, R small) {
      if (!after) {
        pre_x = world.positions;
        pre_v = world.velocities;
        pre_ids = world.ids;
      } else {
        keep(small);
      }
    };
    int32_t st = kOk;
    for (uint32_t k = 0; k < steps && st == kOk; ++k) {
      Report r;
      st = world.step(dt, r);
    }
    world.substep_hook = nullptr;
    return st;
  }

  // Reverse pass. xbar, vbar: per id; outputs per id; returns dL/d(viscosity).
  Q reverse(const std::vector<V>& xbar_id, const std::vector<V>& vbar_id, std::vector<V>& xbar0_id,
            std::vector<V>& vbar0_id) {
    size_t n = world.len();
    std::vector<V> xb(n), vb(n);
    for (size_t i = 0; i < n; ++i) {
      xb[i] = xbar_id[world.ids[i]];
      vb[i] = vbar_id[world.ids[i]];
    }
    Q alpha_bar = Q(0);
    for (size_t s = substeps.size(); s-- > 0;) alpha_bar += reverse_substep(substeps[s], xb, vb);
    xbar0_id.assign(n, V{});
    vbar0_id.assign(n, V{});
    for (size_t i = 0; i < n; ++i) {
      xbar0_id[ids0[i]] = xb[i];
      vbar0_id[ids0[i]] = vb[i];
    }
    return alpha_bar;
  }

 private:
  void keep(R small) {
    const World<R>& w = world;
    size_t n = w.len();
    Substep sub;
    std::vector<uint32_t> pos(n);
    for (size_t i = 0; i < n; ++i) pos[pre_ids[i]] = (uint32_t)i;
    sub.src.resize(n);
    sub.x.resize(n);
    sub.v.resize(n);
    sub.rho.resize(n);
    sub.clamps.assign(n, 0);
    for (size_t i = 0; i < n; ++i) {
      uint32_t k = pos[w.ids[i]];
      sub.src[i] = k;
      VR x = pre_x[k], v = pre_v[k];
      sub.x[i] = up(x);
      sub.v[i] = up(v);
      sub.rho[i] = Q(w.densities[i]);
      // the advance's clamp decisions, recomputed in the forward's
      // arithmetic as advance does them
      if (w.walls.bounds) {
        R half = R(w.settings.spacing / 2.0f);
        VR rest{half, half, half};
        VR low = w.walls.bounds->first, high = w.walls.bounds->second;
        VR velocity = v + w.forces[i] * small;
        VR next = x + velocity * small;
        VR inside = clamp(next, low + rest, high - rest);
        uint8_t c = 0;
        if (inside.x != next.x) c |= 1;
        if (inside.y != next.y) c |= 2;
        if (inside.z != next.z) c |= 4;
        sub.clamps[i] = c;
      }
    }
    sub.fluid = w.neighbors;
    sub.walls = w.wall_neighbors;
    R r = w.radius();
    R cs = R(w.settings.sound_speed);
    R free = R(w.settings.viscosity) * cs * r, cap = R(0.05f) * r * r / small;
    sub.dt = Q(small);
    sub.nu = Q(w.viscosity(small));
    sub.wall_nu = Q(rmin(R(1) * cs * r, R(0.25f) * (r * r / small)));
    sub.dnu = free < cap ? Q(cs) * Q(r) : Q(0);
    substeps.push_back(std::move(sub));
  }

  // W'(d)/d and (d/dd (W'(d)/d))/d for 0 < d < r
  void kern(Q d, Q& f, Q& fpd) const {
    Q r = Q(world.radius()), sc = Q(world.kernel.scale);
    Q q = d / r;
    if (q >= Q(1)) {
      f = fpd = Q(0);
      return;
    }
    Q t = Q(1) - q;
    f = Q(-20) * sc * t * t * t / (r * r);
    fpd = Q(60) * sc * t * t / (r * r * r * d);
  }

  void tait(Q rho, Q& p, Q& dp) const {
    Q rest = Q(world.settings.rest_density), c = Q(world.settings.sound_speed);
    Q B = rest * (c * c) / Q(7);
    Q ratio = rmin(rho / rest, Q(1.1f));
    Q r6 = ratio * ratio * ratio * ratio * ratio * ratio;
    p = rmax(B * (ratio * r6 - Q(1)), Q(0));
    dp = (p > Q(0) && rho / rest < Q(1.1f)) ? Q(7) * B * r6 / rest : Q(0);
  }

  // One substep backwards; xb, vb: adjoints of its output in its order, on
  // return of its input in the previous order. Returns its dL/d alpha.
  Q reverse_substep(const Substep& S, std::vector<V>& xb, std::vector<V>& vb) {
    const World<R>& w = world;
    size_t n = S.x.size();
    Q m = Q(w.mass), rest = Q(w.settings.rest_density);
    Q eps = Q(R(0.01f) * w.radius() * w.radius());  // soften, in the forward's arithmetic
    std::vector<V> walls(w.walls.positions.size());
    for (size_t k = 0; k < walls.size(); ++k) walls[k] = up(w.walls.positions[k]);
    // advance and velocity update
    std::vector<V> X(n), Vv(n), A(n);
    for (size_t i = 0; i < n; ++i) {
      Q x[3] = {xb[i].x, xb[i].y, xb[i].z}, v[3] = {vb[i].x, vb[i].y, vb[i].z};
      for (int c = 0; c < 3; ++c) {
        if (S.clamps[i] & (1 << c)) {
          x[c] = Q(0);
          v[c] = Q(0);
        } else {
          v[c] = v[c] + x[c] * S.dt;
        }
      }
      X[i] = {x[0], x[1], x[2]};
      Vv[i] = {v[0], v[1], v[2]};
      A[i] = Vv[i] * S.dt;
    }
    std::vector<Q> scaled(n), P(n), dP(n), rb(n), nus(n);
    for (size_t i = 0; i < n; ++i) {
      tait(S.rho[i], P[i], dP[i]);
      scaled[i] = P[i] / (S.rho[i] * S.rho[i]);
    }
    // force loop
    world.dispatch.run(n, [&](size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) {
        V xi = S.x[i], vi = S.v[i], ai = A[i];
        Q rho = S.rho[i], si = scaled[i];
        V xbi = X[i], vbi = Vv[i];
        Q sbar = 0, pbar = 0, rbar = 0, nsum = 0;
        for (uint32_t k = S.fluid.start[i]; k < S.fluid.start[i + 1]; ++k) {
          uint32_t j = S.fluid.items[k];
          if (j == i) continue;
          V o = xi - S.x[j];
          Q sq = length_squared(o);
          if (sq <= Q(0)) continue;
          Q d = std::sqrt(sq), f, fpd;
          kern(d, f, fpd);
          V g = o * f;
          V dv = vi - S.v[j];
          Q u = dot(dv, o);
          Q t = rmin(u, Q(0));
          Q den = sq + eps, rsum = rho + S.rho[j];
          Q cd = -S.nu * Q(2) / (den * rsum);
          Q D = cd * t;
          Q K = m * (si + scaled[j] + D);
          V delta = A[j] - ai;
          Q G = dot(g, delta), od = dot(o, delta);
          sbar += m * G;
          if (planted_mistake() != 4) rbar += m * G * (-D / rsum);
          V dx = o * (m * G * (Q(-2) * D / den)) + (delta * f + o * (fpd * od)) * K;
          if (u < Q(0)) {
            vbi = vbi + o * (m * cd * G);
            dx = dx + dv * (m * G * cd);
          }
          xbi = xbi + dx;
          nsum += m * dot(g, ai) * (Q(2) * t / (den * rsum));
        }
        if (!walls.empty() && S.walls.start.size() == n + 1) {
          Q S_w = si + P[i] / (rest * rest);
          V away{};
          for (uint32_t k = S.walls.start[i]; k < S.walls.start[i + 1]; ++k) {
            V o = xi - walls[S.walls.items[k]];
            Q sq = length_squared(o);
            if (sq <= Q(0)) continue;
            Q f, fpd;
            kern(std::sqrt(sq), f, fpd);
            away = away - o * f;
          }
          Q len = length(away);
          Q rcp = Q(1) / len;
          bool valid = std::isfinite(rcp) && rcp > Q(0);
          V nrm = valid ? away * rcp : V{};
          Q vn = dot(vi, nrm);
          Q tau = rmin(vn, Q(0));
          V U = nrm * tau;
          V ub{};
          for (uint32_t k = S.walls.start[i]; k < S.walls.start[i + 1]; ++k) {
            V o = xi - walls[S.walls.items[k]];
            Q sq = length_squared(o);
            if (sq <= Q(0)) continue;
            Q f, fpd;
            kern(std::sqrt(sq), f, fpd);
            V g = o * f;
            Q ga = dot(g, ai), oa = dot(o, ai);
            V hq = ai * f + o * (fpd * oa);
            sbar -= m * ga;
            pbar -= m * ga / (rest * rest);
            xbi = xbi - hq * (m * S_w);
            Q omega = dot(U, o);
            if (omega < Q(0)) {
              Q denw = sq + eps;
              Q kappa = -S.wall_nu / (denw * rho);
              rbar += m * kappa * omega * ga / rho;
              ub = ub - o * (m * kappa * ga);
              xbi = xbi - (o * (Q(-2) * kappa * omega * ga / denw) + U * (kappa * ga) + hq * (kappa * omega)) * m;
            }
          }
          if (valid && planted_mistake() != 5) {
            Q taub = dot(ub, nrm);
            V nb = ub * tau;
            if (vn < Q(0)) {
              nb = nb + vi * taub;
              vbi = vbi + nrm * taub;
            }
            V awb = (nb - nrm * dot(nrm, nb)) * rcp;
            for (uint32_t k = S.walls.start[i]; k < S.walls.start[i + 1]; ++k) {
              V o = xi - walls[S.walls.items[k]];
              Q sq = length_squared(o);
              if (sq <= Q(0)) continue;
              Q f, fpd;
              kern(std::sqrt(sq), f, fpd);
              xbi = xbi - (awb * f + o * (fpd * dot(o, awb)));
            }
          } else if (valid) {
            // planted mistake 5: the normal's own adjoint still reaches v
            Q taub = dot(ub, nrm);
            if (vn < Q(0)) vbi = vbi + nrm * taub;
          }
        }
        rbar += sbar * (dP[i] / (rho * rho) - Q(2) * P[i] / (rho * rho * rho)) + pbar * dP[i];
        X[i] = xbi;
        Vv[i] = vbi;
        rb[i] = rbar;
        nus[i] = nsum;
      }
    });
    // density sum, then back through the sort
    std::vector<V> xprev(n), vprev(n);
    world.dispatch.run(n, [&](size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) {
        V xi = S.x[i];
        V xbi = X[i];
        for (uint32_t k = S.fluid.start[i]; k < S.fluid.start[i + 1]; ++k) {
          uint32_t j = S.fluid.items[k];
          if (j == i) continue;
          V o = xi - S.x[j];
          Q sq = length_squared(o);
          if (sq <= Q(0)) continue;
          Q f, fpd;
          kern(std::sqrt(sq), f, fpd);
          xbi = xbi + o * (m * f * (rb[i] + rb[j]));
        }
        if (!walls.empty() && S.walls.start.size() == n + 1)
          for (uint32_t k = S.walls.start[i]; k < S.walls.start[i + 1]; ++k) {
            V o = xi - walls[S.walls.items[k]];
            Q sq = length_squared(o);
            if (sq <= Q(0)) continue;
            Q f, fpd;
            kern(std::sqrt(sq), f, fpd);
            xbi = xbi + o * (m * f * rb[i]);
          }
        xprev[S.src[i]] = xbi;
        vprev[S.src[i]] = Vv[i];
      }
    });
    xb.swap(xprev);
    vb.swap(vprev);
    Q nu_bar = Q(0);
    for (size_t i = 0; i < n; ++i) nu_bar += nus[i];
    return nu_bar * S.dnu;
  }
};

}  // namespace oracle
