// ORACLE — test infrastructure only.  CPU restatement of the learned form's
// pieces that have no reference counterpart: the indicator's cell features,
// the training labels taken from the plain method (the reference's MP limiter
// activity, include/quasar/numerics/mp_limiter.hpp:66-148, its reconstruction
// fallback chain, src/backend/hip/mhd/mhd_reconstruct.hip:611-663, and the gap
// between its HLLD and HLL fluxes, include/quasar/numerics/hlld_core.hpp:
// 157-192, 213-457), and the 100-32-32-1 network with its loss gradient.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_capi.h"
#include "oracle_grid.hpp"

namespace oracle {
namespace {

Grid grid_of(const OrcGrid* g) {
  Grid out;
  out.nx = g->nx;
  out.ny = g->ny;
  out.ng = g->ng;
  out.lx = g->lx;
  out.ly = g->ly;
  out.ox = g->ox;
  out.oy = g->oy;
  out.cell_dx = g->cell_dx;
  out.cell_dy = g->cell_dy;
  return out;
}

Real comp8(const State& s, int c) {
  const Real v[8] = {s.rho, s.mx, s.my, s.mz, s.energy, s.bx, s.by, s.bz};
  return v[c];
}

Field field_of(double* const* p) {
  Field f;
  for (int c = 0; c < 8; ++c) f.c[c] = p[c];
  return f;
}

// HLL with the Davis speeds HLLD computes after its normal-field averaging
// (hlld_core.hpp:241-290): the flux HLLD falls back to.
Flux hll_davis(const State& Lin, const State& Rin, Real gamma) {
  const Real bn = Real{0.5} * Lin.bx + Real{0.5} * Rin.bx;
  State L = Lin, R = Rin;
  const Real half_bn2 = half_squared_norm3(bn, Real{0}, Real{0});
  L.energy = (L.energy - half_squared_norm3(L.bx, Real{0}, Real{0})) + half_bn2;
  R.energy = (R.energy - half_squared_norm3(R.bx, Real{0}, Real{0})) + half_bn2;
  L.bx = bn;
  R.bx = bn;
  const Real vxL = L.mx / L.rho, vxR = R.mx / R.rho;
  const Real cfL = fast_speed(L, 0, gamma), cfR = fast_speed(R, 0, gamma);
  const Real cfmax = cfL > cfR ? cfL : cfR;
  const Real sl = (vxL < vxR ? vxL : vxR) - cfmax;
  const Real sr = (vxL > vxR ? vxL : vxR) + cfmax;
  return hll_flux_x(L, R, sl, sr, gamma);
}

// Cell quantities rho, p, |B|, |v| (collocated B) in FP64, in this order of
// operations, rounded to FP32.
void quantities(const Grid& g, const Field& u, const std::vector<Real>& bxc, const std::vector<Real>& byc,
                Real gamma, int i, int j, float q[4]) {
  const std::size_t k = g.index(i, j);
  const Real rho = u.c[0][k], mx = u.c[1][k], my = u.c[2][k], mz = u.c[3][k], en = u.c[4][k];
  const Real bx = bxc[k], by = byc[k], bz = u.c[7][k];
  const Real m2 = (mx * mx + my * my) + mz * mz;
  const Real b2 = (bx * bx + by * by) + bz * bz;
  q[0] = static_cast<float>(rho);
  q[1] = static_cast<float>((gamma - 1.0) * ((en - 0.5 * m2 / rho) - 0.5 * b2));
  q[2] = static_cast<float>(std::sqrt(b2));
  q[3] = static_cast<float>(std::sqrt(m2) / rho);
}

float round_bf16(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return x;
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

constexpr int kIn = 100, kH = 32;
constexpr int kW1 = 0, kB1 = kIn * kH, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH, kB3 = kW3 + kH;
constexpr int kParams = kB3 + 1;

struct Pass {
  Real x[kIn], z1[kH], a1[kH], z2[kH], a2[kH], z3;
};

// Forward pass of one sample; `q(v)` applies the operand rounding.
template <class P, class Q>
void forward(const P& p, const float* xrow, Q q, Pass& s) {
  for (int f = 0; f < kIn; ++f) s.x[f] = q(xrow[f]);
  for (int h = 0; h < kH; ++h) {
    Real acc = 0;
    for (int f = 0; f < kIn; ++f) acc += s.x[f] * q(p(kW1 + f * kH + h));
    s.z1[h] = acc + p(kB1 + h);
    s.a1[h] = q(s.z1[h] > 0 ? s.z1[h] : Real{0});
  }
  for (int h2 = 0; h2 < kH; ++h2) {
    Real acc = 0;
    for (int h1 = 0; h1 < kH; ++h1) acc += s.a1[h1] * q(p(kW2 + h1 * kH + h2));
    s.z2[h2] = acc + p(kB2 + h2);
    s.a2[h2] = s.z2[h2] > 0 ? s.z2[h2] : Real{0};
  }
  Real z = 0;
  for (int h = 0; h < kH; ++h) z += p(kW3 + h) * s.a2[h];
  s.z3 = z + p(kB3);
}

}  // namespace
}  // namespace oracle

using namespace oracle;

extern "C" {

void orc_hll_davis_flux_x(const double* l, const double* r, double gamma, double* f) {
  const State L{l[0], l[1], l[2], l[3], l[4], l[5], l[6], l[7]};
  const State R{r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7]};
  const Flux out = hll_davis(L, R, gamma);
  for (int c = 0; c < 8; ++c) f[c] = comp8(out, c);
}

void orc_cell_features(const OrcGrid* og, double gamma, double* const* u8, float* out) {
  const Grid g = grid_of(og);
  const Field u = field_of(u8);
  const std::size_t n = g.size();
  std::vector<Real> bxc(n), byc(n);
  for (int j = -g.ng; j < g.ny + g.ng; ++j)
    for (int i = -g.ng; i < g.nx + g.ng; ++i) {
      bxc[g.index(i, j)] = cell_bx(g, u.c[5], i, j, 0);
      byc[g.index(i, j)] = cell_by(g, u.c[6], i, j, 0);
    }
#pragma omp parallel for schedule(static)
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      float q[5][5][4];
      for (int dy = 0; dy < 5; ++dy)
        for (int dx = 0; dx < 5; ++dx) quantities(g, u, bxc, byc, gamma, i + dx - 2, j + dy - 2, q[dy][dx]);
      float* o = out + (static_cast<std::size_t>(j) * g.nx + i) * kIn;
      for (int m = 0; m < 4; ++m) {
        // FP32 normalisation: (Q_k - Q_c) / (|Q_c| + max |Q_k - Q_c| + 1e-30).
        const float c = q[2][2][m];
        float dev = 0.f;
        for (int dy = 0; dy < 5; ++dy)
          for (int dx = 0; dx < 5; ++dx) dev = std::fmax(dev, std::fabs(q[dy][dx][m] - c));
        const float scale = 1.f / ((std::fabs(c) + dev) + 1e-30f);
        for (int dy = 0; dy < 5; ++dy)
          for (int dx = 0; dx < 5; ++dx) o[25 * m + 5 * dy + dx] = (q[dy][dx][m] - c) * scale;
      }
    }
}

void orc_cell_labels(const OrcGrid* og, int order, const int* flags, double gamma, double* const* u8,
                     double label_tol, uint8_t* cell_bits) {
  const Grid g = grid_of(og);
  const Field u = field_of(u8);
  const std::size_t n = g.size();
  std::vector<uint8_t> fb[2] = {std::vector<uint8_t>(n, 0), std::vector<uint8_t>(n, 0)};
  for (int dir = 0; dir < 2; ++dir) {
    const int slo = dir == 0 ? flags[0] : flags[2];
    const int shi = dir == 0 ? flags[1] : flags[3];
    const int i1 = dir == 0 ? g.nx : g.nx - 1, j1 = dir == 0 ? g.ny - 1 : g.ny;
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j <= j1; ++j)
      for (int i = 0; i <= i1; ++i) {
        unsigned bits = 0;
        State ul, ur;
        reconstruct_one_face(g, dir, i, j, order, slo, shi, gamma, u, ul, ur, nullptr, &bits, label_tol);
        const State lr = rotate(ul, dir), rr = rotate(ur, dir);
        const Flux fd = hlld_flux_x(lr, rr, gamma);
        const Flux fh = hll_davis(lr, rr, gamma);
        Real gap = 0, mag = 0;
        for (int c = 0; c < 8; ++c) {
          gap = std::fmax(gap, std::fabs(comp8(fd, c) - comp8(fh, c)));
          mag = std::fmax(mag, std::fabs(comp8(fd, c)));
        }
        if (!(gap <= label_tol * (mag + 1e-300))) bits |= 8u;
        fb[dir][g.index(i, j)] = static_cast<uint8_t>(bits);
      }
  }
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i)
      cell_bits[static_cast<std::size_t>(j) * g.nx + i] =
          fb[0][g.index(i, j)] | fb[0][g.index(i + 1, j)] | fb[1][g.index(i, j)] | fb[1][g.index(i, j + 1)];
}

void orc_mlp_forward(const float* params, const float* x, long long n, int bf16, double* logits) {
  auto p = [&](int k) { return static_cast<Real>(params[k]); };
  auto q = [&](Real v) { return bf16 ? static_cast<Real>(round_bf16(static_cast<float>(v))) : v; };
  for (long long s = 0; s < n; ++s) {
    Pass ps;
    forward(p, x + s * kIn, q, ps);
    logits[s] = ps.z3;
  }
}

// Loss sum_s w_s BCE(z_s, y_s) and its gradient (grad[kParams]).  params64,
// when given, overrides params (exact network for finite differences).
double orc_mlp_loss_grad(const float* params, const double* params64, const float* x, const float* y, const float* w,
                         long long n, int bf16, double* grad) {
  auto p = [&](int k) { return params64 ? params64[k] : static_cast<Real>(params[k]); };
  auto q = [&](Real v) { return bf16 ? static_cast<Real>(round_bf16(static_cast<float>(v))) : v; };
  std::vector<Real> gsum(kParams, 0.0);
  Real loss = 0;
  for (long long s = 0; s < n; ++s) {
    Pass ps;
    forward(p, x + s * kIn, q, ps);
    const Real z = ps.z3, yt = y[s], wt = w[s];
    loss += wt * (std::fmax(z, 0.0) - z * yt + std::log1p(std::exp(-std::fabs(z))));
    const Real dl = wt * (1.0 / (1.0 + std::exp(-z)) - yt);
    gsum[kB3] += dl;
    Real d2[kH], d2q[kH], d1[kH];
    for (int h = 0; h < kH; ++h) {
      gsum[kW3 + h] += ps.a2[h] * dl;
      d2[h] = ps.z2[h] > 0 ? p(kW3 + h) * dl : 0.0;
      gsum[kB2 + h] += d2[h];
      d2q[h] = q(d2[h]);
    }
    for (int h1 = 0; h1 < kH; ++h1) {
      Real acc = 0;
      for (int h2 = 0; h2 < kH; ++h2) acc += q(p(kW2 + h1 * kH + h2)) * d2q[h2];
      d1[h1] = ps.z1[h1] > 0 ? acc : 0.0;
      gsum[kB1 + h1] += d1[h1];
      for (int h2 = 0; h2 < kH; ++h2) gsum[kW2 + h1 * kH + h2] += ps.a1[h1] * d2q[h2];
    }
    for (int h = 0; h < kH; ++h) {
      const Real d1q = q(d1[h]);
      for (int f = 0; f < kIn; ++f) gsum[kW1 + f * kH + h] += ps.x[f] * d1q;
    }
  }
  for (int k = 0; k < kParams; ++k) grad[k] = gsum[k];
  return loss;
}

}  // extern "C"
