// Test-infrastructure driver: exposes the reference's own header-only numerics
// (include/quasar/numerics/*.hpp, include/quasar/physics/mhd/mhd_staggering.hpp,
// compiled in place from the reference tree, never copied) through a C ABI so
// that the oracle restatement can be compared with them bit for bit.  Built by
// oracle/Makefile into oracle/_ref/ only when the reference tree is present.

#include <cstring>

#include "quasar/core/grid.hpp"
#include "quasar/numerics/characteristic_projection.hpp"
#include "quasar/numerics/finite_volume_quadrature.hpp"
#include "quasar/numerics/hlld_core.hpp"
#include "quasar/numerics/mhd_eigensystem.hpp"
#include "quasar/numerics/mhd_state.hpp"
#include "quasar/numerics/mp_limiter.hpp"
#include "quasar/numerics/scaled_arithmetic.hpp"
#include "quasar/physics/mhd/mhd_staggering.hpp"

namespace {
using quasar::Real;
using quasar::numerics::MhdFlux;
using quasar::numerics::MhdState;

MhdState st(const double* u) { return MhdState{u[0], u[1], u[2], u[3], u[4], u[5], u[6], u[7]}; }
void out(const MhdFlux& f, double* o) {
  o[0] = f.rho;
  o[1] = f.mx;
  o[2] = f.my;
  o[3] = f.mz;
  o[4] = f.energy;
  o[5] = f.bx;
  o[6] = f.by;
  o[7] = f.bz;
}
quasar::Grid2D grid(int nx, int ny, int ng, double lx, double ly, double ox, double oy) {
  quasar::Grid2D g;
  g.nx = nx;
  g.ny = ny;
  g.nghost = ng;
  g.lx = lx;
  g.ly = ly;
  g.origin_x = ox;
  g.origin_y = oy;
  return g;
}
}  // namespace

extern "C" {

double ref_product_sum(const double* a, const double* b, int n) {
  return quasar::numerics::scaled_product_sum(a, b, n);
}
double ref_pressure(const double* u, double gamma) {
  return quasar::numerics::pressure(st(u), gamma);
}
void ref_physical_flux_x(const double* u, double gamma, double* f) {
  out(quasar::numerics::hlld::physical_flux_x(st(u), gamma), f);
}
void ref_hll_flux_x(const double* l, const double* r, double sl, double sr, double gamma, double* f) {
  out(quasar::numerics::hlld::hll_flux_x(st(l), st(r), sl, sr, gamma), f);
}
void ref_lf_flux_x(const double* l, const double* r, double gamma, double* f) {
  out(quasar::numerics::hlld::lax_friedrichs_flux_x(st(l), st(r), gamma), f);
}
void ref_hlld_flux_x(const double* l, const double* r, double gamma, double* f) {
  out(quasar::numerics::hlld::hlld_flux_x(st(l), st(r), gamma), f);
}
double ref_fast_speed(const double* u, int dir, double gamma) {
  return quasar::numerics::fast_magnetosonic_speed(st(u), dir, gamma);
}
double ref_mp_limit(double vl, const double* v) {
  return quasar::numerics::mp_limit(vl, v[0], v[1], v[2], v[3], v[4]);
}
double ref_mp5_interp(const double* v) { return quasar::numerics::mp5_interp(v[0], v[1], v[2], v[3], v[4]); }
double ref_mp7_interp(const double* v) {
  return quasar::numerics::mp7_interp(v[0], v[1], v[2], v[3], v[4], v[5], v[6]);
}
double ref_mp5_reconstruct(const double* v) {
  return quasar::numerics::mp5_reconstruct(v[0], v[1], v[2], v[3], v[4]);
}
double ref_mp7_reconstruct(const double* v) {
  return quasar::numerics::mp7_reconstruct(v[0], v[1], v[2], v[3], v[4], v[5], v[6]);
}
void ref_eigen_build(const double* u, double gamma, double* L, double* R, double* speeds) {
  quasar::numerics::MhdEigensystem e;
  e.build(st(u), 0, gamma);
  for (int k = 0; k < 7; ++k) {
    const Real* row = e.left_row(k);
    for (int v = 0; v < 7; ++v) L[k * 7 + v] = row[v];
    const Real* col = e.right_col(k);
    for (int v = 0; v < 7; ++v) R[v * 7 + k] = col[v];
    speeds[k] = e.wave_speed(k);
  }
}
void ref_to_char(const double* d8, const double* u_ref, double gamma, double* w7) {
  quasar::numerics::MhdEigensystem e;
  e.build(st(u_ref), 0, gamma);
  const auto w = quasar::numerics::CharacteristicProjector::to_char(st(d8), e);
  for (int k = 0; k < 7; ++k) w7[k] = w.w[k];
}
void ref_from_char(const double* w7, const double* u_ref, double gamma, double* d8) {
  quasar::numerics::MhdEigensystem e;
  e.build(st(u_ref), 0, gamma);
  quasar::numerics::CharVec7 w;
  for (int k = 0; k < 7; ++k) w.w[k] = w7[k];
  const MhdState d = quasar::numerics::CharacteristicProjector::from_char(w, e);
  d8[0] = d.rho;
  d8[1] = d.mx;
  d8[2] = d.my;
  d8[3] = d.mz;
  d8[4] = d.energy;
  d8[5] = d.bx;
  d8[6] = d.by;
  d8[7] = d.bz;
}
double ref_cell_bx(int nx, int ny, int ng, double lx, double ly, double ox, double oy,
                   const double* face, int i, int j, int co) {
  return quasar::mhd::cell_bx(grid(nx, ny, ng, lx, ly, ox, oy), face, i, j, co);
}
double ref_cell_by(int nx, int ny, int ng, double lx, double ly, double ox, double oy,
                   const double* face, int i, int j, int co) {
  return quasar::mhd::cell_by(grid(nx, ny, ng, lx, ly, ox, oy), face, i, j, co);
}
double ref_point_value7(int node, const double* samples) {
  return quasar::numerics::weighted_quadrature_value(
      quasar::numerics::kMp7TransversePointWeights[node], [&](int k) { return samples[k]; });
}
double ref_point_value5(int node, const double* samples) {
  return quasar::numerics::weighted_quadrature_value(
      quasar::numerics::kMp5TransversePointWeights[node], [&](int k) { return samples[k]; });
}
double ref_gauss_average(int order, const double* samples) {
  if (order == 5)
    return quasar::numerics::weighted_quadrature_value(quasar::numerics::kMp5TransverseGaussWeights,
                                                       [&](int q) { return samples[q]; });
  return quasar::numerics::weighted_quadrature_value(quasar::numerics::kMp7TransverseGaussWeights,
                                                     [&](int q) { return samples[q]; });
}
double ref_x_at_cell_center(int nx, double lx, double ox, int i) {
  return grid(nx, 1, 1, lx, 1.0, ox, 0.0).x_at_cell_center(i);
}

}  // extern "C"
