// Device data views and host launchers for the MHD step kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "qmhd_numerics.hpp"

namespace qmhd {

struct GridD {
  int nx, ny, ng, pitch;
  double dx, dy, ox, oy, lx, ly;
};

struct Flags4 {
  int side[4];
};

struct StateView {
  const double* c[8];
};
struct StateOut {
  double* c[8];
};
template <class T>
struct IfaceView {
  T* c[16];
};
template <class T>
struct FluxView {
  T* c[8];
};

// Device-resident per-step status, reduced with order-independent atomics on
// the bit patterns of non-negative doubles.
struct StepStatus {
  unsigned long long theta[4];     // per-stage admissible fraction (min); [3] = stage 2, second collocation
  unsigned long long live_theta0;  // live state, high-order collocation (min)
  unsigned long long live_theta1;  // live state, two-face collocation (min)
  unsigned long long divb_residual;  // max unexplained |div B| (relative check)
  unsigned long long divb_scale;     // max directional scale
  unsigned long long max_rate;       // CFL additive rate (max)
  double dt;                         // device-side dt of the auto-dt loop
  double time;                       // accumulated simulated time
  double pad[5];
};

struct IcSpecD {
  int kind;
  double gamma, magnetic_scale, interface_x;
  double left[8], right[8];
  double rho, pressure, b0, total_b0, amplitude;
  long long wavenumber;
  double magnetic_velocity_scale;
  double b_uniform[3];
  double center[2];
  double r_in, r0, r1, rho_in, rho_out, p_in, p_out, p_core, p_ambient, rho_ambient, u0;
  double ot_rho, ot_p, noise_amplitude;
  unsigned long long noise_seed;
};

hipError_t upload_quad_rows(const QuadRows& rows);
void build_quad_rows(QuadRows& rows);

// Launchers (all asynchronous on `stream`).
void launch_prep(const GridD& g, const StateView& u, double* bxc, double* byc, const StateView& base,
                 const double* base_bxc, const double* base_byc, int theta_slot, int theta_co, int theta_both,
                 int live_checks, int solver_owned, StepStatus* status, hipStream_t stream);
void launch_collocate(const GridD& g, int axis, int co, const double* face, double* cell, hipStream_t stream);
// The face-level launchers are instantiated for T = double (the FP64 form)
// and T = float (the FP32 form: interface states, fluxes and their arithmetic
// in FP32; state, collocation, cell/corner E_z, rates and the update in FP64).
template <class T>
void launch_reconstruct(const GridD& g, int dir, int order, double gamma, const Flags4& flags, const StateView& u,
                        const double* bxc, const double* byc, const IfaceView<T>& out, int* eq,
                        unsigned long long* fb_count, unsigned long long* fb_list, hipStream_t stream);
template <class T>
void launch_flux(const GridD& g, int dir, int order, int hll_only, const Flags4& flags, double gamma,
                 const IfaceView<T>& in, const FluxView<T>& flux, double* rate, hipStream_t stream);
void launch_cfl_reduce(const GridD& g, const double* rate_x, const double* rate_y, unsigned long long* max_rate,
                       hipStream_t stream);
template <class T>
void launch_cell_emf(const GridD& g, int order, const Flags4& flags, double gamma, const StateView& u,
                     const double* bxc, const double* byc, double* cell_ez, hipStream_t stream);
template <class T>
void launch_corner(const GridD& g, int order, const Flags4& flags, const FluxView<T>& fx, const FluxView<T>& fy,
                   const double* cell_ez, const int* eqx, const int* eqy, double* ez, hipStream_t stream);
template <class T>
void launch_update(const GridD& g, double a, double b, double cw, const double* dt_dev, const StateView& un,
                   const StateView& us, const FluxView<T>& fx, const FluxView<T>& fy, const double* ez,
                   const StateOut& out, hipStream_t stream);
void launch_rk_combine(const GridD& g, double a, double b, double c, const StateView& un, const StateView& us,
                       const StateView& dudt, const StateOut& out, hipStream_t stream);
void launch_ghosts(const GridD& g, const Flags4& fluid, const Flags4& field, const StateOut& u, hipStream_t stream);
void launch_dt(const unsigned long long* max_rate_bits, double effective_cfl, double dx, double dy, double* dt_out,
               double* t_accum, hipStream_t stream);
void launch_seed(const GridD& g, const IcSpecD& s, const StateOut& out, double* raw_mag, int* status,
                 hipStream_t stream);
void launch_divb_linf(const GridD& g, const StateView& u, unsigned long long* out, hipStream_t stream);
void launch_theta(const GridD& g, const StateView& base, const StateView& cand, int co, unsigned long long* out,
                  hipStream_t stream);
void launch_cell_sums(const GridD& g, const StateView& u, double* partials, int blocks, hipStream_t stream);
void launch_eq_flags(const GridD& g, int dir, const IfaceView<double>& in, int* eq, hipStream_t stream);
void launch_flux_difference(const GridD& g, int dir, const FluxView<double>& f, const StateOut& d,
                            hipStream_t stream);
void launch_curl(const GridD& g, const double* ez, double* dbx, double* dby, hipStream_t stream);
void launch_status_reset(StepStatus* s, hipStream_t stream);

}  // namespace qmhd
