// ORACLE — test infrastructure only. Never linked into the product library.
//
// Grid-level CPU restatement of quasar's 2D constrained-transport ideal-MHD
// step: staggered face->cell collocation, characteristic MP5/MP7 (and MUSCL /
// piecewise-constant) face reconstruction, transverse-Gauss HLLD face fluxes,
// the upwind high-order CT corner EMF with its face-B curl rate, the SSP-RK3
// stage combine, ghost fills, the additive CFL rate, the positivity fraction,
// the div(B) diagnostics and the analytic initial conditions.  Each function
// cites the reference kernel it restates.
//
// Parity status: the per-state numerics it calls are pinned bit-for-bit against
// the reference headers (oracle/_ref).  The composite kernel bodies restated here
// exist only in the reference's .hip device code, which cannot execute on a
// CPU; they are pinned by the reference's own invariants (constant-state
// exactness, conservation, div(B) preservation, 1-D shortcuts, polynomial
// exactness of the quadrature rows) in tests/test_oracle_invariants.py.

#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "oracle_numerics.hpp"

namespace oracle {

struct Grid {
  int nx{1}, ny{1};
  Real lx{1}, ly{1}, ox{0}, oy{0};
  int ng{1};
  Real cell_dx{0}, cell_dy{0};

  Real dx() const { return cell_dx > 0 ? cell_dx : lx / static_cast<Real>(nx); }
  Real dy() const { return cell_dy > 0 ? cell_dy : ly / static_cast<Real>(ny); }
  int pitch() const { return nx + 2 * ng; }
  int height() const { return ny + 2 * ng; }
  std::size_t size() const {
    return static_cast<std::size_t>(pitch()) * static_cast<std::size_t>(height());
  }
  std::size_t index(int i, int j) const {
    return static_cast<std::size_t>(i + ng) +
           static_cast<std::size_t>(pitch()) * static_cast<std::size_t>(j + ng);
  }
  // include/quasar/core/grid.hpp:167-173 (fused multiply-add centre).
  Real xc(int i) const { return std::fma(static_cast<Real>(i) + Real{0.5}, dx(), ox); }
  Real yc(int j) const { return std::fma(static_cast<Real>(j) + Real{0.5}, dy(), oy); }
};

// Eight conserved planes in the reference's component order
// (include/quasar/physics/mhd/mhd_field.hpp:16-40).
struct Field {
  Real* c[8];  // rho, mx, my, mz, energy, bx_face, by_face, bz_cell
};

// Sixteen interface planes: L (0..7) then R (8..15)
// (include/quasar/numerics/interface_states.hpp:30-88).
struct Iface {
  Real* c[16];
};

// Side modes: 0 periodic, 1 outflow, 2 wall, 3 axis, 4 internal
// (include/quasar/physics/mhd/kernels.hpp:43-51).
struct Flags {
  int side[4];
};

// Instrumentation counters used for label generation and fallback statistics.
struct Counters {
  long long mp_clips{0};
  long long muscl_fallbacks{0};
  long long pc_fallbacks{0};
  long long sentinels{0};
  long long point_fallbacks{0};
  long long hll_fallbacks{0};
};

// Collocation (include/quasar/physics/mhd/mhd_staggering.hpp:42-283).
Real cell_bx(const Grid& g, const Real* bx_face, int i, int j, int collocation_order);
Real cell_by(const Grid& g, const Real* by_face, int i, int j, int collocation_order);
State load_cell_state(const Grid& g, const Field& u, int i, int j, int collocation_order);

// Transverse recovery rows (include/quasar/numerics/finite_volume_quadrature.hpp:12-68).
Real transverse_point_value(int order, int node, const Real* samples);
Real transverse_gauss_average(int order, const Real* samples);

// Kernels.
void reconstruct_one_face(const Grid& g, int dir, int i, int j, int order, int side_lo, int side_hi, Real gamma,
                          const Field& u, State& ul, State& ur, Counters* counters, unsigned* bits,
                          Real label_tol);
void reconstruct(const Grid& g, int dir, int order, const Flags& flags, Real gamma,
                 const Field& u, Iface& out, Counters* counters);
void face_flux(const Grid& g, int dir, int order, int hll_only, const Flags& flags,
               Real gamma, const Iface& in, Field& flux, Counters* counters);
void flux_difference(const Grid& g, int dir, const Field& flux, Field& dudt);
void ct_emf(const Grid& g, int order, int hll_only, const Flags& flags, Real gamma,
            const Field& u, const Iface& ifx, const Iface& ify, Real* cell_ez,
            Real* xface_ez, Real* yface_ez, int* xface_no_jump, int* yface_no_jump,
            Real* ez_edge);
void emf_curl_rate(const Grid& g, const Real* ez_edge, Field& dudt);
void rk_stage(const Grid& g, Real a, Real b, Real c, const Field& un, const Field& us,
              const Field& dudt, Field& out);
void fill_ghosts(const Grid& g, const Flags& fluid, const Flags& field, Field& u);
Real cfl_max_rate(const Grid& g, int order, const Flags& flags, Real gamma,
                  const Iface& ifx, const Iface& ify);
Real admissible_fraction(const Grid& g, const Field& base, const Field& cand,
                         Real rho_floor, Real q_floor, int collocation_order);
Real divb_linf(const Grid& g, const Field& u);
Real divb_relative_linf(const Grid& g, const Field& u, bool solver_owned);
void conserved_totals(const Grid& g, const Field& u, Real* mass, Real* energy);

// Analytic initial conditions (src/backend/hip/mhd/mhd_initial_conditions.hip).
struct IcSpec {
  int kind{0};  // 0 brio_wu, 1 alfven_wave, 2 orszag_tang, 3 blast, 4 rotor, 5 confined_blob
  Real gamma{Real{5} / Real{3}};
  Real magnetic_scale{1};
  Real interface_x{0.5};
  Real left[8]{}, right[8]{};  // rho, p, vx, vy, vz, bx, by, bz
  Real rho{1}, pressure{0.1}, b0{1}, total_b0{1}, amplitude{1e-6};
  long long wavenumber{1};
  Real magnetic_velocity_scale{1};
  Real b_uniform[3]{};
  Real center[2]{};
  Real r_in{0.1}, r0{0.1}, r1{0.115}, rho_in{10}, rho_out{1}, p_in{1}, p_out{0.1},
      p_core{10}, p_ambient{0.1}, rho_ambient{1}, u0{2};
  // Orszag-Tang background density/pressure; zero selects the reference
  // generator's gamma^2 / gamma.
  Real ot_rho{0}, ot_p{0};
  // Seeded white-noise perturbation on rho, vx, vy (0 = none).
  Real noise_amplitude{0};
  std::uint64_t noise_seed{0};
  Real blob_half{0.25};  // confined_blob square half-width
};
int seed_initial_state(const Grid& g, const IcSpec& s, Field& out);

// Counter-based generator shared with the product for seeded perturbations:
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw 2011).
void philox4x32_10(std::uint32_t ctr[4], const std::uint32_t key[2]);
Real philox_uniform_pm1(std::uint64_t seed, std::uint32_t component, std::uint64_t cell);

}  // namespace oracle
