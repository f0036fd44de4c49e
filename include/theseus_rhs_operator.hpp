/*
 * C++ face of the C-ABI (theseus_rhs.h) in the shape of the reference's
 * operator interface (RHSOperator.hpp:19-188, EulerOperator.hpp:12-49,
 * NSOperator.hpp): an owning, move-only handle whose Mult(u, dudt) evaluates
 * the DGSEM right-hand side on device-resident L-vectors (byNODES layout,
 * U[q * ndof_total + dof]), EstimateStability returns the reference's
 * StabilityEstimate triple, and Step advances one SSP-RK3 step in place.
 *
 * Errors become exceptions, as the reference raises them from Finalize
 * (RHSOperator_impl.hpp:291-306): std::invalid_argument for a rejected
 * case descriptor or argument, std::runtime_error for a device failure.
 * Like the reference's operator, Mult is const but not reentrant: calls on
 * one operator must not overlap.
 *
 * Header-only; link against libtheseus_rhs.so.  MFEM is not required; the
 * adapter a maintainer adds inside the reference wraps this class in an
 * mfem::TimeDependentOperator.
 */
#ifndef THESEUS_RHS_OPERATOR_HPP
#define THESEUS_RHS_OPERATOR_HPP

#include <algorithm>
#include <stdexcept>
#include <string>

#include "theseus_rhs.h"

namespace theseus {

/* StabilityEstimate.hpp:40-55 */
struct StabilityEstimate {
  double advective_rate = 0.0;
  double surface_rate = 0.0;
  double diffusive_rate = 0.0;
  double TotalRate() const { return std::max(advective_rate, surface_rate) + diffusive_rate; }
};

inline void check_status(int rc) {
  if (rc == THESEUS_OK) return;
  const std::string msg = theseus_last_error();
  if (rc == THESEUS_ERR_INVALID_ARGUMENT || rc == THESEUS_ERR_UNSUPPORTED) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

/* theseus_orient_elements (see theseus_rhs.h): aligns a quad/hex mesh given
 * each element's corner vertex ids and coordinates in its own lexicographic
 * corner order; throws std::invalid_argument when it cannot. */
inline void OrientElements(int dim, int num_elements, const long long *elem_vertices, const double *corner_coords,
                           unsigned char *elem_orient, int *face_neighbor, double *elem_jac, double *elem_adj) {
  check_status(theseus_orient_elements(dim, num_elements, elem_vertices, corner_coords, elem_orient, face_neighbor,
                                       elem_jac, elem_adj));
}

/* Storage corner (lexicographic) of aligned corner c of an element with
 * orientation `code`: the corners are the nodes of order 1. */
inline int StorageCorner(int dim, int code, int c) {
  const int r = theseus_orient_node(dim, 1, code, c);
  check_status(r < 0 ? r : THESEUS_OK);
  return r;
}

class RHSOperator {
 public:
  /* An empty operator, created later by move-assignment (the reference
   * builds its operator before the boundary data exists and completes it in
   * Finalize, src/Simulation.cpp:493,996-997). */
  RHSOperator() = default;
  /* MakeRHSOperator + Finalize (SimFactory.hpp:22-31, RHSOperator_impl.hpp:273-314) */
  explicit RHSOperator(const theseus_case_desc &desc) { check_status(theseus_rhs_create(&desc, &h_)); }
  /* The same with the tabulated LTE gas (LTEGasModel.hpp:339: the LTEEC,
   * Pirozzoli, HLL or LLF flux pairs, Euler or Navier-Stokes;
   * theseus_rhs_create_lte). */
  RHSOperator(const theseus_case_desc &desc, const theseus_lte_table &table) {
    check_status(theseus_rhs_create_lte(&desc, &table, &h_));
  }
  explicit operator bool() const { return h_ != nullptr; }
  ~RHSOperator() {
    if (h_) theseus_rhs_destroy(h_);
  }
  RHSOperator(const RHSOperator &) = delete;
  RHSOperator &operator=(const RHSOperator &) = delete;
  RHSOperator(RHSOperator &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
  RHSOperator &operator=(RHSOperator &&o) noexcept {
    if (this != &o) {
      if (h_) theseus_rhs_destroy(h_);
      h_ = o.h_;
      o.h_ = nullptr;
    }
    return *this;
  }

  /* Scalar dofs per equation and number of equations; Height() = their product
   * (the mfem::TimeDependentOperator height of the reference). */
  long long NumDofs() const {
    long long n = 0;
    int q = 0;
    check_status(theseus_rhs_sizes(h_, &n, &q));
    return n;
  }
  int NumEquations() const {
    long long n = 0;
    int q = 0;
    check_status(theseus_rhs_sizes(h_, &n, &q));
    return q;
  }
  long long Height() const { return NumDofs() * NumEquations(); }

  /* RHSOperator<P>::Mult (RHSOperator_impl.hpp:719-793): dudt = L(u).
   * `stream` is a hipStream_t (nullptr: the default stream); asynchronous. */
  void Mult(const double *u, double *dudt, void *stream = nullptr) const {
    check_status(theseus_rhs_eval(h_, u, dudt, stream));
  }

  /* RHSOperator<P>::EstimateStability (RHSOperator_impl.hpp:10-270); synchronises `stream`. */
  StabilityEstimate EstimateStability(const double *u, void *stream = nullptr) const {
    theseus_stability s{};
    check_status(theseus_rhs_estimate_stability(h_, u, &s, stream));
    StabilityEstimate out;
    out.advective_rate = s.advective_rate;
    out.surface_rate = s.surface_rate;
    out.diffusive_rate = s.diffusive_rate;
    return out;
  }

  /* One RK3SSPSolver step on u in place (src/Simulation.cpp:265-268,1255); returns t + dt. */
  double Step(double *u, double t, double dt, void *stream = nullptr) {
    check_status(theseus_rhs_step_ssprk3(h_, u, t, dt, stream));
    return t + dt;
  }

  /* One variable step with dt chosen on the device (the reference's
   * variable_dt loop, src/Simulation.cpp:1136-1158,1240-1301); `clock` is a
   * device-resident theseus_clock.  Asynchronous. */
  void StepCFL(double *u, theseus_clock *clock, void *stream = nullptr) {
    check_status(theseus_rhs_step_ssprk3_cfl(h_, u, clock, stream));
  }

  /* RHSOperator<P>::ComputeIntegralMeasures (RHSOperator_impl.hpp:530-717); synchronises `stream`. */
  theseus_integral_measures ComputeIntegralMeasures(const double *u, void *stream = nullptr) const {
    theseus_integral_measures m{};
    check_status(theseus_rhs_integral_measures(h_, u, &m, stream));
    return m;
  }

  /* RHSOperator::GetIntegralMeasuresBaseline (RHSOperator.hpp:92). */
  theseus_integral_measures GetIntegralMeasuresBaseline() const {
    theseus_integral_measures m{};
    check_status(theseus_rhs_integral_measures_baseline(h_, &m));
    return m;
  }

  /* FP64 entropy production (sum w J S(u).dudt, sum w J |S(u).dudt|) of a
   * right-hand side; no counterpart in the reference (theseus_rhs_entropy_production). */
  theseus_entropy_production EntropyProduction(const double *u, const double *dudt, void *stream = nullptr) const {
    theseus_entropy_production e{};
    check_status(theseus_rhs_entropy_production(h_, u, dudt, &e, stream));
    return e;
  }

  /* Blending coefficient per element (device array of num_elements doubles). */
  void Alpha(double *alpha_device, void *stream = nullptr) const {
    check_status(theseus_rhs_get_alpha(h_, alpha_device, stream));
  }

  theseus_rhs *handle() const { return h_; }

 private:
  theseus_rhs *h_ = nullptr;
};

}  // namespace theseus

#endif /* THESEUS_RHS_OPERATOR_HPP */
