/*
 * theseus_rhs.h -- C-ABI of the MI355X DGSEM right-hand-side engine.
 *
 * This is the drop-in boundary for the reference's RHS evaluation
 * (chess-uiuc/Theseus):
 *
 *   reference                                   | here
 *   --------------------------------------------+------------------------------------
 *   MakeRHSOperator(runTime, vfes, ...)         | theseus_rhs_create(desc, &handle)
 *     include/SimFactory.hpp:22-31,             |
 *     include/SimFactory_impl.hpp:63-488        |
 *   RHSOperator::Finalize(t)                    | (done inside theseus_rhs_create)
 *     include/RHSOperator_impl.hpp:273-314      |
 *   RHSOperator<P>::Mult(u, dudt)               | theseus_rhs_eval(h, u, dudt, stream)
 *     include/RHSOperator_impl.hpp:719-793      |
 *   RHSOperator<P>::EstimateStability(u)        | theseus_rhs_estimate_stability(h, u, r)
 *     include/RHSOperator_impl.hpp:10-270       |
 *   mfem::RK3SSPSolver::Step(x, t, dt) driving  | theseus_rhs_step_ssprk3(h, u, t, dt, s)
 *     Mult (src/Simulation.cpp:265-268,1255)    |
 *   RHSOperatorBase::~RHSOperatorBase           | theseus_rhs_destroy(h)
 *
 * Data layout of every state vector (u, dudt): the reference's L-vector,
 * "byNODES" (include/Simulation.hpp:68-69):
 *     u[q * (ne * ndof) + e * ndof + node]
 * with q the equation (rho, rho*u_d..., rho*E), e the element and node the
 * lexicographic GLL node index i + Np*(j + Np*k) within the element.
 *
 * All device pointers are plain HIP device pointers (hipMalloc'd or any other
 * device allocation); no torch or MFEM types cross this boundary.  Every call
 * returns 0 on success or a negative THESEUS_ERR_* code; theseus_last_error()
 * returns a thread-local message for the last failure.  A handle is not
 * reentrant: calls on one handle must not overlap (the reference's Mult is
 * likewise const but not thread-safe, include/RHSOperator.hpp:129-130).
 */
#ifndef THESEUS_RHS_H
#define THESEUS_RHS_H

#if defined(__GNUC__) || defined(__clang__)
#define THESEUS_API __attribute__((visibility("default")))
#else
#define THESEUS_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define THESEUS_OK 0
#define THESEUS_ERR_INVALID_ARGUMENT (-1)
#define THESEUS_ERR_UNSUPPORTED (-2)
#define THESEUS_ERR_DEVICE (-3)
#define THESEUS_ERR_NONFINITE (-4)
#define THESEUS_ERR_CAPACITY (-5) /* serving pool: not enough free elements or case slots */

/* ---- flow model (SimFactory_impl.hpp:119-123 "flow_model") ------------- */
#define THESEUS_FLOW_EULER 0
#define THESEUS_FLOW_NAVIER_STOKES 1

/* ---- two-point volume fluxes (A3) -------------------------------------- */
#define THESEUS_VOLUME_CENTRAL 0       /* RoeFlux.hpp:18-45, HLLFlux.hpp:19-51, LaxFriedrichsFlux.hpp:20-52 */
#define THESEUS_VOLUME_CHANDRASHEKAR 1 /* ChandrashekarFlux.hpp:20-89 */
#define THESEUS_VOLUME_RANOCHA 2       /* RanochaFlux.hpp:20-79 */
#define THESEUS_VOLUME_PIROZZOLI 3     /* PirozzoliFlux.hpp:19-71 */
#define THESEUS_VOLUME_LTEEC 4         /* LTEECFlux.hpp:11-99, tabulated LTE gas (theseus_rhs_create_lte) */

/* ---- surface (face) fluxes (A5) ---------------------------------------- */
#define THESEUS_SURFACE_CHANDRASHEKAR 0 /* EC + Scalar_dissipation, ChandrashekarFlux.hpp:91-155 */
#define THESEUS_SURFACE_RANOCHA 1       /* EC + Roe_dissipation, RanochaFlux.hpp:81-134 */
#define THESEUS_SURFACE_PIROZZOLI 2     /* KEP + Roe_dissipation, PirozzoliFlux.hpp:73-120 */
#define THESEUS_SURFACE_ROE 3           /* RoeFlux.hpp:50-157 */
#define THESEUS_SURFACE_HLL 4           /* HLLFlux.hpp:53-116 */
#define THESEUS_SURFACE_LLF 5           /* LaxFriedrichsFlux.hpp:54-77 */
#define THESEUS_SURFACE_LTEEC 6         /* LTEEC + Lax dissipation, LTEECFlux.hpp:11-99, DissipationOperator.hpp:15-28 */

/* ---- boundary condition types (BCType, bc_cache_utilities.hpp:11-23) ----
 * Inviscid fluxes as ApplyBoundaryConditionInviscid (bc_kernels.hpp:219-287),
 * Navier-Stokes fluxes as ApplyViscousBoundaryCondition (:289-375) and the
 * BR1 gradient's boundary jumps as ComputeBdrFaceGradFlux (:32-135).  Each
 * boundary index b owns THESEUS_BC_DATA_STRIDE doubles of bc_data
 * (bc_data[b * THESEUS_BC_DATA_STRIDE + k]), the reference's bc_vector_data
 * payload of that boundary (Simulation.cpp:575-820):
 *   SUPERSONIC_INFLOW, PRESCRIBED_STATE: the exterior conservative state
 *     [rho, rho*u_0..rho*u_{dim-1}, rho*E];
 *   NO_SLIP_ADIABATIC: wall velocity [dim], then the wall heat flux;
 *   NO_SLIP_ISOTHERMAL: wall velocity [dim], then the wall temperature (> 0);
 *   SLIP_WALL, SUPERSONIC_OUTFLOW, SYMMETRY: none.
 * In the Euler operator the no-slip types give a zero flux, as the
 * reference's inviscid switch does (its default branch). */
#define THESEUS_BC_SLIP_WALL 0          /* (0, p* n, 0), p* from the 1-D wall Riemann problem, Flow.hpp:55-170 */
#define THESEUS_BC_SUPERSONIC_INFLOW 1  /* F*(q, q_bc, n) */
#define THESEUS_BC_SUPERSONIC_OUTFLOW 2 /* F*(q, q, n), bc_kernels.hpp:236-238 */
#define THESEUS_BC_PRESCRIBED_STATE 3   /* F*(q, q_bc, n); NS: - n.Fv(q, grad q) */
#define THESEUS_BC_SYMMETRY 4           /* F*(q, q mirrored about the unit normal, n) */
#define THESEUS_BC_AXIS 5               /* axisymmetric builds only: rejected (THESEUS_ERR_UNSUPPORTED) */
#define THESEUS_BC_NO_SLIP_ADIABATIC 6  /* NS wall, bc_kernels.hpp:154-184 */
#define THESEUS_BC_NO_SLIP_ISOTHERMAL 7 /* NS wall, bc_kernels.hpp:186-216 */
#define THESEUS_BC_DATA_STRIDE 8

/* ---- precision modes ---------------------------------------------------- */
#define THESEUS_PRECISION_FP64 0
#define THESEUS_PRECISION_MIXED 1 /* FP32 flux arithmetic, gradients and LDS staging; FP64 state, SSP-RK3
                                     stage combination, blending indicator and stability estimate.  The
                                     momentum fluxes leave out a background pressure (the pressure at the
                                     first node of the first state the handle evaluates, fixed for the
                                     handle's lifetime), formed in FP64 before rounding to FP32. */

/*
 * Case descriptor.  Host pointers; copied into device memory by create, so
 * the caller may free them afterwards.
 *
 * Geometry is affine per element (J and the adjugate of dx/dxi on the
 * reference cube [0,1]^dim), which is what the reference's cache holds for
 * every node of an affine element (dgsem_cache_utilities.hpp:469-514).
 *
 * Connectivity: face_neighbor[e*2*dim + 2*d + s] for local face (d, s),
 * d = reference direction, s = 0 (xi_d = 0 side) or 1 (xi_d = 1 side):
 *   >= 0 : the neighbouring element, whose opposite face (d, 1-s) is
 *          matched node-for-node (same in-face lexicographic order);
 *   <  0 : a boundary face with bc index -(value+1) into bc_type[].
 */
typedef struct theseus_case_desc {
  int dim;           /* 2 or 3 */
  int order;         /* polynomial order p (Np = p+1 GLL nodes per direction) */
  int num_elements;  /* ne */
  int flow_model;    /* THESEUS_FLOW_* */
  int volume_flux;   /* THESEUS_VOLUME_* */
  int surface_flux;  /* THESEUS_SURFACE_* */
  int precision;     /* THESEUS_PRECISION_* */
  int blending;      /* 1: Persson-Peraire indicator + subcell FV blending (SUBCELL_FV_BLENDING) */

  /* gas and transport constants (Physics.hpp:13-43) */
  double gamma;
  double Pr;
  double R_gas;
  double mu;
  double mu_bulk;    /* reference default 2/3 */

  /* blending constants (RHSOperator.hpp:33-35,53,64) */
  double alpha_max;  /* reference default 0.5 */
  double alpha_min;  /* reference default 0.001 */

  const double *elem_jac;      /* [ne] */
  const double *elem_adj;      /* [ne][dim][dim], row = reference direction */
  const int *face_neighbor;    /* [ne][2*dim] */

  int num_bc;
  const int *bc_type;          /* [num_bc] THESEUS_BC_* */
  const double *bc_data;       /* [num_bc][THESEUS_BC_DATA_STRIDE]; may be NULL when no boundary needs data */

  /* Element orientation (may be NULL: every element is stored in the
   * aligned frame).  The kernels need neighbouring elements to share their
   * faces with the same in-face node order (the "aligned" frame, which
   * face_neighbor, elem_jac and elem_adj describe).  A mesh whose elements
   * keep their own local frames (an MFEM mesh whose faces carry
   * orientations, dgsem_cache_utilities.hpp:533-541) is described in the
   * aligned frame plus, per element, the symmetry of the reference
   * square/cube that maps its storage frame onto the aligned one:
   * elem_orient[e] = code (theseus_orient_elements computes them), and the
   * state vectors keep the storage order: the node with aligned
   * lexicographic index n of element e is stored at
   * e * ndof + theseus_orient_node(dim, order, elem_orient[e], n).
   * Code c = perm * 2^dim + flip (perm: the index of an axis permutation
   * in lexicographic order; flip: bit dim-1-k set when aligned axis k runs
   * opposite to its storage axis): aligned index a_k = s_{perm[k]}, or
   * p - s_{perm[k]} when flipped, from the storage multi-index s.  A
   * handle with any nonzero code permutes its inputs into the aligned frame
   * and its outputs back in two extra passes over the state (16 bytes per
   * value each way); a handle with none runs as without the array. */
  const unsigned char *elem_orient; /* [ne] or NULL */
} theseus_case_desc;

typedef struct theseus_rhs theseus_rhs; /* opaque */

/* Stability rates (StabilityEstimate.hpp:40-55): total = max(adv, surf) + diff */
typedef struct theseus_stability {
  double advective_rate;
  double surface_rate;
  double diffusive_rate;
} theseus_stability;

/* Replaces MakeRHSOperator + RHSOperator::Finalize (SimFactory_impl.hpp:63-488,
 * RHSOperator_impl.hpp:273-314).  Allocates the device-resident cache. */
THESEUS_API int theseus_rhs_create(const theseus_case_desc *desc, theseus_rhs **out);

/* Six-plane Helmholtz table of a tabulated LTE gas (LTEHelmholtzTable.hpp:13-31,
 * LTETable.hpp:65-72): grids and planes as the reference's HelmholtzTable::Data
 * holds them (host pointers, copied by create). */
typedef struct theseus_lte_table {
  int nx, ny;
  const double *rho_grid; /* [nx] strictly increasing, > 0 */
  const double *T_grid;   /* [ny] strictly increasing, > 0 */
  const double *e_grid;   /* [ny] strictly increasing: the inverse table's energy axis */
  const double *planes;   /* [6][ny][nx]: H, H_rho, H_T, H_rhoT, mu, lambda */
  const double *inverse;  /* [ny][nx]: temperature guess at (rho_i, e_j) */
  /* Viscosity of the Navier-Stokes form (LTETransport.hpp:22-43): 0 = the
   * mu plane, bilinear (the reference's default build); 1 = Sutherland's law
   * mu0 (T0 + Ts) (T/T0)^1.5 / (T + Ts) at the recovered temperature (its
   * SUTHERLAND build; the reference takes mu0 from the runTime key "mu0",
   * default 1.716e-5, and T0 = 273.15, Ts = 110.4, Physics.hpp:40-42).  The
   * conductivity stays the lambda plane either way. */
  int viscosity_model;
  double mu0, T0, Ts;
} theseus_lte_table;

/* MakeRHSOperator for the LTEHelmholtzGas (LTEGasModel.hpp:339) with one of
 * the flux pairs the reference's factory builds for it
 * (SimFactory_impl.hpp:140-245), as (volume_flux, surface_flux):
 * (THESEUS_VOLUME_LTEEC, THESEUS_SURFACE_LTEEC), LTEECFlux.hpp:93-107;
 * (THESEUS_VOLUME_PIROZZOLI, THESEUS_SURFACE_PIROZZOLI);
 * (THESEUS_VOLUME_CENTRAL, THESEUS_SURFACE_HLL) or
 * (THESEUS_VOLUME_CENTRAL, THESEUS_SURFACE_LLF); any other pair is
 * THESEUS_ERR_INVALID_ARGUMENT.  The last three take pressure, sound speed
 * and the Roe dissipation's gamma = cp/cv from the table at each node's
 * recovered temperature.  No blending.
 * precision FP64, or MIXED: FP32 pair and face fluxes (LTEEC: the
 * discrete-gradient means of the interpolant on an FP32 copy of the planes),
 * with the temperature recovery, each node's pressure and sound speed, the
 * stage combination and the stability estimate in FP64.  Euler, or
 * Navier-Stokes (flow_model THESEUS_FLOW_NAVIER_STOKES, periodic meshes):
 * BR1 on the gas's entropy variables ((g - |u|^2/2)/T, u/T, -1/T), viscosity
 * and conductivity bilinear from the table's mu and lambda planes
 * (LTETransport.hpp:32-65, the build without SUTHERLAND), bulk-viscosity
 * ratio mu_bulk from the descriptor; the descriptor's mu and Pr are unused.
 * A state outside the table (or whose temperature recovery does not
 * converge) yields non-finite values instead of the reference's trap. */
THESEUS_API int theseus_rhs_create_lte(const theseus_case_desc *desc, const theseus_lte_table *table,
                                       theseus_rhs **out);

/* Replaces RHSOperatorBase::~RHSOperatorBase. */
THESEUS_API int theseus_rhs_destroy(theseus_rhs *h);

/* Number of scalar dofs per equation (ne * Np^dim) and equations (dim+2). */
THESEUS_API int theseus_rhs_sizes(const theseus_rhs *h, long long *ndof_total, int *num_equations);

/* Replaces RHSOperator<P>::Mult(u, dudt) (RHSOperator_impl.hpp:719-793):
 * dudt = L(u) with u, dudt FP64 L-vectors in device memory.  Enqueued on
 * `stream` (a hipStream_t; NULL = default stream), asynchronous. */
THESEUS_API int theseus_rhs_eval(theseus_rhs *h, const double *u, double *dudt, void *stream);

/* Replaces RHSOperator<P>::EstimateStability(u) (RHSOperator_impl.hpp:10-270)
 * followed by the host max-reduction; the reduction runs on the device and
 * only three doubles come back.  Synchronises `stream`. */
THESEUS_API int theseus_rhs_estimate_stability(theseus_rhs *h, const double *u,
                                   theseus_stability *out, void *stream);

/* One Shu-Osher SSP-RK3 step of u in place (what mfem::RK3SSPSolver::Step
 * does around Mult, src/Simulation.cpp:1255), stage times t, t+dt, t+dt/2.
 * Uses handle-owned stage buffers; asynchronous on `stream`. */
THESEUS_API int theseus_rhs_step_ssprk3(theseus_rhs *h, double *u, double t, double dt, void *stream);

/* Copies the per-element blending coefficient alpha of the last eval/step
 * (ne doubles, the reference's operator_cache.alpha,
 * RHSOperator_impl.hpp:411-437) into the device buffer `alpha`; fails with
 * THESEUS_ERR_INVALID_ARGUMENT when blending is off.  Asynchronous. */
THESEUS_API int theseus_rhs_get_alpha(const theseus_rhs *h, double *alpha, void *stream);

/* Integral measures (dgsem_cache.hpp:16-26): node-quadrature integrals
 * (weights J w_i w_j w_k, dgsem_cache_utilities.hpp:485-491) of the density,
 * the kinetic energy density 0.5|m|^2/rho and the total energy density, and
 * the nodal extremes of pressure, temperature and density.  As in the
 * reference the maxima start at 0 and the minima at 1e32. */
typedef struct theseus_integral_measures {
  double mass, ke, en;
  double max_press, min_press;
  double max_temp, min_temp;
  double max_dens, min_dens;
} theseus_integral_measures;

/* Replaces RHSOperator<P>::ComputeIntegralMeasures(u, diag)
 * (RHSOperator_impl.hpp:530-717): per-workgroup partials and a fixed-order
 * second pass on the device (the result does not depend on scheduling);
 * nine doubles come back.  The first call whose mass is nonzero becomes the
 * baseline (RHSOperator_impl.hpp:711-713).  Synchronises `stream`. */
THESEUS_API int theseus_rhs_integral_measures(theseus_rhs *h, const double *u,
                                  theseus_integral_measures *out, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the integral measures of theseus_rhs_integral_measures,
 * computed by the same kernels in the same fixed order (bit-identical
 * results), written as nine doubles in theseus_integral_measures order to
 * the DEVICE pointer `out`, asynchronously on `stream`, with no host
 * synchronisation and no allocation.  The workgroup partials live in a fixed
 * device array of the library (65,536 workgroups * 9 doubles = 4.5 MiB per
 * GPU, held until the process ends): a case needing more than 65,536
 * workgroups (more than 65,536 elements at 3-D p >= 4, 131,072 at 3-D
 * p = 3 with its two elements per workgroup, more at lower orders) fails
 * with THESEUS_ERR_CAPACITY.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams, other threads, other handles) race and
 * must be ordered by the caller.  The handle's device must be current.  The
 * baseline of theseus_rhs_integral_measures_baseline is not updated. */
THESEUS_API int theseus_rhs_integral_measures_noalloc(theseus_rhs *h, const double *u, double *out, void *stream);
/* Entropy production of a right-hand side, computed in FP64 whatever the
 * handle's precision: production = sum over elements e and nodes i of
 * w_i J_e S(u_i) . dudt_i, magnitude = the same sum of |S . dudt|, with S
 * the entropy variables ((gamma - s)/(gamma-1) - beta |v|^2/2, beta v, -beta),
 * s = ln p - gamma ln rho, beta = rho/p (EOS.hpp:206-234), and w_i the GLL
 * quadrature weights on [0,1]^dim.  It is d/dt of the total mathematical
 * entropy sum w J (-rho s/(gamma-1)) under du/dt = dudt: <= 0 (up to
 * rounding) for an entropy-stable discretisation on a periodic mesh, so it
 * checks a mixed-precision right-hand side in FP64.  No counterpart in the
 * reference.  u and dudt are device L-vectors in the same layout (file or
 * aligned frames alike: the sum does not depend on the node order inside an
 * element).  With the tabulated gas S = ((g - |v|^2/2)/T, v/T, -1/T) at the
 * temperature recovered from u (LTEHelmholtzGasEOS.hpp:144-157), the
 * gradient of the entropy density rho H_T(rho, T).
 * Per-workgroup partials and a fixed-order second pass (deterministic);
 * synchronises `stream`. */
typedef struct theseus_entropy_production {
  double production, magnitude;
} theseus_entropy_production;
THESEUS_API int theseus_rhs_entropy_production(theseus_rhs *h, const double *u, const double *dudt,
                                               theseus_entropy_production *out, void *stream);
/* RHSOperator::GetIntegralMeasuresBaseline (RHSOperator.hpp:92); all zero
 * before the first theseus_rhs_integral_measures call. */
THESEUS_API int theseus_rhs_integral_measures_baseline(const theseus_rhs *h, theseus_integral_measures *out);

/* Device-resident clock of a variable-step run (the host loop of
 * src/Simulation.cpp:1136-1158 and 1240-1301, TimeStepControl.hpp:15-48, kept
 * on the device so that steps need no host round trip and can be captured in
 * a graph).  The caller allocates it in device memory and initialises it. */
typedef struct theseus_clock {
  double t;             /* time reached */
  double dt;            /* nominal step: cfl / total rate, at most max_dt_growth
                           times the previous one; 0 before the first step */
  double dt_real;       /* length of the last step: dt shortened to land on
                           t_final; 0 when no step was taken */
  double t_final;       /* end time; +inf for none */
  double cfl;           /* target CFL (runTime "cfl") */
  double max_dt_growth; /* runTime "max_dt_growth" (>= 1; <= 0: no limit) */
  double rate;          /* total stability rate of the last estimate */
  long long steps;      /* steps taken */
} theseus_clock;

/* One variable SSP-RK3 step of u in place with the step chosen on the
 * device: stability estimate of u, dt = cfl / rate (limited to
 * max_dt_growth * the previous dt), done when t >= t_final - 1e-8 dt (then no
 * step is taken and u is left bit for bit as it is), else
 * dt_real = min(dt, time to t_final), the three stages, t += dt_real.
 * A total rate that is not positive and finite takes no step (rate records
 * it).  Asynchronous on `stream`; `clock` is a device pointer. */
THESEUS_API int theseus_rhs_step_ssprk3_cfl(theseus_rhs *h, double *u, theseus_clock *clock, void *stream);

/* ---- instrumentation (no counterpart in the reference) -------------------
 * Per-kernel HIP-event timing for measurement tools: when enabled, every
 * kernel the handle enqueues is bracketed by hipEvents on its own stream, up
 * to `capacity` launches per kernel kind.  Kinds: 0 = the fused
 * volume+face kernel (Euler) / Navier-Stokes pass 2, 1 = Navier-Stokes
 * pass 1 (BR1 gradient + viscous flux) or, for the tabulated LTE gas, the
 * per-node temperature recovery.  enable(h, 0) disables and clears. */
THESEUS_API int theseus_rhs_enable_kernel_timing(theseus_rhs *h, int capacity);
/* Synchronises the recorded events; returns the mean duration in ms and the
 * number of timed launches of one kernel kind. */
THESEUS_API int theseus_rhs_kernel_timing(theseus_rhs *h, int kind, double *mean_ms, int *count);

/* ---- element orientation (the counterpart of MFEM's face orientations,
 * dgsem_cache_utilities.hpp:533-541, GetPermutedIndex) -------------------
 * Aligns the elements of a conforming quad/hex mesh: from each element's
 * corner vertex ids in its own lexicographic corner order (corner c has
 * local coordinates ((c>>0)&1, (c>>1)&1, (c>>2)&1); MFEM's counterclockwise
 * quad corners 0,1,2,3 are lexicographic 0,1,3,2, its hex corners 0..7 are
 * 0,1,3,2,4,5,7,6), with periodic vertices already identified, a
 * breadth-first walk over face neighbours chooses for every element the
 * symmetry (elem_orient code, as in theseus_case_desc) under which each
 * shared face is seen with the same in-face corner order from both sides.
 * Outputs, each may be NULL: elem_orient[ne]; face_neighbor[ne][2*dim] in
 * the aligned frame, -1 on faces without a neighbour (the caller puts its
 * boundary index there); and, when corner_coords ([ne][2^dim][dim], the
 * corners' physical coordinates in the same storage order) is given,
 * elem_jac[ne] and elem_adj[ne][dim][dim] of the aligned frame, after
 * turning every connected component whose aligned frame is mirrored
 * (negative Jacobians) around.  Fails with THESEUS_ERR_UNSUPPORTED when the
 * mesh cannot be aligned (unstructured vertices), a face has more than two
 * elements, an element is not affine or the Jacobian signs disagree. */
THESEUS_API int theseus_orient_elements(int dim, int num_elements, const long long *elem_vertices,
                                        const double *corner_coords, unsigned char *elem_orient, int *face_neighbor,
                                        double *elem_jac, double *elem_adj);
/* Storage index within its element of the aligned node n (lexicographic) of
 * an element with orientation `code` (see theseus_case_desc); negative on
 * invalid arguments. */
THESEUS_API int theseus_orient_node(int dim, int order, int code, int node);

/* Thread-local message of the last failing call. */
THESEUS_API const char *theseus_last_error(void);

/* ---- serving pool (no counterpart in the reference, whose driver runs one
 * case per process, src/Simulation.cpp:1255) ---------------------------------
 * Many independent cases of one (dim, order, flux pairing, precision) share
 * one device-resident element array: each admitted case occupies a
 * contiguous range of element slots (first fit), its connectivity offset
 * into the array, so one kernel launch per stage advances every case and a
 * case's result does not depend on what else is in the pool.  The pool's
 * capacity is set by an HBM budget; admission fails with
 * THESEUS_ERR_CAPACITY when no free range is large enough, and a released
 * case's slots are reused.  The SSP-RK3 step of all admitted cases is one
 * hipGraph over the occupied slots' high-water mark, rounded up to a ladder
 * of ranges (1024 slots, then steps of 25 %); the graph of each range is
 * captured at its first step and kept, so admissions and releases never
 * re-capture a range seen before.  Each case advances with its own dt,
 * computed on the device at admission from its stability estimate.
 * Admission, step and release only enqueue work on `stream` (host tables go
 * through pinned staging); the calls on one pool must be ordered, e.g. all
 * made on one stream.  The tile
 * shape (elements per workgroup, block order) is picked at create time by
 * timing the candidates on a synthetic case (tile/xcd_remap = -1), or fixed. */
typedef struct theseus_pool theseus_pool; /* opaque */

typedef struct theseus_pool_desc {
  int dim, order, flow_model, volume_flux, surface_flux, precision, blending;
  double gamma, Pr, R_gas, mu, mu_bulk, alpha_max, alpha_min;
  long long hbm_budget_bytes; /* device memory the pool may hold */
  int max_cases;              /* case slots */
  int tile;                   /* -1: autotune; 0: default elements per workgroup; 1: twice that (Euler) */
  int xcd_remap;              /* -1: autotune; 1: contiguous element ranges per XCD; 0: launch order */
} theseus_pool_desc;

THESEUS_API int theseus_pool_create(const theseus_pool_desc *desc, theseus_pool **out);
THESEUS_API int theseus_pool_destroy(theseus_pool *p);
/* Capacity and occupancy; any output may be NULL. */
THESEUS_API int theseus_pool_info(const theseus_pool *p, int *capacity_elements, int *free_elements,
                                  int *largest_free_range, int *active_cases, int *tile, int *xcd_remap);
/* Admit one case (physics and kernel fields must match the pool's; its
 * geometry/connectivity arrays and boundary table (at most
 * THESEUS_POOL_MAX_BC boundaries) are host pointers as for
 * theseus_rhs_create), with initial state u0 (its L-vector, host or device
 * memory) and dt = cfl / total stability rate of u0.  Asynchronous on
 * `stream` when `dt` is NULL and u0 is device or pinned memory (pageable u0
 * is copied as hipMemcpyAsync copies it, which may wait for the stream); a
 * case whose rate is not positive and finite then stays in place with dt = 0.
 * With `dt` non-NULL the call waits for the estimate, returns it, and
 * refuses such a case with THESEUS_ERR_NONFINITE. */
#define THESEUS_POOL_MAX_BC 16
THESEUS_API int theseus_pool_admit(theseus_pool *p, const theseus_case_desc *c, const double *u0, double cfl,
                                   int *case_id, double *dt, void *stream);
/* One SSP-RK3 step of every admitted case (graph replay on `stream`). */
THESEUS_API int theseus_pool_step(theseus_pool *p, void *stream);
/* Copy a case's current state (its L-vector) to `u` (host or device memory); synchronises `stream`. */
THESEUS_API int theseus_pool_read(theseus_pool *p, int case_id, double *u, void *stream);
/* Integral measures of one case's current state (its element range only),
 * as theseus_rhs_integral_measures computes them.  Synchronises `stream`. */
THESEUS_API int theseus_pool_integral_measures(theseus_pool *p, int case_id, theseus_integral_measures *out,
                                               void *stream);
/* Free a case's slots for later admissions (asynchronous on `stream`). */
THESEUS_API int theseus_pool_release(theseus_pool *p, int case_id, void *stream);

/* Pools of different polynomial order in one launch: up to three pools of
 * orders 2, 3 and 4 (one each; 3-D Euler, one flux pairing -- compiled for
 * Ranocha + Lax-Friedrichs -- and one precision) are advanced by one kernel
 * launch per SSP-RK3 stage, the workgroups of each order running that
 * order's code (128 lanes: four elements per workgroup at p = 2, two at
 * p = 3, one at p = 4).  Each pool keeps its element array, admission and
 * release; theseus_pool_group_step replaces theseus_pool_step for the
 * member pools (a captured graph per combination of their launch ranges, kept
 * until the group is destroyed), on the stream their other calls use.  A
 * case's result is that of theseus_pool_step.  The group does not own the
 * pools: destroy it first. */
typedef struct theseus_pool_group theseus_pool_group; /* opaque */
THESEUS_API int theseus_pool_group_create(theseus_pool *const *pools, int n, theseus_pool_group **out);
THESEUS_API int theseus_pool_group_step(theseus_pool_group *g, void *stream);
THESEUS_API int theseus_pool_group_destroy(theseus_pool_group *g);

#ifdef __cplusplus
}
#endif

#endif /* THESEUS_RHS_H */
