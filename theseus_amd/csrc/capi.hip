// C-ABI of the gfx950 DGSEM RHS engine (include/theseus_rhs.h).
//
// theseus_rhs_create builds the device-resident operator cache that the
// reference assembles in RHSOperator::Finalize (RHSOperator_impl.hpp:273-314,
// dgsem_cache_utilities.hpp:18-215,867-943): GLL basis, derivative
// operators, per-element affine geometry, face connectivity and the
// Persson-Peraire modal transform (1-D inverse Legendre Vandermonde).  Evaluation enqueues one kernel for Euler
// and two for Navier-Stokes; the SSP-RK3 step fuses each stage's vector
// update into the RHS kernel's epilogue.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/theseus_rhs.h"
#include "registry.hpp"

namespace thx {

namespace {
std::map<std::tuple<int, int, int, int>, FluxKernels> &flux_table() {
  static std::map<std::tuple<int, int, int, int>, FluxKernels> t;
  return t;
}
std::map<std::tuple<int, int>, ShapeKernels> &shape_table() {
  static std::map<std::tuple<int, int>, ShapeKernels> t;
  return t;
}
}  // namespace

void register_flux(int dim, int p, int vol, int surf, FluxKernels k) { flux_table()[{dim, p, vol, surf}] = k; }
void register_shape(int dim, int p, ShapeKernels k) { shape_table()[{dim, p}] = k; }
const FluxKernels *find_flux(int dim, int p, int vol, int surf) {
  auto it = flux_table().find({dim, p, vol, surf});
  return it == flux_table().end() ? nullptr : &it->second;
}
const ShapeKernels *find_shape(int dim, int p) {
  auto it = shape_table().find({dim, p});
  return it == shape_table().end() ? nullptr : &it->second;
}

// GLL nodes and weights on [0,1]: endpoints and the roots of P'_p, by Newton
// iteration in long double, symmetrised (x_{p-i} = 1 - x_i).
static void gll01(int p, std::vector<double> &x, std::vector<double> &w) {
  const int Np = p + 1, N = p;
  x.assign(Np, 0.0);
  w.assign(Np, 0.0);
  for (int i = 0; i <= N / 2; ++i) {
    long double z = -cosl(3.14159265358979323846264338327950288L * i / N);
    auto legendre = [&](long double zz, long double &PN, long double &PN1) {
      long double P0 = 1, P1 = zz;
      for (int k = 1; k < N; ++k) {
        const long double P2 = ((2 * k + 1) * zz * P1 - k * P0) / (k + 1);
        P0 = P1;
        P1 = P2;
      }
      PN = P1;
      PN1 = P0;
    };
    if (i > 0) {
      for (int it = 0; it < 100; ++it) {
        long double PN, PN1;
        legendre(z, PN, PN1);
        const long double dP = N * (z * PN - PN1) / (z * z - 1);
        const long double d2P = (2 * z * dP - N * (N + 1) * PN) / (1 - z * z);
        const long double dz = dP / d2P;
        z -= dz;
        if (fabsl(dz) < 1e-19L) break;
      }
    }
    long double PN, PN1;
    legendre(z, PN, PN1);
    const long double wt = 2.0L / (N * (N + 1) * PN * PN);
    x[i] = (double)((1.0L + z) / 2.0L);
    w[i] = (double)(wt / 2.0L);
    x[N - i] = 1.0 - x[i];
    w[N - i] = w[i];
  }
  x[0] = 0.0;
  x[N] = 1.0;
  if (Np % 2 == 1) x[N / 2] = 0.5;
}

// Derivative operators exactly as dgsem_cache_utilities.hpp:124-181 forms
// them: barycentric weights, negative-sum diagonal, corner corrections.
static void derivative_ops(int p, const std::vector<double> &x, const std::vector<double> &w, BasisArgs &b) {
  const int Np = p + 1;
  std::vector<double> wB(Np, 1.0), DT(Np * Np, 0.0);
  for (int i = 1; i < Np; i++)
    for (int j = 0; j < i; j++) {
      wB[j] *= (x[j] - x[i]);
      wB[i] *= (x[i] - x[j]);
    }
  for (int i = 0; i < Np; ++i) wB[i] = 1.0 / wB[i];
  for (int iL = 0; iL < Np; iL++)
    for (int i = 0; i < Np; i++)
      if (iL != i) {
        DT[i * Np + iL] = wB[iL] / wB[i] / (x[i] - x[iL]);
        DT[i * Np + i] -= DT[i * Np + iL];
      }
  std::memset(&b, 0, sizeof(b));
  for (int k = 0; k < Np * Np; ++k) {
    b.D[k] = DT[k];
    b.Dhat[k] = DT[k];
    b.Dhat2[k] = DT[k] * 2.0;
  }
  b.Dhat[0] += 1.0 / w[0];
  b.Dhat[(Np - 1) * Np + Np - 1] -= 1.0 / w[Np - 1];
  b.Dhat2[0] += 1.0 / w[0];
  b.Dhat2[(Np - 1) * Np + Np - 1] -= 1.0 / w[Np - 1];
  for (int i = 0; i < Np; ++i) b.w[i] = w[i];
}

// Inverse of the 1-D Legendre Vandermonde V(i,a) = P_a(2 x_i - 1) on the
// GLL nodes (Gauss-Jordan with partial pivoting in long double).  The
// reference inverts the full tensor-product Vandermonde of the element
// (src/ModalBasis.cpp:47-140, MFEM Poly_1D::CalcLegendre on [0,1]); that
// matrix is the Kronecker product of this one.
static void legendre_vandermonde_inverse(int p, const std::vector<double> &x, double *Vinv) {
  const int n = p + 1;
  std::vector<long double> V(n * n), I(n * n, 0.0L);
  for (int i = 0; i < n; ++i) {
    const long double z = 2.0L * x[i] - 1.0L;
    long double P0 = 1.0L, P1 = z;
    V[i * n + 0] = 1.0L;
    if (n > 1) V[i * n + 1] = z;
    for (int a = 1; a + 1 < n; ++a) {
      const long double P2 = ((2 * a + 1) * z * P1 - a * P0) / (a + 1);
      V[i * n + a + 1] = P2;
      P0 = P1;
      P1 = P2;
    }
    I[i * n + i] = 1.0L;
  }
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (fabsl(V[r * n + c]) > fabsl(V[piv * n + c])) piv = r;
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(V[c * n + k], V[piv * n + k]);
        std::swap(I[c * n + k], I[piv * n + k]);
      }
    const long double inv = 1.0L / V[c * n + c];
    for (int k = 0; k < n; ++k) {
      V[c * n + k] *= inv;
      I[c * n + k] *= inv;
    }
    for (int r = 0; r < n; ++r)
      if (r != c) {
        const long double f = V[r * n + c];
        for (int k = 0; k < n; ++k) {
          V[r * n + k] -= f * V[c * n + k];
          I[r * n + k] -= f * I[c * n + k];
        }
      }
  }
  // I = V^{-1}: mode a = sum_i I[a][i] u_i
  for (int k = 0; k < n * n; ++k) Vinv[k] = (double)I[k];
}

// StabilityEstimate.hpp:61-91
static double advection_scale(int order) {
  static const double c[] = {0.0,           2.19706891727, 5.41995189335, 9.64849524786, 14.7297419321,
                             20.5984751391, 27.2141622680, 34.5459214016, 42.5691944324, 51.2637780299,
                             60.6126232210, 70.6010566044, 81.2162499718};
  if (order > 0 && order < 13) return 1.05 * c[order];
  const double pts = order + 1 > 1 ? order + 1 : 1;
  return 0.65 * pts * pts;
}
static double diffusion_scale(int order) {
  static const double c[] = {0.0,           4.0,           25.5969429388, 82.9000427145, 203.470468982,
                             426.230394541, 799.733889878, 1383.0975821,  2244.38177282, 3461.2550528,
                             5121.29136549, 7321.77632194, 10169.7148876};
  if (order > 0 && order < 13) return 1.25 * c[order];
  const double pts = order + 1 > 1 ? order + 1 : 1;
  return 0.5 * pts * pts * pts * pts;
}

// Clears the stability accumulators with a kernel (no memset node, so a
// captured graph replays it correctly).
__global__ void k_zero_u64(unsigned long long *p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0ull;
}

// Background pressure of the mixed mode: the pressure of the first node of
// the state (one lane; a vector store).
template <int DIM>
__global__ void k_pref(const double *U, long long Ntot, GasC<double> g, double *out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  Pt<double, DIM> s;
#pragma unroll
  for (int q = 0; q < DIM + 2; ++q) s.q[q] = U[q * Ntot];
  derive(g, s);
  out[0] = s.p;
}

}  // namespace thx

using namespace thx;

static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                                   \
  do {                                                                                                 \
    hipError_t _e = (expr);                                                                            \
    if (_e != hipSuccess) return fail(THESEUS_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(_e)); \
  } while (0)

struct theseus_rhs {
  theseus_case_desc d;
  int Np, ndof, nfp, neq;
  long long Ntot;
  const FluxKernels *fk = nullptr;
  const ShapeKernels *sk = nullptr;
  BasisArgs basis;
  GasC<double> gas;
  GasC<float> gas32;
  int pm = 0;  // kernel variant: 0 FP64, 1 FP32 flux arithmetic
  double adv_scale, diff_scale, surf_scale;
  double *d_J = nullptr, *d_adj = nullptr;
  int *d_nbr = nullptr;
  double *d_alpha = nullptr;
  double *d_Fv = nullptr;       // NS: viscous volume divergence, then n.Fv at face points
  double *d_pref = nullptr;     // mixed mode: background pressure (one double)
  bool pref_set = false;
  double *d_stage[2] = {nullptr, nullptr};
  unsigned long long *d_rates = nullptr;
  bool has_boundary = false;
  // per-kernel event timing (instrumentation)
  int timing_capacity = 0;
  std::vector<hipEvent_t> ev[2][2];  // [kind][start/stop]
  int ev_used[2] = {0, 0};

  RHSArgs args(const double *U, double *out) const {
    RHSArgs a;
    std::memset(&a, 0, sizeof(a));
    a.U = U;
    a.out = out;
    a.mode = 0;
    if (d_Fv) {
      // Vd then Gf, in the kernels' compute type
      const size_t esz = pm ? sizeof(float) : sizeof(double);
      char *base = reinterpret_cast<char *>(d_Fv);
      a.Vd = base;
      a.Gf = base + esz * (size_t)(d.dim + 1) * Ntot;
      a.Vd_out = base;
      a.Gf_out = base + esz * (size_t)(d.dim + 1) * Ntot;
    }
    a.alpha = d_alpha;
    a.Ntot = Ntot;
    a.ne = d.num_elements;
    a.blending = d.blending;
    a.alpha_max = d.alpha_max;
    a.alpha_min = d.alpha_min;
    a.sharp = 9.21024;                                                 // RHSOperator.hpp:33
    a.thresh = 0.5 * std::pow(10.0, -1.8 * std::pow(d.order, 0.25));  // RHSOperator.hpp:53
    a.J = d_J;
    a.adj = d_adj;
    a.nbr = d_nbr;
    a.gas = gas;
    a.gas32 = gas32;
    a.pref = d_pref;
    a.basis = basis;
    return a;
  }
  int nblocks() const { return (d.num_elements + sk->epb - 1) / sk->epb; }
};

template <typename F>
static void timed(theseus_rhs *h, int kind, hipStream_t s, F &&launch) {
  const int slot = h->ev_used[kind];
  const bool rec = slot < h->timing_capacity;
  if (rec) hipEventRecord(h->ev[kind][0][slot], s);
  launch();
  if (rec) {
    hipEventRecord(h->ev[kind][1][slot], s);
    h->ev_used[kind] = slot + 1;
  }
}

static int enqueue_rhs(theseus_rhs *h, RHSArgs a, hipStream_t s) {
  const int nb = h->nblocks();
  if (h->pm && !h->pref_set) {
    // the mixed mode's background pressure, fixed from the first state evaluated
    if (h->d.dim == 2) hipLaunchKernelGGL(k_pref<2>, dim3(1), dim3(64), 0, s, a.U, h->Ntot, h->gas, h->d_pref);
    else hipLaunchKernelGGL(k_pref<3>, dim3(1), dim3(64), 0, s, a.U, h->Ntot, h->gas, h->d_pref);
    h->pref_set = true;
  }
  if (h->d.flow_model == THESEUS_FLOW_NAVIER_STOKES) {
    timed(h, 1, s, [&] { h->sk->ns_grad[h->pm](a, nb, s); });
    timed(h, 0, s, [&] { h->fk->ns[h->pm](a, nb, s); });
  } else {
    timed(h, 0, s, [&] { h->fk->euler[h->pm](a, nb, s); });
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(THESEUS_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  return THESEUS_OK;
}

template <typename T>
static int upload(T **dst, const T *src, size_t n) {
  HIP_TRY(hipMalloc((void **)dst, n * sizeof(T) + 8));
  if (src) HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return THESEUS_OK;
}

extern "C" {

const char *theseus_last_error(void) { return g_last_error.c_str(); }

int theseus_rhs_create(const theseus_case_desc *desc, theseus_rhs **out) {
  if (!desc || !out) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null descriptor or output");
  *out = nullptr;
  const theseus_case_desc &d = *desc;
  if (d.dim != 2 && d.dim != 3) return fail(THESEUS_ERR_INVALID_ARGUMENT, "dim must be 2 or 3");
  if (d.order < 1 || d.order > 7) return fail(THESEUS_ERR_INVALID_ARGUMENT, "order must be in [1,7]");
  if (d.num_elements < 1) return fail(THESEUS_ERR_INVALID_ARGUMENT, "num_elements must be positive");
  if (!d.elem_jac || !d.elem_adj || !d.face_neighbor)
    return fail(THESEUS_ERR_INVALID_ARGUMENT, "geometry and connectivity arrays are required");
  if (d.flow_model != THESEUS_FLOW_EULER && d.flow_model != THESEUS_FLOW_NAVIER_STOKES)
    return fail(THESEUS_ERR_INVALID_ARGUMENT, "unknown flow_model");
  if (d.precision != THESEUS_PRECISION_FP64 && d.precision != THESEUS_PRECISION_MIXED)
    return fail(THESEUS_ERR_UNSUPPORTED, "precision must be THESEUS_PRECISION_FP64 or THESEUS_PRECISION_MIXED");
  if (!(d.gamma > 1.0) || !std::isfinite(d.gamma)) return fail(THESEUS_ERR_INVALID_ARGUMENT, "gamma must be > 1");
  if (d.flow_model == THESEUS_FLOW_NAVIER_STOKES && (!(d.Pr > 0) || !(d.mu >= 0)))
    return fail(THESEUS_ERR_INVALID_ARGUMENT, "Navier-Stokes needs Pr > 0 and mu >= 0");
  const FluxKernels *fk = find_flux(d.dim, d.order, d.volume_flux, d.surface_flux);
  const ShapeKernels *sk = find_shape(d.dim, d.order);
  if (!fk || !sk) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "no kernel compiled for dim=%d order=%d volume_flux=%d surface_flux=%d", d.dim,
                  d.order, d.volume_flux, d.surface_flux);
    return fail(THESEUS_ERR_UNSUPPORTED, buf);
  }
  const int ne = d.num_elements, dim = d.dim;
  bool has_boundary = false;
  for (long long f = 0; f < (long long)ne * 2 * dim; ++f) {
    const int nb = d.face_neighbor[f];
    const int e = (int)(f / (2 * dim)), lf = (int)(f % (2 * dim));
    if (nb >= ne) return fail(THESEUS_ERR_INVALID_ARGUMENT, "face_neighbor index out of range");
    if (nb >= 0) {
      const int opp = (lf & 1) ? lf - 1 : lf + 1;
      if (d.face_neighbor[(long long)nb * 2 * dim + opp] != e)
        return fail(THESEUS_ERR_INVALID_ARGUMENT, "face_neighbor is not symmetric (element " + std::to_string(e) +
                                                      ", face " + std::to_string(lf) + ")");
    } else {
      has_boundary = true;
      const int bc = -(nb + 1);
      if (bc >= d.num_bc || !d.bc_type) return fail(THESEUS_ERR_INVALID_ARGUMENT, "boundary index out of range");
      if (d.bc_type[bc] != THESEUS_BC_SUPERSONIC_OUTFLOW)
        return fail(THESEUS_ERR_UNSUPPORTED, "only supersonic-outflow (transmissive) boundaries are implemented");
    }
  }
  if (has_boundary && d.flow_model == THESEUS_FLOW_NAVIER_STOKES)
    return fail(THESEUS_ERR_UNSUPPORTED, "Navier-Stokes boundary conditions are not implemented (periodic only)");
  for (int e = 0; e < ne; ++e)
    if (!(d.elem_jac[e] > 0) || !std::isfinite(d.elem_jac[e]))
      return fail(THESEUS_ERR_INVALID_ARGUMENT, "element Jacobian must be positive and finite");

  theseus_rhs *h = new theseus_rhs();
  h->d = d;
  h->d.elem_jac = nullptr;
  h->d.elem_adj = nullptr;
  h->d.face_neighbor = nullptr;
  h->d.bc_type = nullptr;
  h->fk = fk;
  h->sk = sk;
  h->Np = d.order + 1;
  h->ndof = dim == 2 ? h->Np * h->Np : h->Np * h->Np * h->Np;
  h->nfp = dim == 2 ? h->Np : h->Np * h->Np;
  h->neq = dim + 2;
  h->Ntot = (long long)ne * h->ndof;
  h->has_boundary = has_boundary;
  std::vector<double> x, w;
  gll01(d.order, x, w);
  derivative_ops(d.order, x, w, h->basis);
  legendre_vandermonde_inverse(d.order, x, h->basis.Vinv);
  // PhysicsConstants (Physics.hpp:33-38)
  h->gas.gamma = d.gamma;
  h->gas.gammaM1 = d.gamma - 1.0;
  h->gas.gammaM1Inverse = 1.0 / h->gas.gammaM1;
  h->gas.cp = (d.gamma * h->gas.gammaM1Inverse) * d.R_gas;
  h->gas.PrInverse = d.Pr > 0 ? 1.0 / d.Pr : 0.0;
  h->gas.mu = d.mu;
  h->gas.mu_bulk = d.mu_bulk;
  h->gas.R_gas = d.R_gas;
  h->gas32.gamma = (float)h->gas.gamma;
  h->gas32.gammaM1 = (float)h->gas.gammaM1;
  h->gas32.gammaM1Inverse = (float)h->gas.gammaM1Inverse;
  h->gas32.cp = (float)h->gas.cp;
  h->gas32.PrInverse = (float)h->gas.PrInverse;
  h->gas32.mu = (float)h->gas.mu;
  h->gas32.mu_bulk = (float)h->gas.mu_bulk;
  h->gas32.R_gas = (float)h->gas.R_gas;
  h->pm = d.precision == THESEUS_PRECISION_MIXED ? 1 : 0;
  h->adv_scale = advection_scale(d.order);
  h->diff_scale = diffusion_scale(d.order);
  h->surf_scale = h->adv_scale / (1.0 / w[0]);

  int rc = THESEUS_OK;
  const size_t M = (size_t)h->Ntot * h->neq;
  if ((rc = upload(&h->d_J, d.elem_jac, ne)) || (rc = upload(&h->d_adj, d.elem_adj, (size_t)ne * dim * dim)) ||
      (rc = upload(&h->d_nbr, d.face_neighbor, (size_t)ne * 2 * dim)) ||
      (rc = upload<double>(&h->d_alpha, nullptr, ne)) ||
      (rc = upload<unsigned long long>(&h->d_rates, nullptr, 4)) || (rc = upload<double>(&h->d_pref, nullptr, 1)) ||
      (rc = upload<double>(&h->d_stage[0], nullptr, M)) || (rc = upload<double>(&h->d_stage[1], nullptr, M))) {
    theseus_rhs_destroy(h);
    return rc;
  }
  if (d.flow_model == THESEUS_FLOW_NAVIER_STOKES) {
    const size_t nv = (size_t)(dim + 1) * ((size_t)h->Ntot + (size_t)ne * 2 * dim * h->nfp);
    if ((rc = upload<double>(&h->d_Fv, nullptr, nv))) {
      theseus_rhs_destroy(h);
      return rc;
    }
  }
  hipError_t e = hipMemset(h->d_alpha, 0, ne * sizeof(double));
  if (e != hipSuccess) {
    theseus_rhs_destroy(h);
    return fail(THESEUS_ERR_DEVICE, std::string("hipMemset: ") + hipGetErrorString(e));
  }
  *out = h;
  return THESEUS_OK;
}

static void release_events(theseus_rhs *h) {
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j) {
      for (hipEvent_t e : h->ev[k][j]) hipEventDestroy(e);
      h->ev[k][j].clear();
    }
  h->ev_used[0] = h->ev_used[1] = 0;
  h->timing_capacity = 0;
}

int theseus_rhs_enable_kernel_timing(theseus_rhs *h, int capacity) {
  if (!h || capacity < 0) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null handle or negative capacity");
  release_events(h);
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j) {
      h->ev[k][j].resize(capacity);
      for (int i = 0; i < capacity; ++i) HIP_TRY(hipEventCreate(&h->ev[k][j][i]));
    }
  h->timing_capacity = capacity;
  return THESEUS_OK;
}

int theseus_rhs_kernel_timing(theseus_rhs *h, int kind, double *mean_ms, int *count) {
  if (!h || kind < 0 || kind > 1 || !mean_ms || !count) return fail(THESEUS_ERR_INVALID_ARGUMENT, "bad arguments");
  const int n = h->ev_used[kind];
  double tot = 0.0;
  for (int i = 0; i < n; ++i) {
    HIP_TRY(hipEventSynchronize(h->ev[kind][1][i]));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev[kind][0][i], h->ev[kind][1][i]));
    tot += ms;
  }
  *count = n;
  *mean_ms = n ? tot / n : 0.0;
  return THESEUS_OK;
}

int theseus_rhs_destroy(theseus_rhs *h) {
  if (!h) return THESEUS_OK;
  release_events(h);
  hipFree(h->d_J);
  hipFree(h->d_adj);
  hipFree(h->d_nbr);
  hipFree(h->d_alpha);
  hipFree(h->d_Fv);
  hipFree(h->d_pref);
  hipFree(h->d_stage[0]);
  hipFree(h->d_stage[1]);
  hipFree(h->d_rates);
  delete h;
  return THESEUS_OK;
}

int theseus_rhs_sizes(const theseus_rhs *h, long long *ndof_total, int *num_equations) {
  if (!h) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null handle");
  if (ndof_total) *ndof_total = h->Ntot;
  if (num_equations) *num_equations = h->neq;
  return THESEUS_OK;
}

int theseus_rhs_eval(theseus_rhs *h, const double *u, double *dudt, void *stream) {
  if (!h || !u || !dudt) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null handle or vector");
  if ((const double *)dudt == u) return fail(THESEUS_ERR_INVALID_ARGUMENT, "u and dudt must not alias");
  RHSArgs a = h->args(u, dudt);
  return enqueue_rhs(h, a, (hipStream_t)stream);
}

int theseus_rhs_step_ssprk3(theseus_rhs *h, double *u, double t, double dt, void *stream) {
  (void)t;  // the operator is autonomous (no time-dependent sources or boundary data)
  if (!h || !u) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null handle or vector");
  if (!std::isfinite(dt)) return fail(THESEUS_ERR_INVALID_ARGUMENT, "dt must be finite");
  hipStream_t s = (hipStream_t)stream;
  double *A = h->d_stage[0], *B = h->d_stage[1];
  int rc;
  // stage 1: A = u + dt L(u)
  RHSArgs a = h->args(u, A);
  a.mode = 1;
  a.dt = dt;
  if ((rc = enqueue_rhs(h, a, s))) return rc;
  // stage 2: B = 3/4 u + 1/4 (A + dt L(A))
  a = h->args(A, B);
  a.mode = 2;
  a.dt = dt;
  a.X = u;
  a.cx = 3. / 4;
  a.cy = 1. / 4;
  if ((rc = enqueue_rhs(h, a, s))) return rc;
  // stage 3: u = 1/3 u + 2/3 (B + dt L(B))
  a = h->args(B, u);
  a.mode = 2;
  a.dt = dt;
  a.X = u;
  a.cx = 1. / 3;
  a.cy = 2. / 3;
  return enqueue_rhs(h, a, s);
}

int theseus_rhs_estimate_stability(theseus_rhs *h, const double *u, theseus_stability *out, void *stream) {
  if (!h || !u || !out) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null handle, vector or output");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, h->d_rates, 3);
  StabArgs a;
  std::memset(&a, 0, sizeof(a));
  a.U = u;
  a.rates = h->d_rates;
  a.Ntot = h->Ntot;
  a.ne = h->d.num_elements;
  a.viscous = h->d.flow_model == THESEUS_FLOW_NAVIER_STOKES;
  a.adv_scale = h->adv_scale;
  a.diff_scale = h->diff_scale;
  a.surf_scale = h->surf_scale;
  a.J = h->d_J;
  a.adj = h->d_adj;
  a.nbr = h->d_nbr;
  a.gas = h->gas;
  a.w0 = h->basis.w[0];
  h->sk->stability(a, h->nblocks(), s);
  HIP_TRY(hipGetLastError());
  unsigned long long r[3];
  HIP_TRY(hipMemcpyAsync(r, h->d_rates, sizeof r, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  double v[3];
  std::memcpy(v, r, sizeof v);
  out->advective_rate = v[0];
  out->surface_rate = v[1];
  out->diffusive_rate = v[2];
  return THESEUS_OK;
}

int theseus_rhs_get_alpha(const theseus_rhs *h, double *alpha, void *stream) {
  if (!h || !alpha) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null handle or output");
  if (!h->d.blending) return fail(THESEUS_ERR_INVALID_ARGUMENT, "blending is off for this operator");
  HIP_TRY(hipMemcpyAsync(alpha, h->d_alpha, sizeof(double) * h->d.num_elements, hipMemcpyDeviceToDevice,
                         (hipStream_t)stream));
  return THESEUS_OK;
}

}  // extern "C"
