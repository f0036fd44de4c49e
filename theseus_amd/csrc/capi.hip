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

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <climits>
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
std::map<std::tuple<int, int, int, int>, TileKernels> &tile_table() {
  static std::map<std::tuple<int, int, int, int>, TileKernels> t;
  return t;
}
void register_tiles(int dim, int p, int vol, int surf, TileKernels k) { tile_table()[{dim, p, vol, surf}] = k; }
const TileKernels *find_tiles(int dim, int p, int vol, int surf) {
  auto it = tile_table().find({dim, p, vol, surf});
  return it == tile_table().end() ? nullptr : &it->second;
}
std::map<std::tuple<int, int>, OrdersKernels> &orders_table() {
  static std::map<std::tuple<int, int>, OrdersKernels> t;
  return t;
}
void register_orders(int vol, int surf, OrdersKernels k) { orders_table()[{vol, surf}] = k; }
const OrdersKernels *find_orders(int vol, int surf) {
  auto it = orders_table().find({vol, surf});
  return it == orders_table().end() ? nullptr : &it->second;
}
std::map<std::tuple<int, int>, LteKernels> &lte_table() {
  static std::map<std::tuple<int, int>, LteKernels> t;
  return t;
}
// merges: the LTEEC file registers the shared kernels, the others one pairing each
void register_lte(int dim, int p, LteKernels k) {
  LteKernels &e = lte_table()[{dim, p}];
  for (int fx = 0; fx < LTE_NFLUX; ++fx)
    for (int pm = 0; pm < 2; ++pm) {
      if (k.rhs[fx][pm]) e.rhs[fx][pm] = k.rhs[fx][pm];
      if (k.ns[fx][pm]) e.ns[fx][pm] = k.ns[fx][pm];
    }
  if (k.nodes) {
    e.nodes = k.nodes;
    e.ns_grad[0] = k.ns_grad[0];
    e.ns_grad[1] = k.ns_grad[1];
    e.pref = k.pref;
    e.tpb = k.tpb;
    e.epb = k.epb;
  }
}
// The pairings of SimFactory_impl.hpp:140-245 (LteFlux), -1 for any other
int lte_flux_of(int vol, int surf) {
  if (vol == THESEUS_VOLUME_LTEEC && surf == THESEUS_SURFACE_LTEEC) return LTE_FLUX_LTEEC;
  if (vol == THESEUS_VOLUME_PIROZZOLI && surf == THESEUS_SURFACE_PIROZZOLI) return LTE_FLUX_PIROZZOLI;
  if (vol == THESEUS_VOLUME_CENTRAL && surf == THESEUS_SURFACE_HLL) return LTE_FLUX_HLL;
  if (vol == THESEUS_VOLUME_CENTRAL && surf == THESEUS_SURFACE_LLF) return LTE_FLUX_LLF;
  return -1;
}
const LteKernels *find_lte(int dim, int p) {
  auto it = lte_table().find({dim, p});
  return it == lte_table().end() ? nullptr : &it->second;
}
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
// One workgroup of 256: thread t combines partials t, t+256, ... in order,
// then a fixed binary tree.
__global__ void __launch_bounds__(256) k_measures_final(const double *partial, int nb, double *out) {
  __shared__ double red[256][NMEAS];
  const int tid = threadIdx.x;
  double m[NMEAS];
  meas_identity(m);
  for (int b = tid; b < nb; b += 256) meas_combine(m, partial + (size_t)b * NMEAS);
#pragma unroll
  for (int c = 0; c < NMEAS; ++c) red[tid][c] = m[c];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) {
      meas_combine(m, red[tid + st]);
#pragma unroll
      for (int c = 0; c < NMEAS; ++c) red[tid][c] = m[c];
    }
    __syncthreads();
  }
  if (tid < NMEAS) out[tid] = red[0][tid];
}

// Entropy production (theseus_rhs_entropy_production): one lane per node,
// FP64; per-workgroup sums of w J S.dU and w J |S.dU| in a fixed tree order.
// With the tabulated gas (lte.nx > 0) S = ((g - |u|^2/2)/T, u/T, -1/T) at the
// node's recovered temperature (LTEHelmholtzGasEOS.hpp:144-157).
struct GllWeights {
  double w[MAXNP];  // GLL weights on [0,1]
};
template <int DIM>
__global__ void __launch_bounds__(256) k_entropy_production(const double *U, const double *dU, long long Ntot,
                                                             int ndof, int Np, const double *J, GasC<double> gas,
                                                             lte::View lv, GllWeights gw, double *partial) {
  __shared__ double red[256][2];
  const int tid = threadIdx.x;
  const long long t = (long long)blockIdx.x * 256 + tid;
  double pr = 0.0, mg = 0.0;
  if (t < Ntot) {
    double q[DIM + 2], S[DIM + 2];
#pragma unroll
    for (int c = 0; c < DIM + 2; ++c) q[c] = U[c * Ntot + t];
    if (lv.nx > 0) {
      const lte::Eval<double> ev = lte::state_eval<DIM>(lv, q);
      const double beta = 1 / ev.T;
      double m2 = 0.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) m2 += q[1 + d] * q[1 + d];
      const double kinetic = 0.5 * m2 / q[0] / q[0];
      S[0] = beta * (ev.gibbs() - kinetic);
#pragma unroll
      for (int d = 0; d < DIM; ++d) S[1 + d] = beta * (q[1 + d] / q[0]);
      S[DIM + 1] = -beta;
    } else {
      entropy_state<double, DIM>(gas, q, S);
    }
    double sd = 0.0;
#pragma unroll
    for (int c = 0; c < DIM + 2; ++c) sd += S[c] * dU[c * Ntot + t];
    const int e = (int)(t / ndof);
    const int node = (int)(t - (long long)e * ndof);
    double w = gw.w[node % Np];
    if (DIM > 1) w *= gw.w[(node / Np) % Np];
    if (DIM > 2) w *= gw.w[node / (Np * Np)];
    const double wj = w * J[e];
    pr = wj * sd;
    mg = wj * fabs(sd);
  }
  red[tid][0] = pr;
  red[tid][1] = mg;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) {
      red[tid][0] += red[tid + st][0];
      red[tid][1] += red[tid + st][1];
    }
    __syncthreads();
  }
  if (tid == 0) {
    partial[2 * (size_t)blockIdx.x] = red[0][0];
    partial[2 * (size_t)blockIdx.x + 1] = red[0][1];
  }
}

// One workgroup of 256: thread t sums pairs t, t+256, ... in order, then a
// fixed binary tree; out[0..1] = the totals.
__global__ void __launch_bounds__(256) k_sum_pairs_final(const double *partial, int nb, double *out) {
  __shared__ double red[256][2];
  const int tid = threadIdx.x;
  double a = 0.0, b = 0.0;
  for (int k = tid; k < nb; k += 256) {
    a += partial[2 * (size_t)k];
    b += partial[2 * (size_t)k + 1];
  }
  red[tid][0] = a;
  red[tid][1] = b;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) {
      red[tid][0] += red[tid + st][0];
      red[tid][1] += red[tid + st][1];
    }
    __syncthreads();
  }
  if (tid < 2) out[tid] = red[0][tid];
}

// rates[q] = max over the slots of rate q (one 64-lane workgroup).
__global__ void __launch_bounds__(64) k_rates_fold(unsigned long long *rates) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    unsigned long long v = t < kRateSlots ? rates[3 + q * kRateSlots + t] : 0ull;
    // non-negative doubles order like their bit patterns
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(v, off, 64);
      v = o > v ? o : v;
    }
    if (t == 0) rates[q] = v;
  }
}

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

// ---------------------------------------------------------------------------
// Element orientation: the symmetries of [0,1]^dim (codes as in
// theseus_case_desc) and the pass that moves a state between the storage
// frames of the elements and the aligned frame the kernels work in.
// ---------------------------------------------------------------------------
struct Sym {
  int perm[3];  // aligned axis k is storage axis perm[k] ...
  int flip[3];  // ... reversed when flip[k]
};

// In code order: axis permutations in lexicographic order, then the flips
// with aligned axis 0 the slowest bit.
static std::vector<Sym> symmetries(int dim) {
  std::vector<Sym> out;
  int perm[3] = {0, 1, 2};
  do {
    for (int f = 0; f < (1 << dim); ++f) {
      Sym y{};
      for (int k = 0; k < dim; ++k) {
        y.perm[k] = perm[k];
        y.flip[k] = (f >> (dim - 1 - k)) & 1;
      }
      out.push_back(y);
    }
  } while (std::next_permutation(perm, perm + dim));
  return out;
}

// Storage corner of aligned corner c (corner bits = local coordinates).
static int sym_corner(const Sym &y, int dim, int c) {
  int old = 0;
  for (int k = 0; k < dim; ++k) old |= (((c >> k) & 1) ^ y.flip[k]) << y.perm[k];
  return old;
}

// Storage node index of aligned node n (lexicographic, Np nodes per axis).
static int sym_node(const Sym &y, int dim, int Np, int n) {
  int sidx[3] = {0, 0, 0};
  for (int k = 0, r = n; k < dim; ++k, r /= Np) {
    const int a = r % Np;
    sidx[y.perm[k]] = y.flip[k] ? Np - 1 - a : a;
  }
  int out = 0;
  for (int k = dim - 1; k >= 0; --k) out = out * Np + sidx[k];
  return out;
}

// dst <- src between storage and aligned frames: one lane per (element,
// aligned node), all equations.  tab[code * ND + n] = storage index of
// aligned node n.
template <bool TO_ALIGNED>
__global__ void __launch_bounds__(256) k_orient(const double *src, double *dst, const unsigned char *code,
                                                const int *tab, int ND, long long Ntot, int neq) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Ntot) return;
  const long long e = t / ND;
  const int n = (int)(t - e * ND);
  const long long st = e * ND + tab[(int)code[e] * ND + n];
  for (int q = 0; q < neq; ++q) {
    if (TO_ALIGNED) dst[q * Ntot + t] = src[q * Ntot + st];
    else dst[q * Ntot + st] = src[q * Ntot + t];
  }
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
  const LteKernels *lk = nullptr;  // tabulated LTE gas
  int lfx = 0;                     // its flux pairing (LteFlux)
  lte::View ltev;                  // its device arrays
  double *d_lte = nullptr;
  float *d_lte32 = nullptr;   // mixed precision: the planes in FP32
  double *d_lteT = nullptr;   // per node: recovered temperature of the stage state
  int *d_ltecell = nullptr;   // per node: its table cell
  double *d_ltepcg = nullptr; // per node, pairs other than LTEEC: pressure, sound speed (, gamma)
  BasisArgs basis;
  GasC<double> gas;
  GasC<float> gas32;
  int pm = 0;  // kernel variant: 0 FP64, 1 FP32 flux arithmetic
  double adv_scale, diff_scale, surf_scale;
  double *d_J = nullptr, *d_adj = nullptr;
  int *d_nbr = nullptr;
  int *d_bc_type = nullptr;     // [max(num_bc, 1)]
  double *d_bc_data = nullptr;  // [max(num_bc, 1)][THESEUS_BC_DATA_STRIDE]
  double *d_alpha = nullptr;
  double *d_Fv = nullptr;       // NS: viscous volume divergence, then n.Fv at face points
  double *d_pref = nullptr;     // mixed mode: background pressure (one double)
  bool pref_set = false;
  double *d_stage[2] = {nullptr, nullptr};
  unsigned long long *d_rates = nullptr;
  double *d_meas = nullptr;           // integral measures: [nblocks][9] partials, then the 9 results
  double *d_ent = nullptr;            // entropy production: [nodes/256][2] partials, then the 2 results
  theseus_integral_measures diag0{};  // baseline: the first measures with nonzero mass
  bool has_boundary = false;
  int device = 0;  // the device current at create time, which holds the handle's arrays
  // element orientation: codes, node tables, and the aligned copies of the
  // input and output state (only when some element is not stored aligned)
  unsigned char *d_orient = nullptr;
  int *d_otab = nullptr;
  double *d_ual = nullptr, *d_dual = nullptr;
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
    a.bc_type = d_bc_type;
    a.bc_data = d_bc_data;
    a.gas = gas;
    a.gas32 = gas32;
    a.pref = d_pref;
    a.xcd_remap = 1;
    a.lte = ltev;
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
    // the mixed mode's background pressure, fixed from the first state
    // evaluated; that evaluation must run eagerly: recorded into a graph, the
    // kernel computing it would rerun from each replay's state
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    const hipError_t ce = hipStreamIsCapturing(s, &cs);
    if (ce != hipSuccess) return fail(THESEUS_ERR_DEVICE, std::string("hipStreamIsCapturing: ") + hipGetErrorString(ce));
    if (cs != hipStreamCaptureStatusNone)
      return fail(THESEUS_ERR_INVALID_ARGUMENT,
                  "mixed precision: the first evaluation of a handle fixes its background pressure and must not be "
                  "captured into a graph; evaluate or step once before capturing");
    if (h->lk) h->lk->pref(a, 1, s);  // the tabulated gas's pressure (FP64 recovery)
    else if (h->d.dim == 2) hipLaunchKernelGGL(k_pref<2>, dim3(1), dim3(64), 0, s, a.U, h->Ntot, h->gas, h->d_pref);
    else hipLaunchKernelGGL(k_pref<3>, dim3(1), dim3(64), 0, s, a.U, h->Ntot, h->gas, h->d_pref);
    h->pref_set = true;
  }
  if (h->lk) {
    if (h->d.flow_model == THESEUS_FLOW_NAVIER_STOKES) {
      h->lk->nodes(a, nb, s);
      timed(h, 1, s, [&] { h->lk->ns_grad[h->pm](a, nb, s); });
      timed(h, 0, s, [&] { h->lk->ns[h->lfx][h->pm](a, nb, s); });
    } else {
      timed(h, 1, s, [&] { h->lk->nodes(a, nb, s); });
      timed(h, 0, s, [&] { h->lk->rhs[h->lfx][h->pm](a, nb, s); });
    }
  } else if (h->d.flow_model == THESEUS_FLOW_NAVIER_STOKES) {
    timed(h, 1, s, [&] { h->sk->ns_grad[h->pm][h->has_boundary](a, nb, s); });
    timed(h, 0, s, [&] { h->fk->ns[h->pm][h->has_boundary](a, nb, s); });
  } else {
    timed(h, 0, s, [&] { h->fk->euler[h->pm][h->has_boundary](a, nb, s); });
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(THESEUS_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  return THESEUS_OK;
}

// The orientation pass of a handle whose elements are not all aligned.
static int orient_pass(const theseus_rhs *h, const double *src, double *dst, bool to_aligned, hipStream_t s) {
  const int nb = (int)((h->Ntot + 255) / 256);
  if (to_aligned)
    hipLaunchKernelGGL(k_orient<true>, dim3(nb), dim3(256), 0, s, src, dst, h->d_orient, h->d_otab, h->ndof, h->Ntot,
                       h->neq);
  else
    hipLaunchKernelGGL(k_orient<false>, dim3(nb), dim3(256), 0, s, src, dst, h->d_orient, h->d_otab, h->ndof,
                       h->Ntot, h->neq);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(THESEUS_ERR_DEVICE, std::string("orientation pass: ") + hipGetErrorString(e));
  return THESEUS_OK;
}

// u in the aligned frame: u itself, or its aligned copy in d_ual.
static int aligned_input(theseus_rhs *h, const double *u, const double **out, hipStream_t s) {
  *out = u;
  if (!h->d_orient) return THESEUS_OK;
  *out = h->d_ual;
  return orient_pass(h, u, h->d_ual, true, s);
}

template <typename T>
static int upload(T **dst, const T *src, size_t n) {
  HIP_TRY(hipMalloc((void **)dst, n * sizeof(T) + 8));
  if (src) HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return THESEUS_OK;
}

// One boundary's type and payload, as the reference's configuration checks
// them (Simulation.cpp:555-820: positive wall temperature, a physical
// exterior state).  Axis boundaries need the axisymmetric formulation.
// The boundary table on the device (at least one entry, zeros where the
// descriptor has no payload).
static int upload_bc(int **type, double **data, const theseus_case_desc &d) {
  const int nb = std::max(d.num_bc, 1);
  std::vector<int> t(nb, THESEUS_BC_SUPERSONIC_OUTFLOW);
  std::vector<double> v((size_t)nb * THESEUS_BC_DATA_STRIDE, 0.0);
  for (int b = 0; b < d.num_bc; ++b) t[b] = d.bc_type[b];
  if (d.bc_data && d.num_bc > 0) std::copy(d.bc_data, d.bc_data + (size_t)d.num_bc * THESEUS_BC_DATA_STRIDE, v.begin());
  int rc = upload(type, t.data(), t.size());
  if (rc) return rc;
  return upload(data, v.data(), v.size());
}

static int validate_bc(const theseus_case_desc &d, int b) {
  const int t = d.bc_type[b];
  const std::string who = "boundary " + std::to_string(b) + ": ";
  if (t == THESEUS_BC_AXIS)
    return fail(THESEUS_ERR_UNSUPPORTED, who + "axis boundaries need the axisymmetric formulation, which is not built");
  if (t < THESEUS_BC_SLIP_WALL || t > THESEUS_BC_NO_SLIP_ISOTHERMAL)
    return fail(THESEUS_ERR_INVALID_ARGUMENT, who + "unknown boundary type " + std::to_string(t));
  const bool needs_data = t == THESEUS_BC_SUPERSONIC_INFLOW || t == THESEUS_BC_PRESCRIBED_STATE ||
                          t == THESEUS_BC_NO_SLIP_ADIABATIC || t == THESEUS_BC_NO_SLIP_ISOTHERMAL;
  if (!needs_data) return THESEUS_OK;
  if (!d.bc_data) return fail(THESEUS_ERR_INVALID_ARGUMENT, who + "this boundary type needs bc_data");
  const double *v = d.bc_data + (size_t)b * THESEUS_BC_DATA_STRIDE;
  const int nv = (t == THESEUS_BC_SUPERSONIC_INFLOW || t == THESEUS_BC_PRESCRIBED_STATE) ? d.dim + 2 : d.dim + 1;
  for (int k = 0; k < nv; ++k)
    if (!std::isfinite(v[k])) return fail(THESEUS_ERR_INVALID_ARGUMENT, who + "payload must be finite");
  if (t == THESEUS_BC_SUPERSONIC_INFLOW || t == THESEUS_BC_PRESCRIBED_STATE) {
    double m2 = 0.0;
    for (int k = 0; k < d.dim; ++k) m2 += v[1 + k] * v[1 + k];
    const double p = (d.gamma - 1.0) * (v[d.dim + 1] - 0.5 * m2 / v[0]);
    if (!(v[0] > 0) || !(p > 0)) return fail(THESEUS_ERR_INVALID_ARGUMENT, who + "exterior state needs positive density and pressure");
  }
  if (t == THESEUS_BC_NO_SLIP_ISOTHERMAL && !(v[d.dim] > 0))
    return fail(THESEUS_ERR_INVALID_ARGUMENT, who + "wall temperature must be positive");
  return THESEUS_OK;
}

static int validate_mesh(const theseus_case_desc &d, bool &has_boundary) {
  const int ne = d.num_elements, dim = d.dim;
  has_boundary = false;
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
    }
  }
  if (d.num_bc < 0) return fail(THESEUS_ERR_INVALID_ARGUMENT, "num_bc must not be negative");
  for (int b = 0; b < d.num_bc; ++b) {
    const int rc = validate_bc(d, b);
    if (rc) return rc;
  }
  for (int e = 0; e < ne; ++e)
    if (!(d.elem_jac[e] > 0) || !std::isfinite(d.elem_jac[e]))
      return fail(THESEUS_ERR_INVALID_ARGUMENT, "element Jacobian must be positive and finite");

  return THESEUS_OK;
}

// PhysicsConstants (Physics.hpp:33-38) in FP64 and rounded to FP32
static void init_gas(const theseus_case_desc &d, GasC<double> &g, GasC<float> &g32) {
  g.gamma = d.gamma;
  g.gammaM1 = d.gamma - 1.0;
  g.gammaM1Inverse = 1.0 / g.gammaM1;
  g.cp = (d.gamma * g.gammaM1Inverse) * d.R_gas;
  g.PrInverse = d.Pr > 0 ? 1.0 / d.Pr : 0.0;
  g.mu = d.mu;
  g.mu_bulk = d.mu_bulk;
  g.R_gas = d.R_gas;
  g32.gamma = (float)g.gamma;
  g32.gammaM1 = (float)g.gammaM1;
  g32.gammaM1Inverse = (float)g.gammaM1Inverse;
  g32.cp = (float)g.cp;
  g32.PrInverse = (float)g.PrInverse;
  g32.mu = (float)g.mu;
  g32.mu_bulk = (float)g.mu_bulk;
  g32.R_gas = (float)g.R_gas;
}

// Workgroup partials of the integral measures for callers that cannot
// allocate (theseus_rhs_integral_measures_noalloc): a file-scope device array
// of NOALLOC_MAX_WORKGROUPS * NMEAS doubles (4.5 MiB), of which each GPU holds
// its own copy until the process ends.  Host code never passes the array's
// name to a kernel (on the host it is only a stand-in symbol): the device
// address of each GPU's copy is looked up once with hipGetSymbolAddress, with
// that device current, and kept per device.
static constexpr int NOALLOC_MAX_WORKGROUPS = 65536;
__device__ __align__(256) double scratch[NOALLOC_MAX_WORKGROUPS * NMEAS];

static std::mutex g_scratch_mu;
static std::map<int, double *> g_scratch_ptr;  // device ordinal -> address of its copy of `scratch`

// The current device's copy of `scratch`; resolved on the first call for
// that device (the caller makes the device current).
static int scratch_for_current_device(double **out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(THESEUS_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    it = g_scratch_ptr.emplace(dev, static_cast<double *>(ptr)).first;
  }
  *out = it->second;
  return THESEUS_OK;
}

extern "C" {

const char *theseus_last_error(void) { return g_last_error.c_str(); }

static int create_impl(const theseus_case_desc *desc, const theseus_lte_table *lt, theseus_rhs **out) {
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
  const LteKernels *lk = nullptr;
  int lfx = -1;
  if (lt) {
    if (d.blending) return fail(THESEUS_ERR_UNSUPPORTED, "the tabulated LTE gas is implemented without blending");
    lfx = lte_flux_of(d.volume_flux, d.surface_flux);
    if (lfx < 0)
      return fail(THESEUS_ERR_INVALID_ARGUMENT,
                  "the tabulated LTE gas takes the LTEEC, Pirozzoli, central/HLL or central/LLF "
                  "volume/surface flux pairs");
    if (lt->nx < 2 || lt->ny < 2 || !lt->rho_grid || !lt->T_grid || !lt->e_grid || !lt->planes || !lt->inverse)
      return fail(THESEUS_ERR_INVALID_ARGUMENT, "LTE table needs nx, ny >= 2 and all arrays");
    if (lt->viscosity_model != 0 && lt->viscosity_model != 1)
      return fail(THESEUS_ERR_INVALID_ARGUMENT, "LTE viscosity_model must be 0 (the mu plane) or 1 (Sutherland)");
    if (lt->viscosity_model == 1 &&
        (!(lt->mu0 >= 0) || !(lt->T0 > 0) || !(lt->Ts >= 0) || !std::isfinite(lt->mu0) || !std::isfinite(lt->T0) ||
         !std::isfinite(lt->Ts)))
      return fail(THESEUS_ERR_INVALID_ARGUMENT, "Sutherland viscosity needs finite mu0 >= 0, T0 > 0, Ts >= 0");
    const double *grids[3] = {lt->rho_grid, lt->T_grid, lt->e_grid};
    const int counts[3] = {lt->nx, lt->ny, lt->ny};
    for (int g = 0; g < 3; ++g)
      for (int i = 0; i < counts[g]; ++i)
        if (!std::isfinite(grids[g][i]) || (i > 0 && !(grids[g][i] > grids[g][i - 1])))
          return fail(THESEUS_ERR_INVALID_ARGUMENT, "LTE table grids must be finite and strictly increasing");
    if (!(lt->rho_grid[0] > 0) || !(lt->T_grid[0] > 0))
      return fail(THESEUS_ERR_INVALID_ARGUMENT, "LTE table density and temperature must be positive");
    lk = find_lte(d.dim, d.order);
    if (!lk || !lk->nodes || !lk->rhs[lfx][0] || !lk->ns[lfx][0])
      return fail(THESEUS_ERR_UNSUPPORTED, "no LTE kernel compiled for this dim/order and flux");
  }
  const FluxKernels *fk = lt ? nullptr : find_flux(d.dim, d.order, d.volume_flux, d.surface_flux);
  const ShapeKernels *sk = find_shape(d.dim, d.order);
  if ((!fk && !lk) || !sk) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "no kernel compiled for dim=%d order=%d volume_flux=%d surface_flux=%d", d.dim,
                  d.order, d.volume_flux, d.surface_flux);
    return fail(THESEUS_ERR_UNSUPPORTED, buf);
  }
  bool has_boundary = false;
  int rc0 = validate_mesh(d, has_boundary);
  if (rc0) return rc0;
  if (lt)
    for (int b = 0; b < d.num_bc; ++b)
      if (d.bc_type[b] != THESEUS_BC_SUPERSONIC_OUTFLOW)
        return fail(THESEUS_ERR_UNSUPPORTED, "the tabulated LTE gas supports supersonic-outflow boundaries only");
  if (lt && d.flow_model == THESEUS_FLOW_NAVIER_STOKES && has_boundary)
    return fail(THESEUS_ERR_UNSUPPORTED, "the tabulated LTE gas with Navier-Stokes is implemented for periodic meshes");
  const int ne = d.num_elements, dim = d.dim;

  int device = 0;
  HIP_TRY(hipGetDevice(&device));
  double *scratch_dev = nullptr;  // set-up: resolve this device's scratch copy once
  const int rs = scratch_for_current_device(&scratch_dev);
  if (rs) return rs;
  theseus_rhs *h = new theseus_rhs();
  h->d = d;
  h->device = device;
  h->d.elem_jac = nullptr;
  h->d.elem_adj = nullptr;
  h->d.face_neighbor = nullptr;
  h->d.bc_type = nullptr;
  h->d.bc_data = nullptr;
  h->fk = fk;
  h->sk = sk;
  h->lk = lk;
  h->lfx = lfx;
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
  init_gas(d, h->gas, h->gas32);
  h->pm = d.precision == THESEUS_PRECISION_MIXED ? 1 : 0;
  h->adv_scale = advection_scale(d.order);
  h->diff_scale = diffusion_scale(d.order);
  h->surf_scale = h->adv_scale / (1.0 / w[0]);

  int rc = THESEUS_OK;
  const size_t M = (size_t)h->Ntot * h->neq;
  if ((rc = upload(&h->d_J, d.elem_jac, ne)) || (rc = upload(&h->d_adj, d.elem_adj, (size_t)ne * dim * dim)) ||
      (rc = upload(&h->d_nbr, d.face_neighbor, (size_t)ne * 2 * dim)) ||
      (rc = upload_bc(&h->d_bc_type, &h->d_bc_data, d)) || (rc = upload<double>(&h->d_alpha, nullptr, ne)) ||
      (rc = upload<unsigned long long>(&h->d_rates, nullptr, kRateWords)) || (rc = upload<double>(&h->d_pref, nullptr, 1)) ||
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
  if (lt) {
    // one device block: rho, T, e grids, the six planes, the inverse table
    const size_t nx = lt->nx, ny = lt->ny;
    std::vector<double> host;
    host.reserve(nx + 2 * ny + 7 * nx * ny);
    host.insert(host.end(), lt->rho_grid, lt->rho_grid + nx);
    host.insert(host.end(), lt->T_grid, lt->T_grid + ny);
    host.insert(host.end(), lt->e_grid, lt->e_grid + ny);
    host.insert(host.end(), lt->planes, lt->planes + 6 * nx * ny);
    host.insert(host.end(), lt->inverse, lt->inverse + nx * ny);
    if ((rc = upload(&h->d_lte, host.data(), host.size()))) {
      theseus_rhs_destroy(h);
      return rc;
    }
    h->ltev.nx = lt->nx;
    h->ltev.ny = lt->ny;
    h->ltev.rho = h->d_lte;
    h->ltev.T = h->d_lte + nx;
    h->ltev.e = h->d_lte + nx + ny;
    h->ltev.planes = h->d_lte + nx + 2 * ny;
    h->ltev.inv = h->d_lte + nx + 2 * ny + 6 * nx * ny;
    if ((rc = upload<double>(&h->d_lteT, nullptr, h->Ntot)) || (rc = upload<int>(&h->d_ltecell, nullptr, h->Ntot))) {
      theseus_rhs_destroy(h);
      return rc;
    }
    h->ltev.node_T = h->d_lteT;
    h->ltev.node_cell = h->d_ltecell;
    if (lfx != LTE_FLUX_LTEEC) {
      const int nf = lfx == LTE_FLUX_PIROZZOLI ? 3 : 2;
      if ((rc = upload<double>(&h->d_ltepcg, nullptr, (size_t)nf * h->Ntot))) {
        theseus_rhs_destroy(h);
        return rc;
      }
      h->ltev.node_p = h->d_ltepcg;
      h->ltev.node_c = h->d_ltepcg + h->Ntot;
      if (nf == 3) h->ltev.node_g = h->d_ltepcg + 2 * h->Ntot;
    }
    h->ltev.sutherland = lt->viscosity_model == 1;
    h->ltev.mu0 = lt->mu0;
    h->ltev.T0 = lt->T0;
    h->ltev.Ts = lt->Ts;
    if (h->pm) {
      // the mixed-precision form's FP32 copy of the planes
      std::vector<float> p32(lt->planes, lt->planes + 6 * nx * ny);
      if ((rc = upload(&h->d_lte32, p32.data(), p32.size()))) {
        theseus_rhs_destroy(h);
        return rc;
      }
      h->ltev.planes32 = h->d_lte32;
    }
  }
  if (d.elem_orient) {
    const std::vector<Sym> syms = symmetries(dim);
    bool any = false;
    for (int el = 0; el < ne; ++el) {
      if (d.elem_orient[el] >= syms.size()) {
        theseus_rhs_destroy(h);
        return fail(THESEUS_ERR_INVALID_ARGUMENT, "elem_orient[" + std::to_string(el) + "] is not a symmetry code (< " +
                                                      std::to_string(syms.size()) + ")");
      }
      any = any || d.elem_orient[el] != 0;
    }
    if (any) {
      std::vector<int> tab(syms.size() * (size_t)h->ndof);
      for (size_t c = 0; c < syms.size(); ++c)
        for (int n = 0; n < h->ndof; ++n) tab[c * h->ndof + n] = sym_node(syms[c], dim, h->Np, n);
      if ((rc = upload(&h->d_orient, d.elem_orient, ne)) || (rc = upload(&h->d_otab, tab.data(), tab.size())) ||
          (rc = upload<double>(&h->d_ual, nullptr, M)) || (rc = upload<double>(&h->d_dual, nullptr, M))) {
        theseus_rhs_destroy(h);
        return rc;
      }
    }
  }
  h->d.elem_orient = nullptr;
  hipError_t e = hipMemset(h->d_alpha, 0, ne * sizeof(double));
  if (e == hipSuccess) e = hipMemset(h->d_pref, 0, sizeof(double));  // defined before the first evaluation sets it
  if (e != hipSuccess) {
    theseus_rhs_destroy(h);
    return fail(THESEUS_ERR_DEVICE, std::string("hipMemset: ") + hipGetErrorString(e));
  }
  *out = h;
  return THESEUS_OK;
}

int theseus_rhs_create(const theseus_case_desc *desc, theseus_rhs **out) { return create_impl(desc, nullptr, out); }

int theseus_rhs_create_lte(const theseus_case_desc *desc, const theseus_lte_table *table, theseus_rhs **out) {
  if (!table) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null LTE table");
  return create_impl(desc, table, out);
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
  hipFree(h->d_bc_type);
  hipFree(h->d_bc_data);
  hipFree(h->d_alpha);
  hipFree(h->d_Fv);
  hipFree(h->d_pref);
  hipFree(h->d_lte);
  hipFree(h->d_lte32);
  hipFree(h->d_lteT);
  hipFree(h->d_ltecell);
  hipFree(h->d_ltepcg);
  hipFree(h->d_stage[0]);
  hipFree(h->d_stage[1]);
  hipFree(h->d_rates);
  hipFree(h->d_meas);
  hipFree(h->d_ent);
  hipFree(h->d_orient);
  hipFree(h->d_otab);
  hipFree(h->d_ual);
  hipFree(h->d_dual);
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
  hipStream_t s = (hipStream_t)stream;
  if (h->d_orient) {
    int rc = orient_pass(h, u, h->d_ual, true, s);
    if (!rc) rc = enqueue_rhs(h, h->args(h->d_ual, h->d_dual), s);
    return rc ? rc : orient_pass(h, h->d_dual, dudt, false, s);
  }
  RHSArgs a = h->args(u, dudt);
  return enqueue_rhs(h, a, s);
}

// The three Shu-Osher stages; dt_dev (device, may be null) overrides dt.
static int enqueue_ssprk3(theseus_rhs *h, double *u, double dt, const double *dt_dev, hipStream_t s) {
  double *A = h->d_stage[0], *B = h->d_stage[1];
  int rc;
  // stage 1: A = u + dt L(u)
  RHSArgs a = h->args(u, A);
  a.mode = 1;
  a.dt = dt;
  a.case_dt = dt_dev;
  if ((rc = enqueue_rhs(h, a, s))) return rc;
  // stage 2: B = 3/4 u + 1/4 (A + dt L(A))
  a = h->args(A, B);
  a.mode = 2;
  a.dt = dt;
  a.case_dt = dt_dev;
  a.X = u;
  a.cx = 3. / 4;
  a.cy = 1. / 4;
  if ((rc = enqueue_rhs(h, a, s))) return rc;
  // stage 3: u = 1/3 u + 2/3 (B + dt L(B))
  a = h->args(B, u);
  a.mode = 2;
  a.dt = dt;
  a.case_dt = dt_dev;
  a.X = u;
  a.cx = 1. / 3;
  a.cy = 2. / 3;
  return enqueue_rhs(h, a, s);
}

// Zeroes the rate slots and launches the stability kernel (no host copy).
static void enqueue_stability(theseus_rhs *h, const double *u, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u64, dim3((kRateWords + 63) / 64), dim3(64), 0, s, h->d_rates, kRateWords);
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
  a.bc_type = h->d_bc_type;
  a.bc_data = h->d_bc_data;
  a.gas = h->gas;
  a.w0 = h->basis.w[0];
  a.lte = h->ltev;
  if (h->lk) {
    // the tabulated gas: each node's temperature and cell recovered once
    // (k_lte_nodes), instead of a Newton solve per node and per face point
    h->lk->nodes(h->args(u, nullptr), 0, s);
    a.lte_nodes = 1;
  }
  h->sk->stability(a, h->nblocks(), s);
  hipLaunchKernelGGL(k_rates_fold, dim3(1), dim3(64), 0, s, h->d_rates);
}

// The variable-step decision of src/Simulation.cpp:1136-1158,1240-1301 and
// TimeStepControl.hpp:15-48 on the device, from the rates just estimated.
__global__ void k_clock(theseus_clock *c, const unsigned long long *rates) {
  if (threadIdx.x != 0) return;
  const double adv = __longlong_as_double((long long)rates[0]);
  const double surf = __longlong_as_double((long long)rates[1]);
  const double diff = __longlong_as_double((long long)rates[2]);
  const double rate = fmax(adv, surf) + diff;  // StabilityEstimate::TotalRate
  c->rate = rate;
  if (!(rate > 0.0) || !isfinite(rate)) {
    c->dt_real = 0.0;
    return;
  }
  const double est = c->cfl / rate;
  double dt = est;
  if (c->dt > 0.0 && c->max_dt_growth > 0.0) dt = fmin(est, c->max_dt_growth * c->dt);
  const double t = c->t;
  double dt_real = 0.0;
  if (!(t >= c->t_final - 1e-8 * dt)) {
    const double tol = 16.0 * 2.220446049250313e-16 * fmax(1.0, fabs(t));  // TimeEventTolerance
    const double delta = c->t_final - t;
    dt_real = (isfinite(c->t_final) && delta > tol) ? fmin(dt, delta) : dt;
  }
  c->dt = dt;
  c->dt_real = dt_real;
  c->t = t + dt_real;
  if (dt_real > 0.0) c->steps += 1;
}

// A pooled case's time step from its stability rates: cfl / total rate, or 0
// (the case does not advance) when the rate is not positive and finite.
__global__ void k_case_dt(const unsigned long long *rates, double cfl, double *dt) {
  if (threadIdx.x != 0) return;
  const double adv = __longlong_as_double((long long)rates[0]);
  const double surf = __longlong_as_double((long long)rates[1]);
  const double diff = __longlong_as_double((long long)rates[2]);
  const double rate = fmax(adv, surf) + diff;
  dt[0] = (rate > 0.0 && isfinite(rate)) ? cfl / rate : 0.0;
}

int theseus_rhs_step_ssprk3(theseus_rhs *h, double *u, double t, double dt, void *stream) {
  (void)t;  // the operator is autonomous (no time-dependent sources or boundary data)
  if (!h || !u) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null handle or vector");
  if (!std::isfinite(dt)) return fail(THESEUS_ERR_INVALID_ARGUMENT, "dt must be finite");
  hipStream_t s = (hipStream_t)stream;
  if (h->d_orient) {
    int rc = orient_pass(h, u, h->d_ual, true, s);
    if (!rc) rc = enqueue_ssprk3(h, h->d_ual, dt, nullptr, s);
    return rc ? rc : orient_pass(h, h->d_ual, u, false, s);
  }
  return enqueue_ssprk3(h, u, dt, nullptr, s);
}

int theseus_rhs_step_ssprk3_cfl(theseus_rhs *h, double *u, theseus_clock *clock, void *stream) {
  if (!h || !u || !clock) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null handle, vector or clock");
  hipStream_t s = (hipStream_t)stream;
  double *ua = u;
  if (h->d_orient) {
    const int rc = orient_pass(h, u, h->d_ual, true, s);
    if (rc) return rc;
    ua = h->d_ual;
  }
  enqueue_stability(h, ua, s);
  hipLaunchKernelGGL(k_clock, dim3(1), dim3(64), 0, s, clock, h->d_rates);
  HIP_TRY(hipGetLastError());
  const int rc = enqueue_ssprk3(h, ua, 0.0, &clock->dt_real, s);
  if (rc || !h->d_orient) return rc;
  return orient_pass(h, h->d_ual, u, false, s);
}

// Integral measures of elements [e0, e0 + ne) of U; partial holds the
// per-workgroup partials followed by the nine results.
static int measures_impl(const ShapeKernels *sk, const double *U, long long Ntot, int e0, int ne, const double *J,
                         const GasC<double> &gas, const BasisArgs &basis, const lte::View &lv, double *partial,
                         theseus_integral_measures *out, hipStream_t s) {
  MeasArgs a;
  std::memset(&a, 0, sizeof(a));
  a.U = U;
  a.partial = partial;
  a.Ntot = Ntot;
  a.e0 = e0;
  a.ne = ne;
  a.J = J;
  a.gas = gas;
  std::memcpy(a.w, basis.w, sizeof(a.w));
  a.lte = lv;
  const int nb = (ne + sk->epb - 1) / sk->epb;
  sk->measures(a, nb, s);
  hipLaunchKernelGGL(k_measures_final, dim3(1), dim3(256), 0, s, partial, nb, partial + (size_t)nb * NMEAS);
  HIP_TRY(hipGetLastError());
  double r[NMEAS];
  HIP_TRY(hipMemcpyAsync(r, partial + (size_t)nb * NMEAS, sizeof r, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  static_assert(sizeof(theseus_integral_measures) == sizeof r, "nine doubles");
  std::memcpy(out, r, sizeof r);
  return THESEUS_OK;
}

int theseus_rhs_integral_measures(theseus_rhs *h, const double *u, theseus_integral_measures *out, void *stream) {
  if (!h || !u || !out) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null handle, vector or output");
  if (!h->d_meas) {
    const int rc = upload<double>(&h->d_meas, nullptr, (size_t)(h->nblocks() + 1) * NMEAS);
    if (rc) return rc;
  }
  const double *ua = u;
  int rc = aligned_input(h, u, &ua, (hipStream_t)stream);
  if (rc) return rc;
  rc = measures_impl(h->sk, ua, h->Ntot, 0, h->d.num_elements, h->d_J, h->gas, h->basis, h->ltev, h->d_meas, out,
                     (hipStream_t)stream);
  if (rc) return rc;
  if (h->diag0.mass == 0.0) h->diag0 = *out;
  return THESEUS_OK;
}

int theseus_rhs_entropy_production(theseus_rhs *h, const double *u, const double *dudt,
                                   theseus_entropy_production *out, void *stream) {
  if (!h || !u || !dudt || !out) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null handle, vector or output");
  const long long nbl = (h->Ntot + 255) / 256;
  if (nbl > INT_MAX) return fail(THESEUS_ERR_INVALID_ARGUMENT, "too many nodes");
  const int nb = (int)nbl;
  if (!h->d_ent) {
    const int rc = upload<double>(&h->d_ent, nullptr, 2 * ((size_t)nb + 1));
    if (rc) return rc;
  }
  hipStream_t s = (hipStream_t)stream;
  const int ndof = h->ndof, Np = h->Np;
  GllWeights gw;
  std::memcpy(gw.w, h->basis.w, sizeof gw.w);
  if (h->d.dim == 2)
    hipLaunchKernelGGL(k_entropy_production<2>, dim3(nb), dim3(256), 0, s, u, dudt, h->Ntot, ndof, Np, h->d_J, h->gas,
                       h->ltev, gw, h->d_ent);
  else
    hipLaunchKernelGGL(k_entropy_production<3>, dim3(nb), dim3(256), 0, s, u, dudt, h->Ntot, ndof, Np, h->d_J, h->gas,
                       h->ltev, gw, h->d_ent);
  hipLaunchKernelGGL(k_sum_pairs_final, dim3(1), dim3(256), 0, s, h->d_ent, nb, h->d_ent + 2 * (size_t)nb);
  HIP_TRY(hipGetLastError());
  double r[2];
  HIP_TRY(hipMemcpyAsync(r, h->d_ent + 2 * (size_t)nb, sizeof r, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  out->production = r[0];
  out->magnitude = r[1];
  return THESEUS_OK;
}

int theseus_rhs_integral_measures_baseline(const theseus_rhs *h, theseus_integral_measures *out) {
  if (!h || !out) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null handle or output");
  *out = h->diag0;
  return THESEUS_OK;
}

int theseus_rhs_integral_measures_noalloc(theseus_rhs *h, const double *u, double *out, void *stream) {
  if (!h || !u || !out) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null handle, vector or output");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev != h->device)
    return fail(THESEUS_ERR_INVALID_ARGUMENT, "the handle's device (" + std::to_string(h->device) +
                                                  ") must be current (device " + std::to_string(dev) + " is)");
  const int nb = h->nblocks();
  if (nb > NOALLOC_MAX_WORKGROUPS)
    return fail(THESEUS_ERR_CAPACITY, "integral measures without allocation: " + std::to_string(nb) +
                                          " workgroups exceed the fixed scratch of " +
                                          std::to_string(NOALLOC_MAX_WORKGROUPS) + " (" +
                                          std::to_string(NOALLOC_MAX_WORKGROUPS * h->sk->epb) +
                                          " elements at this order); use theseus_rhs_integral_measures");
  double *partial = nullptr;
  int rc = scratch_for_current_device(&partial);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const double *ua = u;
  if ((rc = aligned_input(h, u, &ua, s))) return rc;
  // every workgroup writes its own partials, so nothing needs clearing
  MeasArgs a;
  std::memset(&a, 0, sizeof(a));
  a.U = ua;
  a.partial = partial;
  a.Ntot = h->Ntot;
  a.e0 = 0;
  a.ne = h->d.num_elements;
  a.J = h->d_J;
  a.gas = h->gas;
  std::memcpy(a.w, h->basis.w, sizeof(a.w));
  a.lte = h->ltev;
  h->sk->measures(a, nb, s);
  hipLaunchKernelGGL(k_measures_final, dim3(1), dim3(256), 0, s, partial, nb, out);
  HIP_TRY(hipGetLastError());
  return THESEUS_OK;
}

int theseus_rhs_estimate_stability(theseus_rhs *h, const double *u, theseus_stability *out, void *stream) {
  if (!h || !u || !out) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null handle, vector or output");
  hipStream_t s = (hipStream_t)stream;
  const double *ua = u;
  const int rc = aligned_input(h, u, &ua, s);
  if (rc) return rc;
  enqueue_stability(h, ua, s);
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

int theseus_orient_node(int dim, int order, int code, int node) {
  if ((dim != 2 && dim != 3) || order < 1) return fail(THESEUS_ERR_INVALID_ARGUMENT, "dim must be 2 or 3, order >= 1");
  const std::vector<Sym> syms = symmetries(dim);
  const int Np = order + 1, ND = dim == 2 ? Np * Np : Np * Np * Np;
  if (code < 0 || code >= (int)syms.size() || node < 0 || node >= ND)
    return fail(THESEUS_ERR_INVALID_ARGUMENT, "code or node out of range");
  return sym_node(syms[code], dim, Np, node);
}

// Breadth-first alignment of the elements (a restatement for the C-ABI of
// the case reader's orientation walk): the first element of each connected
// component keeps its frame; a neighbour across face (d, s) takes the
// symmetry under which its face (d, 1-s) lists the shared corners in the
// order the aligned element lists them.
int theseus_orient_elements(int dim, int ne, const long long *vids, const double *coords, unsigned char *orient_out,
                            int *nbr_out, double *jac_out, double *adj_out) {
  if ((dim != 2 && dim != 3) || ne < 1 || !vids)
    return fail(THESEUS_ERR_INVALID_ARGUMENT, "dim must be 2 or 3, num_elements >= 1, elem_vertices given");
  const int nc = 1 << dim, nfc = nc / 2;
  const std::vector<Sym> syms = symmetries(dim);
  std::vector<std::array<int, 8>> cmap(syms.size());
  for (size_t c = 0; c < syms.size(); ++c)
    for (int k = 0; k < nc; ++k) cmap[c][k] = sym_corner(syms[c], dim, k);
  int faces[3][2][4];
  for (int d = 0; d < dim; ++d)
    for (int sd = 0; sd < 2; ++sd)
      for (int c = 0, m = 0; c < nc; ++c)
        if (((c >> d) & 1) == sd) faces[d][sd][m++] = c;
  auto face_key = [&](const long long *v) {
    std::vector<long long> k(v, v + nfc);
    std::sort(k.begin(), k.end());
    return k;
  };
  std::map<std::vector<long long>, std::vector<int>> owners;
  for (int e = 0; e < ne; ++e)
    for (int d = 0; d < dim; ++d)
      for (int sd = 0; sd < 2; ++sd) {
        long long fv[4];
        for (int m = 0; m < nfc; ++m) fv[m] = vids[(size_t)e * nc + faces[d][sd][m]];
        std::vector<int> &o = owners[face_key(fv)];
        o.push_back(e);
        if (o.size() > 2) return fail(THESEUS_ERR_UNSUPPORTED, "a face is shared by more than two elements");
      }
  std::vector<int> code(ne, -1), nbr((size_t)ne * 2 * dim, -1), comp(ne, -1);
  auto oriented = [&](int e, long long *ov) {
    for (int c = 0; c < nc; ++c) ov[c] = vids[(size_t)e * nc + cmap[code[e]][c]];
  };
  for (int start = 0; start < ne; ++start) {
    if (code[start] >= 0) continue;
    code[start] = 0;
    std::vector<int> stack{start};
    while (!stack.empty()) {
      const int e = stack.back();
      stack.pop_back();
      comp[e] = start;
      long long ov[8];
      oriented(e, ov);
      for (int d = 0; d < dim; ++d)
        for (int sd = 0; sd < 2; ++sd) {
          long long fv[4];
          for (int m = 0; m < nfc; ++m) fv[m] = ov[faces[d][sd][m]];
          const std::vector<int> &o = owners[face_key(fv)];
          int f = -1;
          for (int x : o)
            if (x != e) f = x;
          if (f < 0) {
            if (o.size() == 2) f = e;  // an element touching itself across a periodic face
            else continue;
          }
          if (code[f] >= 0) {
            long long of[8];
            oriented(f, of);
            for (int m = 0; m < nfc; ++m)
              if (of[faces[d][1 - sd][m]] != fv[m])
                return fail(THESEUS_ERR_UNSUPPORTED, "elements " + std::to_string(e) + " and " + std::to_string(f) +
                                                         " cannot be aligned (the mesh has unstructured vertices)");
          } else {
            int found = -1;
            for (size_t c = 0; c < syms.size() && found < 0; ++c) {
              bool ok = true;
              for (int m = 0; m < nfc && ok; ++m) ok = vids[(size_t)f * nc + cmap[c][faces[d][1 - sd][m]]] == fv[m];
              if (ok) found = (int)c;
            }
            if (found < 0)
              return fail(THESEUS_ERR_UNSUPPORTED, "no orientation of element " + std::to_string(f) +
                                                       " matches its face with element " + std::to_string(e));
            code[f] = found;
            stack.push_back(f);
          }
          nbr[(size_t)e * 2 * dim + 2 * d + sd] = f;
        }
    }
  }
  if (coords) {
    // corners in the aligned frame; a component aligned from a mirrored
    // first element has negative Jacobians throughout: reversing aligned
    // axis 0 in all of it keeps every face pairing and makes them positive
    auto edges = [&](int e, double E[3][3], double X0[3]) {
      const double *c0 = coords + ((size_t)e * nc + cmap[code[e]][0]) * dim;
      for (int k = 0; k < dim; ++k) X0[k] = c0[k];
      for (int r = 0; r < dim; ++r) {
        const double *cr = coords + ((size_t)e * nc + cmap[code[e]][1 << r]) * dim;
        for (int k = 0; k < dim; ++k) E[r][k] = cr[k] - c0[k];
      }
    };
    auto det = [&](double E[3][3]) {
      return dim == 2 ? E[0][0] * E[1][1] - E[0][1] * E[1][0]
                      : E[0][0] * (E[1][1] * E[2][2] - E[1][2] * E[2][1]) -
                            E[0][1] * (E[1][0] * E[2][2] - E[1][2] * E[2][0]) +
                            E[0][2] * (E[1][0] * E[2][1] - E[1][1] * E[2][0]);
    };
    std::map<int, std::pair<int, int>> signs;  // component -> (negative, positive) counts
    std::vector<double> dets(ne);
    for (int e = 0; e < ne; ++e) {
      double E[3][3], X0[3];
      edges(e, E, X0);
      dets[e] = det(E);
      auto &sg = signs[comp[e]];
      (dets[e] < 0 ? sg.first : sg.second)++;
    }
    // symmetry code of "code, then aligned axis 0 reversed"
    auto flip0 = [&](int c) {
      for (size_t k = 0; k < syms.size(); ++k) {
        bool ok = true;
        for (int m = 0; m < nc && ok; ++m) ok = cmap[k][m] == cmap[c][m ^ 1];
        if (ok) return (int)k;
      }
      return -1;
    };
    for (int e = 0; e < ne; ++e) {
      const auto &sg = signs[comp[e]];
      if (sg.first && sg.second)
        return fail(THESEUS_ERR_UNSUPPORTED, "element orientations disagree with the geometry (mixed Jacobian signs)");
      if (sg.first) {
        code[e] = flip0(code[e]);
        std::swap(nbr[(size_t)e * 2 * dim], nbr[(size_t)e * 2 * dim + 1]);
      }
    }
    for (int e = 0; e < ne; ++e) {
      double E[3][3], X0[3];
      edges(e, E, X0);
      double h = 0;
      for (int r = 0; r < dim; ++r) {
        double l2 = 0;
        for (int k = 0; k < dim; ++k) l2 += E[r][k] * E[r][k];
        h = std::max(h, std::sqrt(l2));
      }
      for (int c = 0; c < nc; ++c) {  // affinity: each corner is the origin plus its edges
        const double *xc = coords + ((size_t)e * nc + cmap[code[e]][c]) * dim;
        double dev = 0;
        for (int k = 0; k < dim; ++k) {
          double pred = X0[k];
          for (int r = 0; r < dim; ++r) pred += ((c >> r) & 1) * E[r][k];
          dev += (xc[k] - pred) * (xc[k] - pred);
        }
        if (std::sqrt(dev) > 1e-9 * h)
          return fail(THESEUS_ERR_UNSUPPORTED, "element " + std::to_string(e) + " is not affine");
      }
      const double J = det(E);
      if (!(J > 0)) return fail(THESEUS_ERR_UNSUPPORTED, "element " + std::to_string(e) + " has a non-positive Jacobian");
      if (jac_out) jac_out[e] = J;
      if (adj_out) {
        // adjugate of A (columns = edges): row r = J d(xi_r)/dx
        double *o = adj_out + (size_t)e * dim * dim;
        if (dim == 2) {
          o[0] = E[1][1];
          o[1] = -E[1][0];
          o[2] = -E[0][1];
          o[3] = E[0][0];
        } else {
          for (int r = 0; r < 3; ++r) {
            const int a = (r + 1) % 3, b = (r + 2) % 3;
            o[r * 3 + 0] = E[a][1] * E[b][2] - E[a][2] * E[b][1];
            o[r * 3 + 1] = E[a][2] * E[b][0] - E[a][0] * E[b][2];
            o[r * 3 + 2] = E[a][0] * E[b][1] - E[a][1] * E[b][0];
          }
        }
      }
    }
  }
  if (orient_out)
    for (int e = 0; e < ne; ++e) orient_out[e] = (unsigned char)code[e];
  if (nbr_out) std::copy(nbr.begin(), nbr.end(), nbr_out);
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

// ---------------------------------------------------------------------------
// Serving pool: many cases in one element array (see theseus_rhs.h).
// ---------------------------------------------------------------------------
namespace {

struct PoolStaging {
  char *host = nullptr;  // pinned
  size_t bytes = 0;
  hipEvent_t done = nullptr;  // recorded after the copies out of it
};

struct PoolCase {
  int start = -1, ne = 0;
  bool active = false;
  bool boundary = false;  // the case's mesh has boundary faces
};

// Process-wide memory of the autotuner: (dim, order, vol, surf, precision) -> (tile, xcd_remap)
std::map<std::tuple<int, int, int, int, int>, std::pair<int, int>> &tune_cache() {
  static std::map<std::tuple<int, int, int, int, int>, std::pair<int, int>> c;
  return c;
}
std::mutex &tune_mutex() {
  static std::mutex m;
  return m;
}

}  // namespace

struct theseus_pool {
  theseus_pool_desc d;
  theseus_case_desc cd;  // the pool's physics as a case descriptor (no arrays)
  int Np = 0, ndof = 0, nfp = 0, neq = 0, cap = 0;
  long long Ntot = 0;
  const FluxKernels *fk = nullptr;
  const ShapeKernels *sk = nullptr;
  const TileKernels *tk = nullptr;
  BasisArgs basis;
  GasC<double> gas;
  GasC<float> gas32;
  int pm = 0, tile = 0, xcd = 1;
  int active = 0;  // slots a step launches over: the occupied slots' high-water mark, rounded up
  int bnd = 0;     // 1 while an admitted case has boundary faces: the kernels with boundary conditions
  double adv_scale = 0, diff_scale = 0, surf_scale = 0;
  double *d_J = nullptr, *d_adj = nullptr, *d_alpha = nullptr, *d_Fv = nullptr, *d_U = nullptr;
  double *d_stage[2] = {nullptr, nullptr};
  int *d_nbr = nullptr, *d_elem_case = nullptr;
  // boundary table: case c's boundary b is entry c * THESEUS_POOL_MAX_BC + b
  int *d_bc_type = nullptr;
  double *d_bc_data = nullptr;
  double *d_case_dt = nullptr, *d_case_pref = nullptr;
  unsigned long long *d_rates = nullptr;
  double *d_meas = nullptr;  // integral measures: partials of the largest case, then the results
  std::vector<PoolCase> cases;
  std::vector<unsigned char> slot_used;  // per element slot
  PoolStaging staging[2];
  int next_staging = 0;
  hipStream_t capture_stream = nullptr;
  // captured SSP-RK3 steps, one per (launch range, kernel choice): a step
  // replays the one for the current range, capturing it on first use; none
  // is destroyed before the pool, so a change of range needs no
  // synchronisation with launches still in flight
  std::map<std::pair<int, int>, std::pair<hipGraph_t, hipGraphExec_t>> graphs;

  int epb() const { return (tk && cd.flow_model == THESEUS_FLOW_EULER) ? tk->epb[tile] : sk->epb; }
  int nblocks() const { return (active + epb() - 1) / epb(); }

  RHSArgs args(const double *U, double *out) const {
    RHSArgs a;
    std::memset(&a, 0, sizeof(a));
    a.U = U;
    a.out = out;
    if (d_Fv) {
      const size_t esz = pm ? sizeof(float) : sizeof(double);
      char *base = reinterpret_cast<char *>(d_Fv);
      a.Vd = a.Vd_out = base;
      a.Gf = a.Gf_out = base + esz * (size_t)(cd.dim + 1) * Ntot;
    }
    a.alpha = d_alpha;
    a.Ntot = Ntot;
    a.ne = active;  // every neighbour index of an occupied slot is below the high-water mark
    a.blending = cd.blending;
    a.alpha_max = cd.alpha_max;
    a.alpha_min = cd.alpha_min;
    a.sharp = 9.21024;
    a.thresh = 0.5 * std::pow(10.0, -1.8 * std::pow(cd.order, 0.25));
    a.J = d_J;
    a.adj = d_adj;
    a.nbr = d_nbr;
    a.bc_type = d_bc_type;
    a.bc_data = d_bc_data;
    a.gas = gas;
    a.gas32 = gas32;
    a.elem_case = d_elem_case;
    a.case_dt = d_case_dt;
    a.case_pref = d_case_pref;
    a.xcd_remap = xcd;
    a.basis = basis;
    return a;
  }

  void launch(const RHSArgs &a, hipStream_t s) const {
    if (cd.flow_model == THESEUS_FLOW_NAVIER_STOKES) {
      sk->ns_grad[pm][bnd](a, nblocks(), s);
      fk->ns[pm][bnd](a, nblocks(), s);
    } else if (tk) {
      tk->euler[tile][pm][bnd](a, nblocks(), s);
    } else {
      fk->euler[pm][bnd](a, nblocks(), s);
    }
  }

  // stage st of the SSP-RK3 step of every case (MFEM RK3SSPSolver,
  // src/Simulation.cpp:265-268): u -> A, (u, A) -> B, (u, B) -> u
  RHSArgs stage_args(int st) const {
    double *u = d_U, *A = d_stage[0], *B = d_stage[1];
    RHSArgs a = args(st == 0 ? u : (st == 1 ? A : B), st == 0 ? A : (st == 1 ? B : u));
    a.mode = st == 0 ? 1 : 2;
    if (st > 0) {
      a.X = u;
      a.cx = st == 1 ? 3. / 4 : 1. / 3;
      a.cy = st == 1 ? 1. / 4 : 2. / 3;
    }
    return a;
  }

  void enqueue_step(hipStream_t s) const {
    for (int st = 0; st < 3; ++st) launch(stage_args(st), s);
  }
};

// Pools of orders 2, 3 and 4 stepped by one launch per stage (k_rhs_orders):
// a captured step per combination of the pools' launch ranges.
struct GroupGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static size_t pool_bytes_per_element(const theseus_pool_desc &d, int ndof, int nfp) {
  const int neq = d.dim + 2;
  size_t b = 3 * (size_t)neq * ndof * sizeof(double);              // state + two stages
  b += sizeof(double) * (1 + d.dim * d.dim) + sizeof(int) * (2 * d.dim + 1) + sizeof(double);  // J, adj, nbr, owner, alpha
  if (d.flow_model == THESEUS_FLOW_NAVIER_STOKES)
    b += (size_t)(d.dim + 1) * (ndof + 2 * d.dim * nfp) * sizeof(double);
  return b;
}

// Free slots point at themselves, so the kernels' unconditional neighbour
// loads stay inside the arrays; their owner is -1 (skipped).
__global__ void __launch_bounds__(256) k_pool_reset(int *nbr, int *owner, double *J, double *adj, int start, int n,
                                                    int dim) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int e = start + i;
  for (int f = 0; f < 2 * dim; ++f) nbr[(size_t)e * 2 * dim + f] = e;
  owner[e] = -1;
  J[e] = 1.0;
  for (int r = 0; r < dim; ++r)
    for (int c = 0; c < dim; ++c) adj[(size_t)e * dim * dim + r * dim + c] = r == c ? 1.0 : 0.0;
}

__global__ void __launch_bounds__(256) k_fill_i32(int *p, int n, int v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

static int pool_reset_slots(theseus_pool *p, int start, int n, hipStream_t s) {
  if (n <= 0) return THESEUS_OK;
  hipLaunchKernelGGL(k_pool_reset, dim3((n + 255) / 256), dim3(256), 0, s, p->d_nbr, p->d_elem_case, p->d_J, p->d_adj,
                     start, n, p->cd.dim);
  HIP_TRY(hipGetLastError());
  return THESEUS_OK;
}

// Pinned staging for the host tables of an admission: copies from it are
// asynchronous, where copies from pageable memory make the host wait for the
// stream.  Two buffers alternate; an admission waits only for the copies of
// the admission before the last one (long done in a serving loop).  A buffer
// grows when a larger case arrives; freeing the old one synchronises, once.
static int pool_stage(theseus_pool *p, size_t bytes, char **host, hipEvent_t *done) {
  PoolStaging &b = p->staging[p->next_staging];
  p->next_staging ^= 1;
  if (!b.done) HIP_TRY(hipEventCreateWithFlags(&b.done, hipEventDisableTiming));
  else HIP_TRY(hipEventSynchronize(b.done));
  if (bytes > b.bytes) {
    if (b.host) HIP_TRY(hipHostFree(b.host));
    b.host = nullptr;
    b.bytes = 0;
    const size_t want = std::max(bytes, (size_t)1 << 20);
    HIP_TRY(hipHostMalloc((void **)&b.host, want, hipHostMallocDefault));
    b.bytes = want;
  }
  *host = b.host;
  *done = b.done;
  return THESEUS_OK;
}

// The step's launch range follows the occupied slots: their high-water mark
// rounded up to the next of a geometric ladder of ranges (1024 slots, then
// steps of 25 %), so that first-fit admissions and releases move between a
// few ranges whose captured steps are kept; the kernels with the boundary
// conditions run while an admitted case has boundary faces.
static int pool_update_active(theseus_pool *p) {
  int hwm = 0, bnd = 0;
  for (const PoolCase &c : p->cases)
    if (c.active) {
      hwm = std::max(hwm, c.start + c.ne);
      bnd |= c.boundary ? 1 : 0;
    }
  const int gran = 1024;
  int next = 0;
  if (hwm > 0) {
    next = gran;
    while (next < hwm) next = (int)std::min<long long>(p->cap, ((long long)next * 5 / 4 + gran - 1) / gran * gran);
    next = std::min(next, p->cap);
  }
  p->active = next;
  p->bnd = bnd;
  return THESEUS_OK;
}

static int pool_find_range(const theseus_pool *p, int n) {
  int run = 0;
  for (int e = 0; e < p->cap; ++e) {
    run = p->slot_used[e] ? 0 : run + 1;
    if (run == n) return e - n + 1;
  }
  return -1;
}

// Times the candidate tile shapes on a synthetic periodic box that fills (up
// to 32^3 elements of) the pool: a perturbed uniform flow, 3 warmup + 8
// timed RHS evaluations each; keeps the fastest.
static int pool_autotune(theseus_pool *p, hipStream_t s) {
  const int dim = p->cd.dim;
  int n = 1;
  while (n < 32 && (long long)(n + 1) * (n + 1) * (dim == 3 ? (n + 1) : 1) <= p->cap) ++n;
  const int ne = dim == 3 ? n * n * n : n * n;
  if (ne < 8) return THESEUS_OK;
  std::vector<double> J(ne), adj((size_t)ne * dim * dim, 0.0);
  std::vector<int> nbr((size_t)ne * 2 * dim), owner(ne, 0);
  const double h = 1.0 / n;
  for (int e = 0; e < ne; ++e) {
    const int ix = e % n, iy = (e / n) % n, iz = dim == 3 ? e / (n * n) : 0;
    const int c[3] = {ix, iy, iz};
    J[e] = dim == 3 ? h * h * h : h * h;
    for (int dd = 0; dd < dim; ++dd) adj[(size_t)e * dim * dim + dd * dim + dd] = dim == 3 ? h * h : h;
    for (int dd = 0; dd < dim; ++dd)
      for (int sd = 0; sd < 2; ++sd) {
        int cc[3] = {c[0], c[1], c[2]};
        cc[dd] = (cc[dd] + (sd ? 1 : n - 1)) % n;
        nbr[(size_t)e * 2 * dim + 2 * dd + sd] = cc[0] + n * cc[1] + n * n * cc[2];
      }
  }
  const long long nd = (long long)ne * p->ndof;
  std::vector<double> U((size_t)p->neq * p->Ntot, 0.0);
  for (long long i = 0; i < nd; ++i) {
    const double w = 0.05 * std::sin(0.37 * (double)i);
    const double rho = 1.0 + w, u0 = 0.3 + w, p0 = 1.0 / p->gas.gamma;
    U[0 * p->Ntot + i] = rho;
    U[1 * p->Ntot + i] = rho * u0;
    U[(size_t)(p->neq - 1) * p->Ntot + i] = p0 / p->gas.gammaM1 + 0.5 * rho * u0 * u0;
  }
  const double dt1 = 0.0, pref = 1.0 / p->gas.gamma;
  HIP_TRY(hipMemcpyAsync(p->d_J, J.data(), J.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(p->d_adj, adj.data(), adj.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(p->d_nbr, nbr.data(), nbr.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(p->d_elem_case, owner.data(), owner.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(p->d_U, U.data(), U.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(p->d_case_dt, &dt1, sizeof dt1, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(p->d_case_pref, &pref, sizeof pref, hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));

  const bool tune_tile = p->d.tile < 0 && p->tk && p->cd.flow_model == THESEUS_FLOW_EULER;
  const bool tune_xcd = p->d.xcd_remap < 0;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  float best = 1e30f;
  int best_tile = p->tile, best_xcd = p->xcd;
  for (int t = 0; t < 2; ++t) {
    if (!tune_tile && t != p->tile) continue;
    for (int x = 0; x < 2; ++x) {
      if (!tune_xcd && x != p->xcd) continue;
      p->tile = t;
      p->xcd = x;
      p->active = ne;
      RHSArgs a = p->args(p->d_U, p->d_stage[0]);
      for (int it = 0; it < 3; ++it) p->launch(a, s);
      hipEventRecord(e0, s);
      for (int it = 0; it < 8; ++it) p->launch(a, s);
      hipEventRecord(e1, s);
      HIP_TRY(hipEventSynchronize(e1));
      float ms = 0;
      hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) {
        best = ms;
        best_tile = t;
        best_xcd = x;
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  HIP_TRY(hipGetLastError());
  p->tile = best_tile;
  p->xcd = best_xcd;
  p->active = 0;
  return pool_reset_slots(p, 0, ne, s);
}

extern "C" {

int theseus_pool_create(const theseus_pool_desc *desc, theseus_pool **out) {
  if (!desc || !out) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null descriptor or output");
  *out = nullptr;
  const theseus_pool_desc &d = *desc;
  if (d.dim != 2 && d.dim != 3) return fail(THESEUS_ERR_INVALID_ARGUMENT, "dim must be 2 or 3");
  if (d.order < 1 || d.order > 7) return fail(THESEUS_ERR_INVALID_ARGUMENT, "order must be in [1,7]");
  if (d.flow_model != THESEUS_FLOW_EULER && d.flow_model != THESEUS_FLOW_NAVIER_STOKES)
    return fail(THESEUS_ERR_INVALID_ARGUMENT, "unknown flow_model");
  if (d.precision != THESEUS_PRECISION_FP64 && d.precision != THESEUS_PRECISION_MIXED)
    return fail(THESEUS_ERR_UNSUPPORTED, "precision must be THESEUS_PRECISION_FP64 or THESEUS_PRECISION_MIXED");
  if (!(d.gamma > 1.0) || !std::isfinite(d.gamma)) return fail(THESEUS_ERR_INVALID_ARGUMENT, "gamma must be > 1");
  if (d.max_cases < 1) return fail(THESEUS_ERR_INVALID_ARGUMENT, "max_cases must be positive");
  if (d.tile < -1 || d.tile > 1 || d.xcd_remap < -1 || d.xcd_remap > 1)
    return fail(THESEUS_ERR_INVALID_ARGUMENT, "tile and xcd_remap must be -1, 0 or 1");
  const FluxKernels *fk = find_flux(d.dim, d.order, d.volume_flux, d.surface_flux);
  const ShapeKernels *sk = find_shape(d.dim, d.order);
  if (!fk || !sk) return fail(THESEUS_ERR_UNSUPPORTED, "no kernel compiled for this dim/order/flux pairing");
  const TileKernels *tk = find_tiles(d.dim, d.order, d.volume_flux, d.surface_flux);
  if (d.tile == 1 && (!tk || d.flow_model != THESEUS_FLOW_EULER))
    return fail(THESEUS_ERR_UNSUPPORTED, "tile 1 is compiled only for 3-D Euler, Ranocha + Lax-Friedrichs, p = 2..4");

  theseus_pool *p = new theseus_pool();
  p->d = d;
  std::memset(&p->cd, 0, sizeof(p->cd));
  p->cd.dim = d.dim;
  p->cd.order = d.order;
  p->cd.flow_model = d.flow_model;
  p->cd.volume_flux = d.volume_flux;
  p->cd.surface_flux = d.surface_flux;
  p->cd.precision = d.precision;
  p->cd.blending = d.blending;
  p->cd.gamma = d.gamma;
  p->cd.Pr = d.Pr;
  p->cd.R_gas = d.R_gas;
  p->cd.mu = d.mu;
  p->cd.mu_bulk = d.mu_bulk;
  p->cd.alpha_max = d.alpha_max;
  p->cd.alpha_min = d.alpha_min;
  p->fk = fk;
  p->sk = sk;
  p->tk = d.flow_model == THESEUS_FLOW_EULER ? tk : nullptr;
  p->Np = d.order + 1;
  p->ndof = d.dim == 2 ? p->Np * p->Np : p->Np * p->Np * p->Np;
  p->nfp = d.dim == 2 ? p->Np : p->Np * p->Np;
  p->neq = d.dim + 2;
  p->pm = d.precision == THESEUS_PRECISION_MIXED ? 1 : 0;
  p->tile = d.tile == 1 ? 1 : 0;
  p->xcd = d.xcd_remap == 0 ? 0 : 1;
  // capacity from the budget, a multiple of 16 elements (block counts divisible by 8 for either tile)
  const size_t per = pool_bytes_per_element(d, p->ndof, p->nfp);
  const long long cap = d.hbm_budget_bytes > 0 ? (long long)(d.hbm_budget_bytes / (long long)per) : 0;
  if (cap < 16) {
    delete p;
    return fail(THESEUS_ERR_CAPACITY, "hbm_budget_bytes holds fewer than 16 elements");
  }
  if (cap * (long long)p->ndof > (1LL << 31) / 2) {
    delete p;
    return fail(THESEUS_ERR_INVALID_ARGUMENT, "hbm_budget_bytes exceeds the pool's 32-bit element indexing");
  }
  p->cap = (int)(cap / 16 * 16);
  p->Ntot = (long long)p->cap * p->ndof;
  std::vector<double> x, w;
  gll01(d.order, x, w);
  derivative_ops(d.order, x, w, p->basis);
  legendre_vandermonde_inverse(d.order, x, p->basis.Vinv);
  init_gas(p->cd, p->gas, p->gas32);
  p->adv_scale = advection_scale(d.order);
  p->diff_scale = diffusion_scale(d.order);
  p->surf_scale = p->adv_scale / (1.0 / w[0]);
  p->cases.resize(d.max_cases);
  p->slot_used.assign(p->cap, 0);

  int rc = THESEUS_OK;
  const size_t M = (size_t)p->Ntot * p->neq;
  const int dim = d.dim;
  if ((rc = upload<double>(&p->d_J, nullptr, p->cap)) || (rc = upload<double>(&p->d_adj, nullptr, (size_t)p->cap * dim * dim)) ||
      (rc = upload<int>(&p->d_nbr, nullptr, (size_t)p->cap * 2 * dim)) ||
      (rc = upload<int>(&p->d_elem_case, nullptr, p->cap)) || (rc = upload<double>(&p->d_alpha, nullptr, p->cap)) ||
      (rc = upload<int>(&p->d_bc_type, nullptr, (size_t)d.max_cases * THESEUS_POOL_MAX_BC)) ||
      (rc = upload<double>(&p->d_bc_data, nullptr, (size_t)d.max_cases * THESEUS_POOL_MAX_BC * THESEUS_BC_DATA_STRIDE)) ||
      (rc = upload<double>(&p->d_case_dt, nullptr, d.max_cases)) ||
      (rc = upload<double>(&p->d_case_pref, nullptr, d.max_cases)) ||
      (rc = upload<unsigned long long>(&p->d_rates, nullptr, kRateWords)) || (rc = upload<double>(&p->d_U, nullptr, M)) ||
      (rc = upload<double>(&p->d_stage[0], nullptr, M)) || (rc = upload<double>(&p->d_stage[1], nullptr, M))) {
    theseus_pool_destroy(p);
    return rc;
  }
  if (d.flow_model == THESEUS_FLOW_NAVIER_STOKES &&
      (rc = upload<double>(&p->d_Fv, nullptr, (size_t)(dim + 1) * ((size_t)p->Ntot + (size_t)p->cap * 2 * dim * p->nfp)))) {
    theseus_pool_destroy(p);
    return rc;
  }
  hipError_t e = hipStreamCreate(&p->capture_stream);  // default priority
  if (e != hipSuccess) {
    theseus_pool_destroy(p);
    return fail(THESEUS_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  hipStream_t s = p->capture_stream;
  if ((rc = pool_reset_slots(p, 0, p->cap, s))) {
    theseus_pool_destroy(p);
    return rc;
  }
  // tile shape: remembered per (dim, order, pairing, precision), else timed now
  if (d.tile < 0 || d.xcd_remap < 0) {
    const auto key = std::make_tuple(d.dim, d.order, d.volume_flux, d.surface_flux, d.precision * 2 + (d.flow_model == THESEUS_FLOW_NAVIER_STOKES));
    std::lock_guard<std::mutex> lk(tune_mutex());
    auto it = tune_cache().find(key);
    if (it != tune_cache().end()) {
      if (d.tile < 0) p->tile = it->second.first;
      if (d.xcd_remap < 0) p->xcd = it->second.second;
    } else {
      if ((rc = pool_autotune(p, s))) {
        theseus_pool_destroy(p);
        return rc;
      }
      tune_cache()[key] = {p->tile, p->xcd};
    }
  }
  // the slot reset runs on the pool's own stream: finish it before any call on another
  {
    const hipError_t se = hipStreamSynchronize(s);
    if (se != hipSuccess) {
      theseus_pool_destroy(p);
      return fail(THESEUS_ERR_DEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(se));
    }
  }
  *out = p;
  return THESEUS_OK;
}

int theseus_pool_destroy(theseus_pool *p) {
  if (!p) return THESEUS_OK;
  for (auto &kv : p->graphs) {
    hipGraphExecDestroy(kv.second.second);
    hipGraphDestroy(kv.second.first);
  }
  if (p->capture_stream) hipStreamDestroy(p->capture_stream);
  for (PoolStaging &b : p->staging) {
    if (b.done) {
      hipEventSynchronize(b.done);
      hipEventDestroy(b.done);
    }
    if (b.host) hipHostFree(b.host);
  }
  hipFree(p->d_J);
  hipFree(p->d_adj);
  hipFree(p->d_nbr);
  hipFree(p->d_bc_type);
  hipFree(p->d_bc_data);
  hipFree(p->d_elem_case);
  hipFree(p->d_alpha);
  hipFree(p->d_case_dt);
  hipFree(p->d_case_pref);
  hipFree(p->d_rates);
  hipFree(p->d_meas);
  hipFree(p->d_U);
  hipFree(p->d_stage[0]);
  hipFree(p->d_stage[1]);
  hipFree(p->d_Fv);
  delete p;
  return THESEUS_OK;
}

int theseus_pool_info(const theseus_pool *p, int *capacity_elements, int *free_elements, int *largest_free_range,
                      int *active_cases, int *tile, int *xcd_remap) {
  if (!p) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null pool");
  int fr = 0, run = 0, best = 0, act = 0;
  for (int e = 0; e < p->cap; ++e) {
    if (p->slot_used[e]) {
      run = 0;
    } else {
      ++fr;
      best = std::max(best, ++run);
    }
  }
  for (const PoolCase &c : p->cases) act += c.active;
  if (capacity_elements) *capacity_elements = p->cap;
  if (free_elements) *free_elements = fr;
  if (largest_free_range) *largest_free_range = best;
  if (active_cases) *active_cases = act;
  if (tile) *tile = p->tile;
  if (xcd_remap) *xcd_remap = p->xcd;
  return THESEUS_OK;
}

int theseus_pool_admit(theseus_pool *p, const theseus_case_desc *c, const double *u0, double cfl, int *case_id,
                       double *dt_out, void *stream) {
  if (c && c->elem_orient)
    for (int e = 0; e < c->num_elements; ++e)
      if (c->elem_orient[e] != 0)
        return fail(THESEUS_ERR_UNSUPPORTED, "the serving pool takes aligned meshes (elem_orient all zero or NULL)");
  if (!p || !c || !u0) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null pool, case or state");
  if (!(cfl > 0) || !std::isfinite(cfl)) return fail(THESEUS_ERR_INVALID_ARGUMENT, "cfl must be positive");
  const theseus_case_desc &d = *c;
  if (d.dim != p->cd.dim || d.order != p->cd.order || d.flow_model != p->cd.flow_model ||
      d.volume_flux != p->cd.volume_flux || d.surface_flux != p->cd.surface_flux || d.precision != p->cd.precision ||
      (d.blending != 0) != (p->cd.blending != 0))
    return fail(THESEUS_ERR_INVALID_ARGUMENT, "case kernel fields (dim, order, flow, fluxes, precision, blending) differ from the pool's");
  if (d.num_elements < 1 || !d.elem_jac || !d.elem_adj || !d.face_neighbor)
    return fail(THESEUS_ERR_INVALID_ARGUMENT, "case needs elements, geometry and connectivity");
  bool has_boundary = false;
  int rc = validate_mesh(d, has_boundary);
  if (rc) return rc;
  if (d.num_bc > THESEUS_POOL_MAX_BC)
    return fail(THESEUS_ERR_INVALID_ARGUMENT, "a pooled case has at most " + std::to_string(THESEUS_POOL_MAX_BC) + " boundaries");
  int cid = -1;
  for (int i = 0; i < (int)p->cases.size(); ++i)
    if (!p->cases[i].active) {
      cid = i;
      break;
    }
  if (cid < 0) return fail(THESEUS_ERR_CAPACITY, "all case slots are in use");
  const int ne = d.num_elements, dim = d.dim;
  const int start = pool_find_range(p, ne);
  if (start < 0) return fail(THESEUS_ERR_CAPACITY, "no free range of " + std::to_string(ne) + " elements");
  hipStream_t s = (hipStream_t)stream;
  // geometry and connectivity at the slot offset, staged in pinned memory;
  // boundary codes point into the case's rows of the pool's boundary table
  const size_t nJ = ne, nA = (size_t)ne * dim * dim, nB = (size_t)d.num_bc * THESEUS_BC_DATA_STRIDE;
  const size_t nN = (size_t)ne * 2 * dim, nT = d.num_bc;
  char *host = nullptr;
  hipEvent_t staged = nullptr;
  if ((rc = pool_stage(p, (nJ + nA + nB) * sizeof(double) + (nN + nT) * sizeof(int), &host, &staged))) return rc;
  double *hJ = reinterpret_cast<double *>(host), *hA = hJ + nJ, *hB = hA + nA;
  int *hN = reinterpret_cast<int *>(hB + nB), *hT = hN + nN;
  std::copy(d.elem_jac, d.elem_jac + nJ, hJ);
  std::copy(d.elem_adj, d.elem_adj + nA, hA);
  std::fill(hB, hB + nB, 0.0);
  if (d.bc_data) std::copy(d.bc_data, d.bc_data + nB, hB);
  if (nT) std::copy(d.bc_type, d.bc_type + nT, hT);
  const int boff = cid * THESEUS_POOL_MAX_BC;
  for (size_t f = 0; f < nN; ++f) {
    const int v = d.face_neighbor[f];
    hN[f] = v >= 0 ? v + start : -(boff - (v + 1) + 1);
  }
  if (d.num_bc > 0) {
    HIP_TRY(hipMemcpyAsync(p->d_bc_type + boff, hT, nT * sizeof(int), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(p->d_bc_data + (size_t)boff * THESEUS_BC_DATA_STRIDE, hB, nB * sizeof(double),
                           hipMemcpyHostToDevice, s));
  }
  HIP_TRY(hipMemcpyAsync(p->d_J + start, hJ, nJ * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(p->d_adj + (size_t)start * dim * dim, hA, nA * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(p->d_nbr + (size_t)start * 2 * dim, hN, nN * sizeof(int), hipMemcpyHostToDevice, s));
  HIP_TRY(hipEventRecord(staged, s));
  const long long nd = (long long)ne * p->ndof, off = (long long)start * p->ndof;
  for (int q = 0; q < p->neq; ++q)
    HIP_TRY(hipMemcpyAsync(p->d_U + q * p->Ntot + off, u0 + q * nd, nd * sizeof(double), hipMemcpyDefault, s));
  // stability estimate of the case alone (its element range), then dt and the background pressure
  hipLaunchKernelGGL(k_zero_u64, dim3((kRateWords + 63) / 64), dim3(64), 0, s, p->d_rates, kRateWords);
  StabArgs a;
  std::memset(&a, 0, sizeof(a));
  a.U = p->d_U;
  a.rates = p->d_rates;
  a.Ntot = p->Ntot;
  a.e0 = start;
  a.ne = ne;
  a.viscous = d.flow_model == THESEUS_FLOW_NAVIER_STOKES;
  a.adv_scale = p->adv_scale;
  a.diff_scale = p->diff_scale;
  a.surf_scale = p->surf_scale;
  a.J = p->d_J;
  a.adj = p->d_adj;
  a.nbr = p->d_nbr;
  a.bc_type = p->d_bc_type;
  a.bc_data = p->d_bc_data;
  a.gas = p->gas;
  a.w0 = p->basis.w[0];
  p->sk->stability(a, (ne + p->sk->epb - 1) / p->sk->epb, s);
  hipLaunchKernelGGL(k_rates_fold, dim3(1), dim3(64), 0, s, p->d_rates);
  if (dim == 2) hipLaunchKernelGGL(k_pref<2>, dim3(1), dim3(64), 0, s, p->d_U + off, p->Ntot, p->gas, p->d_case_pref + cid);
  else hipLaunchKernelGGL(k_pref<3>, dim3(1), dim3(64), 0, s, p->d_U + off, p->Ntot, p->gas, p->d_case_pref + cid);
  hipLaunchKernelGGL(k_case_dt, dim3(1), dim3(64), 0, s, p->d_rates, cfl, p->d_case_dt + cid);
  hipLaunchKernelGGL(k_fill_i32, dim3((ne + 255) / 256), dim3(256), 0, s, p->d_elem_case + start, ne, cid);
  HIP_TRY(hipGetLastError());
  double dt = 0.0;
  if (dt_out) {
    // the caller asked for dt: wait for it, and refuse a case whose estimate is not usable
    HIP_TRY(hipMemcpyAsync(&dt, p->d_case_dt + cid, sizeof dt, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (!(dt > 0)) {
      pool_reset_slots(p, start, ne, s);
      return fail(THESEUS_ERR_NONFINITE, "stability estimate of the case is not positive and finite");
    }
  }
  for (int e = 0; e < ne; ++e) p->slot_used[start + e] = 1;
  p->cases[cid].start = start;
  p->cases[cid].ne = ne;
  p->cases[cid].active = true;
  p->cases[cid].boundary = has_boundary;
  {
    const int rc = pool_update_active(p);
    if (rc) return rc;
  }
  if (case_id) *case_id = cid;
  if (dt_out) *dt_out = dt;
  return THESEUS_OK;
}

int theseus_pool_step(theseus_pool *p, void *stream) {
  if (!p) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null pool");
  if (p->active == 0) return THESEUS_OK;  // no case admitted
  auto key = std::make_pair(p->active, p->bnd);
  auto it = p->graphs.find(key);
  if (it == p->graphs.end()) {
    hipStream_t cs = p->capture_stream;
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    p->enqueue_step(cs);
    HIP_TRY(hipStreamEndCapture(cs, &g));
    const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      hipGraphDestroy(g);
      return fail(THESEUS_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    it = p->graphs.emplace(key, std::make_pair(g, x)).first;
  }
  HIP_TRY(hipGraphLaunch(it->second.second, (hipStream_t)stream));
  return THESEUS_OK;
}

}  // extern "C"

struct theseus_pool_group {
  theseus_pool *pools[3] = {nullptr, nullptr, nullptr};  // orders 2, 3, 4
  const OrdersKernels *ok = nullptr;
  int pm = 0;
  BasisArgs *d_basis = nullptr;  // [3]: the bases of orders 2, 3, 4
  hipStream_t capture_stream = nullptr;
  std::map<std::array<int, 4>, GroupGraph> graphs;  // (active of p = 2, 3, 4; bnd)
};

extern "C" {

int theseus_pool_group_create(theseus_pool *const *pools, int n, theseus_pool_group **out) {
  if (!pools || !out || n < 1 || n > 3) return fail(THESEUS_ERR_INVALID_ARGUMENT, "one to three pools and an output");
  *out = nullptr;
  theseus_pool_group *g = new theseus_pool_group();
  const theseus_pool *first = nullptr;
  for (int i = 0; i < n; ++i) {
    theseus_pool *p = pools[i];
    const char *why = nullptr;
    if (!p) why = "null pool";
    else if (p->cd.dim != 3 || p->cd.flow_model != THESEUS_FLOW_EULER) why = "grouped pools are 3-D Euler";
    else if (p->cd.order < 2 || p->cd.order > 4) why = "grouped pools have order 2, 3 or 4";
    else if (g->pools[p->cd.order - 2]) why = "at most one pool per order";
    else if (first && (p->cd.volume_flux != first->cd.volume_flux || p->cd.surface_flux != first->cd.surface_flux ||
                       p->pm != first->pm))
      why = "grouped pools share their flux pairing and precision";
    if (why) {
      delete g;
      return fail(THESEUS_ERR_INVALID_ARGUMENT, why);
    }
    g->pools[p->cd.order - 2] = p;
    if (!first) first = p;
  }
  g->ok = find_orders(first->cd.volume_flux, first->cd.surface_flux);
  if (!g->ok) {
    delete g;
    return fail(THESEUS_ERR_UNSUPPORTED, "the one-launch kernel is compiled for Ranocha + Lax-Friedrichs only");
  }
  g->pm = first->pm;
  BasisArgs bases[3];
  std::memset(bases, 0, sizeof bases);
  for (int i = 0; i < 3; ++i)
    if (g->pools[i]) bases[i] = g->pools[i]->basis;
  int rc = upload<BasisArgs>(&g->d_basis, bases, 3);
  if (rc) {
    delete g;
    return rc;
  }
  const hipError_t e = hipStreamCreate(&g->capture_stream);  // default priority
  if (e != hipSuccess) {
    theseus_pool_group_destroy(g);
    return fail(THESEUS_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  *out = g;
  return THESEUS_OK;
}

int theseus_pool_group_destroy(theseus_pool_group *g) {
  if (!g) return THESEUS_OK;
  for (auto &kv : g->graphs) {
    if (kv.second.exec) hipGraphExecDestroy(kv.second.exec);
    if (kv.second.graph) hipGraphDestroy(kv.second.graph);
  }
  if (g->capture_stream) hipStreamDestroy(g->capture_stream);
  hipFree(g->d_basis);
  delete g;
  return THESEUS_OK;
}

int theseus_pool_group_step(theseus_pool_group *g, void *stream) {
  if (!g) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null group");
  static constexpr int kEpb[3] = {4, 2, 1};  // elements per workgroup of orders 2, 3, 4 in k_rhs_orders
  std::array<int, 4> key = {0, 0, 0, 0};
  for (int i = 0; i < 3; ++i)
    if (g->pools[i]) {
      key[i] = g->pools[i]->active;
      key[3] |= g->pools[i]->bnd;
    }
  if (key[0] + key[1] + key[2] == 0) return THESEUS_OK;  // no case admitted
  auto it = g->graphs.find(key);
  if (it == g->graphs.end()) {
    GroupGraph gg;
    RHSArgs host[3][3];  // [stage][order]
    std::memset(host, 0, sizeof host);
    int nb[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i)
      if (g->pools[i] && key[i] > 0) {
        if (key[i] % kEpb[i]) return fail(THESEUS_ERR_UNSUPPORTED, "launch range not a multiple of the workgroup's elements");
        nb[i] = key[i] / kEpb[i];
        for (int st = 0; st < 3; ++st) host[st][i] = g->pools[i]->stage_args(st);
      }
    const int off[2] = {(nb[0] + 7) / 8 * 8, (nb[0] + 7) / 8 * 8 + (nb[1] + 7) / 8 * 8};
    hipStream_t cs = g->capture_stream;
    const OrdersLaunch launch = g->ok->euler[g->pm][key[3]];
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      for (int st = 0; st < 3; ++st) launch(host[st], g->d_basis, off, nb, cs);
      e = hipStreamEndCapture(cs, &gg.graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&gg.exec, gg.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      if (gg.graph) hipGraphDestroy(gg.graph);
      return fail(THESEUS_ERR_DEVICE, std::string("capturing the grouped step: ") + hipGetErrorString(e));
    }
    it = g->graphs.emplace(key, gg).first;
  }
  HIP_TRY(hipGraphLaunch(it->second.exec, (hipStream_t)stream));
  return THESEUS_OK;
}

int theseus_pool_read(theseus_pool *p, int case_id, double *u, void *stream) {
  if (!p || !u) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null pool or output");
  if (case_id < 0 || case_id >= (int)p->cases.size() || !p->cases[case_id].active)
    return fail(THESEUS_ERR_INVALID_ARGUMENT, "no such active case");
  const PoolCase &c = p->cases[case_id];
  const long long nd = (long long)c.ne * p->ndof, off = (long long)c.start * p->ndof;
  hipStream_t s = (hipStream_t)stream;
  for (int q = 0; q < p->neq; ++q)
    HIP_TRY(hipMemcpyAsync(u + q * nd, p->d_U + q * p->Ntot + off, nd * sizeof(double), hipMemcpyDefault, s));
  HIP_TRY(hipStreamSynchronize(s));
  return THESEUS_OK;
}

int theseus_pool_integral_measures(theseus_pool *p, int case_id, theseus_integral_measures *out, void *stream) {
  if (!p || !out) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null pool or output");
  if (case_id < 0 || case_id >= (int)p->cases.size() || !p->cases[case_id].active)
    return fail(THESEUS_ERR_INVALID_ARGUMENT, "no such active case");
  if (!p->d_meas) {
    const int rc = upload<double>(&p->d_meas, nullptr, (size_t)((p->cap + p->sk->epb - 1) / p->sk->epb + 1) * NMEAS);
    if (rc) return rc;
  }
  const PoolCase &c = p->cases[case_id];
  return measures_impl(p->sk, p->d_U, p->Ntot, c.start, c.ne, p->d_J, p->gas, p->basis, lte::View{}, p->d_meas, out,
                       (hipStream_t)stream);
}

int theseus_pool_release(theseus_pool *p, int case_id, void *stream) {
  if (!p) return fail(THESEUS_ERR_INVALID_ARGUMENT, "null pool");
  if (case_id < 0 || case_id >= (int)p->cases.size() || !p->cases[case_id].active)
    return fail(THESEUS_ERR_INVALID_ARGUMENT, "no such active case");
  PoolCase &c = p->cases[case_id];
  int rc = pool_reset_slots(p, c.start, c.ne, (hipStream_t)stream);
  if (rc) return rc;
  for (int e = 0; e < c.ne; ++e) p->slot_used[c.start + e] = 0;
  c = PoolCase();
  return pool_update_active(p);
}

}  // extern "C"
