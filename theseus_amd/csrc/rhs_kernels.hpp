// gfx950 DGSEM right-hand-side kernels.
//
// One element per workgroup (several small 2-D/low-order elements per
// 64-lane wave), one lane per GLL node.  Each workgroup
//   1. loads its element's conservative state with coalesced loads from the
//      byNODES L-vector (for each equation the element's nodes are
//      contiguous) into LDS, and derives velocity and pressure once per node;
//   2. evaluates the split-form volume term sum_m 2 D(i,m) F#(u_lo,u_hi)
//      along the three GLL lines through its node (partners read from LDS);
//   3. evaluates the Riemann flux at every point of its own faces, reading
//      the neighbour's trace straight from the L-vector (L2-resident), so
//      no face-trace arrays and no scatter/add passes exist;
//   4. optionally evaluates the Persson-Peraire indicator and the subcell
//      FV residual and blends;
//   5. writes dU (or the fused SSP-RK3 stage combination) once per node.
// For Navier-Stokes (BR1) a first kernel computes the entropy-variable
// gradient with its face lift, converts it to primitive gradients and
// stores only the physical viscous flux (symmetric stress + energy flux,
// 9 values/node in 3-D); the second kernel is the kernel above plus the
// viscous volume divergence and the central viscous face term.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "bc_device.hpp"
#include "device_lte.hpp"
#include "device_physics.hpp"

namespace thx {

constexpr int MAXNP = 8;

// EPBX > 0 overrides the elements per workgroup (a tile shape the serving
// pool's autotuner can choose).
template <int DIM, int P, int EPBX = 0>
struct Shape {
  static constexpr int Np = P + 1;
  static constexpr int NDOF = DIM == 2 ? Np * Np : Np * Np * Np;
  static constexpr int NFP = DIM == 2 ? Np : Np * Np;
  static constexpr int NEQ = DIM + 2;
  // elements per workgroup: one wave's worth of nodes, and two elements of 64
  // nodes (3-D p = 3), whose pair and face lists (288 and 96 items) then
  // end on a wave boundary every other element instead of every element
  static constexpr int EPB = EPBX > 0 ? EPBX : (NDOF == 64 ? 2 : NDOF > 64 ? 1 : 64 / NDOF);
  static constexpr int TPB = ((EPB * NDOF + 63) / 64) * 64;
  static constexpr int NPOW2 = NDOF <= 8 ? 8 : NDOF <= 16 ? 16 : NDOF <= 32 ? 32 : NDOF <= 64 ? 64 : NDOF <= 128 ? 128 : 256;
};

struct BasisArgs {
  double D[MAXNP * MAXNP];      // D(i,l) at [i*Np + l]
  double Dhat[MAXNP * MAXNP];   // D + (1/w0) e0 e0^T - (1/wN) eN eN^T
  double Dhat2[MAXNP * MAXNP];  // 2D with the same corner corrections
  double w[MAXNP];              // GLL weights on [0,1]
  double Vinv[MAXNP * MAXNP];   // inverse 1-D Legendre Vandermonde: mode a = sum_i Vinv[a*Np+i] u_i
};

// Everything a kernel launch needs but the basis (RHSArgs adds it): small
// enough that three of them fit in one launch's kernel arguments.
struct RHSCore {
  const double *U;    // stage state (L-vector)
  double *out;        // mode 0: dU; mode 1: U + dt dU; mode 2: cx X + cy (U + dt dU)
  const double *X;    // RK base state for mode 2
  double cx, cy, dt;
  int mode;
  const void *Vd;     // NS: viscous volume divergence, [DIM+1][Ntot] (momentum, energy)
  const void *Gf;     // NS: n.Fv at face points, [DIM+1][ne][2*DIM][NFP]
  void *Vd_out;       // NS pass 1 outputs; element type = the kernels' compute type
  void *Gf_out;
  double *alpha;      // blending coefficient per element (written when blending)
  long long Ntot;     // ne * NDOF
  int ne;
  int blending;
  double alpha_max, alpha_min, sharp, thresh;
  const double *J;      // [ne]
  const double *adj;    // [ne][DIM][DIM]
  const int *nbr;       // [ne][2*DIM]; -(b+1) on a face of boundary b
  const int *bc_type;   // [nbc] boundary type (BC_*)
  const double *bc_data;  // [nbc][BC_STRIDE] boundary payload
  GasC<double> gas;
  GasC<float> gas32;  // the same constants rounded to FP32 (mixed precision)
  const double *pref; // mixed precision: background pressure left out of the momentum fluxes (device scalar)
  // Serving pool: many cases in one element array.  elem_case[e] is the case
  // owning element slot e (< 0: free slot, skipped); case_dt and case_pref
  // hold each case's time step and background pressure.  All null for a
  // single case (dt and pref above apply).
  const int *elem_case;
  const double *case_dt;
  const double *case_pref;
  int xcd_remap;      // block order (logical_block)
  lte::View lte;      // tabulated LTE gas (LTEEC kernels)
};

struct RHSArgs : RHSCore {
  BasisArgs basis;
};

template <typename R>
__device__ __forceinline__ const GasC<R> &gas_of(const RHSCore &a);
template <>
__device__ __forceinline__ const GasC<double> &gas_of<double>(const RHSCore &a) { return a.gas; }
template <>
__device__ __forceinline__ const GasC<float> &gas_of<float>(const RHSCore &a) { return a.gas32; }

// Node index of in-face point t on local face (d, side)
template <int DIM, int Np>
__host__ __device__ constexpr int face_node(int d, int side, int t) {
  const int id = side ? Np - 1 : 0;
  if constexpr (DIM == 2) {
    return d == 0 ? id + Np * t : t + Np * id;
  } else {
    if (d == 0) return id + Np * t;
    if (d == 1) return (t % Np) + Np * id + Np * Np * (t / Np);
    return t + Np * Np * id;
  }
}

// In-face index of node (i,j,k) on a face normal to d
template <int DIM, int Np>
__device__ __forceinline__ int face_index(int d, int i, int j, int k) {
  if constexpr (DIM == 2) {
    return d == 0 ? j : i;
  } else {
    return d == 0 ? j + Np * k : (d == 1 ? i + Np * k : i + Np * j);
  }
}

// Whether lane t of a workgroup of EPB elements of ND nodes holds a node.
__device__ __forceinline__ bool lane_ok_pre(int t, int ND, int EPB) { return t / ND < EPB; }

// Workgroup barrier for LDS traffic only: global loads still in flight stay
// in flight across it and are waited for where their registers are used.
__device__ __forceinline__ void lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// XCD-aware block order: the 8 XCDs each receive a contiguous range of
// elements so that face neighbours share an L2 (hardware dispatches
// workgroups round-robin over the XCDs).  `remap` = 0 keeps the launch order.
__device__ __forceinline__ int logical_block(int remap) {
  const int nb = gridDim.x, b = blockIdx.x;
  if (remap && (nb & 7) == 0) return (b & 7) * (nb >> 3) + (b >> 3);
  return b;
}

// Persson-Peraire indicator of one element (RHSOperator_impl.hpp:355-527):
// modes of the indicator field pressure*density, eta = max(1 - |M1|^2/|m|^2,
// 1 - |M2|^2/|M1|^2) with M1 (M2) the modes of degree <= p-1 (p-2) in every
// direction, alpha = sigmoid(9.21 (eta - theta)/theta) cut at alpha_min and
// alpha_max.  The reference applies the inverse of the full Legendre
// Vandermonde (ModalBasis.cpp:115-140); that matrix is the Kronecker
// product of the 1-D inverse, applied here as one 1-D sweep per direction.
// `buf` holds at least 2*ND + 3*NPOW2 doubles of the element's scratch, the
// first ND of them the element's pressure*density per node (FP64 whatever
// the precision of the flux arithmetic; 0 on lanes without an element),
// written and synchronised by the caller.
template <int DIM, int P>
__device__ __forceinline__ double blending_alpha(const RHSCore &a, double *buf, const double *sVi, int node, int i,
                                                 int j, int k, bool lane_ok, bool valid) {
  using S = Shape<DIM, P>;
  constexpr int Np = S::Np, ND = S::NDOF, NP2 = S::NPOW2;
  double *A = buf, *B = buf + ND;
  // sweep along x: B(a,j,k) = sum_l Vinv(a,l) A(l,j,k)
  if (lane_ok) {
    double acc = 0.0;
    const int base = node - i;
#pragma unroll
    for (int l = 0; l < Np; ++l) acc += sVi[i * Np + l] * A[base + l];
    B[node] = acc;
  }
  lds_sync();
  if (lane_ok) {
    double acc = 0.0;
    const int base = node - j * Np;
#pragma unroll
    for (int l = 0; l < Np; ++l) acc += sVi[j * Np + l] * B[base + l * Np];
    A[node] = acc;
  }
  lds_sync();
  double mode = A[lane_ok ? node : 0];
  if constexpr (DIM == 3) {
    double acc = 0.0;
    const int base = node - k * Np * Np;
#pragma unroll
    for (int l = 0; l < Np; ++l) acc += sVi[k * Np + l] * A[lane_ok ? base + l * Np * Np : 0];
    mode = acc;
  }
  const double mode2 = mode * mode;
  const bool keep1 = i <= P - 1 && j <= P - 1 && k <= P - 1;
  const bool keep2 = i <= P - 2 && j <= P - 2 && k <= P - 2;
  double mm, m1m1, m2m2;
  if constexpr (ND <= 64 && (ND & (ND - 1)) == 0) {
    // An element's ND lanes lie in one wave: a butterfly over them pairs the
    // operands exactly as the tree below does (own + partner at offsets
    // ND/2, ..., 1; IEEE addition is commutative), so every lane ends with
    // the tree's sum, bit for bit, without LDS round trips or barriers.
    mm = valid ? mode2 : 0.0;
    m1m1 = valid && keep1 ? mode2 : 0.0;
    m2m2 = valid && keep2 ? mode2 : 0.0;
#pragma unroll
    for (int off = ND / 2; off > 0; off >>= 1) {
      mm += __shfl_xor(mm, off, 64);
      m1m1 += __shfl_xor(m1m1, off, 64);
      m2m2 += __shfl_xor(m2m2, off, 64);
    }
  } else {
    lds_sync();  // the sweeps' buffers are reused for the reduction below
    double *red = buf;  // 3 rows of NP2
    if (lane_ok) {
      red[node] = valid ? mode2 : 0.0;
      red[NP2 + node] = valid && keep1 ? mode2 : 0.0;
      red[2 * NP2 + node] = valid && keep2 ? mode2 : 0.0;
      for (int t = node + ND; t < NP2; t += ND) red[t] = red[NP2 + t] = red[2 * NP2 + t] = 0.0;
    }
    lds_sync();
    for (int stride = NP2 / 2; stride > 0; stride /= 2) {
      if (lane_ok && node < stride) {
        red[node] += red[node + stride];
        red[NP2 + node] += red[NP2 + node + stride];
        red[2 * NP2 + node] += red[2 * NP2 + node + stride];
      }
      lds_sync();
    }
    mm = red[0];
    m1m1 = red[NP2];
    m2m2 = red[2 * NP2];
  }
  const double eps = 1.0e-30;
  double val = 0.0;
  if (mm > eps) {
    val = 1.0 - m1m1 / mm;
    if (m1m1 > eps) val = fmax(val, 1.0 - m2m2 / m1m1);
    else val = 1.0;
  }
  const double eta = fmin(fmax(val, 0.0), 1.0);
  double al = 1.0 / (1.0 + exp(-a.sharp * (eta - a.thresh) / a.thresh));
  if (al < a.alpha_min) al = 0.0;
  else if (al > (1.0 - a.alpha_min)) al = 1.0;
  return fmin(al, a.alpha_max);
}

// ---------------------------------------------------------------------------
// Fused volume + face kernel (Euler; Navier-Stokes pass 2 when VISC)
// ---------------------------------------------------------------------------

// Index of the pair (a, b), a < b, among the Np(Np-1)/2 pairs of one GLL line.
template <int Np>
__device__ __forceinline__ int pair_index(int a, int b) {
  return a * Np - (a * (a + 1)) / 2 + (b - a - 1);
}
template <int Np>
__host__ __device__ constexpr void pair_nodes(int pr, int &a, int &b) {
  int aa = 0, rem = pr;
#pragma unroll
  for (int t = 0; t < Np - 1; ++t) {
    const int cnt = Np - 1 - t;
    if (aa == t && rem >= cnt) {
      rem -= cnt;
      aa = t + 1;
    }
  }
  a = aa;
  b = aa + 1 + rem;
}

template <int DIM>
__device__ __forceinline__ int vis_index(int q, int dd) {
  return q == DIM + 1 ? VisComp<DIM>::en(dd) : VisComp<DIM>::tau(q - 1, dd);
}

// Staged node arrays leave one slot free after every x-line of Np nodes
// (index n + n / Np): the nodes that a round of pair loads touches then fall
// into distinct LDS banks.
constexpr bool kLdsPad = true;
template <int Np, int ND>
struct StagedIndex {
  static constexpr int size = kLdsPad ? ND + ND / Np : ND;
  __host__ __device__ static constexpr int at(int n) { return kLdsPad ? n + n / Np : n; }
};

// Flux scratch of the volume phase of one workgroup.  The pair list is
// direction-major (direction d: pairs r = pe * NPAIR + line * NPL + pair of
// its NPT = EPB * NPAIR) and processed in rounds of TPB lanes; direction d is
// gathered after the round holding its last pair.  When d's first round
// also holds the previous direction's last pairs, those first nshared(d)
// pairs of d go to an overflow region (one per direction parity, so that a
// direction gathered in that round is not overwritten); every other pair of
// every direction goes to row r of a common region of NPT rows, which the
// direction before has released by then.  That is NPT + the overflow rows
// instead of the 2 * NPT of one region per direction in flight (3-D p = 3,
// two elements: 256 rows instead of 384).
template <int DIM, int P, int EPBX>
struct VolPlan {
  using S = Shape<DIM, P, EPBX>;
  static constexpr int Np = S::Np, ND = S::NDOF, EPB = S::EPB, TPB = S::TPB;
  static constexpr int NPL = Np * (Np - 1) / 2, NPAIR = (ND / Np) * NPL, NPT = EPB * NPAIR;
  static constexpr int NRND = (DIM * NPT + TPB - 1) / TPB;
  __host__ __device__ static constexpr int first(int d) { return d * NPT / TPB; }
  __host__ __device__ static constexpr int last(int d) { return ((d + 1) * NPT - 1) / TPB; }
  __host__ __device__ static constexpr int nshared(int d) {
    if (d == 0 || first(d) != last(d - 1)) return 0;
    const int n = (first(d) + 1) * TPB - d * NPT;
    return n < NPT ? n : NPT;
  }
  __host__ __device__ static constexpr int bsize(int parity) {
    int m = 0;
    for (int d = parity; d < DIM; d += 2) m = nshared(d) > m ? nshared(d) : m;
    return m;
  }
  __host__ __device__ static constexpr int boff(int d) { return NPT + ((d & 1) ? bsize(0) : 0); }
  static constexpr int ROWS = NPT + bsize(0) + bsize(1);
  __host__ __device__ static constexpr int phys(int d, int r) { return r < nshared(d) ? boff(d) + r : r; }
  // an overflow region is reused only once the direction before has been gathered
  __host__ __device__ static constexpr bool valid() {
    for (int d = 2; d < DIM; ++d)
      if (nshared(d) > 0 && nshared(d - 2) > 0 && !(last(d - 2) < first(d))) return false;
    return true;
  }
};

// The volume pair list of one workgroup, decoded at compile time: entry w
// of the direction-major list (pair r of direction dir in element pe) holds
// the staged LDS indices of its two nodes (bits 0-15, 16-31), its scratch
// row VolPlan::phys(dir, r) (bits 32-43), the element (44-47) and the
// direction (48-49).  The pair loop reads one 64-bit word per pair instead
// of decoding w with integer divisions and branches on the direction.
template <int DIM, int P, int EPBX>
struct PairTable {
  using S = Shape<DIM, P, EPBX>;
  static constexpr int Np = S::Np, ND = S::NDOF, EPB = S::EPB;
  static constexpr int NPL = Np * (Np - 1) / 2, NPAIR = (ND / Np) * NPL, NPT = EPB * NPAIR, N = DIM * NPT;
  static_assert(VolPlan<DIM, P, EPBX>::ROWS <= 4096 && EPB <= 16, "pair table fields");
  static_assert(VolPlan<DIM, P, EPBX>::valid(), "volume flux scratch plan");
  uint64_t v[N];
  constexpr PairTable() : v{} {
    for (int w = 0; w < N; ++w) {
      const int dir = w / NPT, rem = w - dir * NPT;
      const int pe = rem / NPAIR, r = rem - pe * NPAIR;
      const int line = r / NPL;
      int pa = 0, pb = 0;
      pair_nodes<Np>(r - line * NPL, pa, pb);
      const int sdir = dir == 0 ? 1 : (dir == 1 ? Np : Np * Np);
      const int base = face_node<DIM, Np>(dir, 0, line);
      const uint64_t sa = (uint64_t)StagedIndex<Np, ND>::at(base + pa * sdir);
      const uint64_t sb = (uint64_t)StagedIndex<Np, ND>::at(base + pb * sdir);
      const uint64_t row = (uint64_t)VolPlan<DIM, P, EPBX>::phys(dir, rem);
      v[w] = sa | (sb << 16) | (row << 32) | ((uint64_t)pe << 44) | ((uint64_t)dir << 48);
    }
  }
};
template <int DIM, int P, int EPBX>
__device__ constexpr PairTable<DIM, P, EPBX> kPairTable{};

// The face-point list of one workgroup, decoded at compile time for the
// gradient pass: entry fi (point t of local face f of element pe) holds the
// element's own node there, face_node(d, s, t) (bits 0-7), the neighbour's
// node across the face, face_node(d, 1-s, t) (8-15), the point's row
// f*NF + t (16-23), the element (24-27) and the face (28-30).
template <int DIM, int P, int EPBX>
struct FaceTable {
  using S = Shape<DIM, P, EPBX>;
  static constexpr int Np = S::Np, ND = S::NDOF, NF = S::NFP, EPB = S::EPB;
  static constexpr int NFACE = 2 * DIM * NF, N = EPB * NFACE;
  static_assert(ND <= 256 && NFACE <= 256 && EPB <= 16, "face table fields");
  uint32_t v[N];
  constexpr FaceTable() : v{} {
    for (int fi = 0; fi < N; ++fi) {
      const int pe = fi / NFACE, fl = fi - pe * NFACE;
      const int f = fl / NF, t = fl - f * NF;
      const int d = f >> 1, s = f & 1;
      const uint32_t own = (uint32_t)face_node<DIM, Np>(d, s, t);
      const uint32_t nb = (uint32_t)face_node<DIM, Np>(d, 1 - s, t);
      v[fi] = own | (nb << 8) | ((uint32_t)fl << 16) | ((uint32_t)pe << 24) | ((uint32_t)f << 28);
    }
  }
};
template <int DIM, int P, int EPBX>
__device__ constexpr FaceTable<DIM, P, EPBX> kFaceTable{};

// A node's point sample from the staged state: velocity is m * (1/rho),
// the same product derive() forms, so staged and fresh samples agree bitwise.
// `n` is a staged index (StagedIndex::at of the node).
template <typename R, int DIM, int ND, int NQ>
__device__ __forceinline__ void load_pt(const R (*sq)[ND], const R *sp, const R *sir, const R *spd, int n,
                                        Pt<R, DIM> &o) {
#pragma unroll
  for (int q = 0; q < NQ; ++q) o.q[q] = sq[q][n];
  o.p = sp[n];
  o.ir = sir[n];
  o.pd = spd[n];
#pragma unroll
  for (int d = 0; d < DIM; ++d) o.v[d] = o.q[1 + d] * o.ir;
}

template <typename R, int DIM>
__device__ __forceinline__ void pt_select(bool c, const Pt<R, DIM> &a, const Pt<R, DIM> &b, Pt<R, DIM> &o) {
#pragma unroll
  for (int q = 0; q < DIM + 2; ++q) o.q[q] = c ? a.q[q] : b.q[q];
#pragma unroll
  for (int d = 0; d < DIM; ++d) o.v[d] = c ? a.v[d] : b.v[d];
  o.p = c ? a.p : b.p;
  o.ir = c ? a.ir : b.ir;
  o.pd = c ? a.pd : b.pd;
}

// Per-element geometry staged in LDS: adjugate rows, the canonical area
// vector of each face (the low-side element's adjugate row; the outward one
// on a boundary) and the Jacobian determinant.
template <int DIM, int EPB>
struct ElemMeta {
  double adj[EPB][DIM * DIM];
  double nrm[EPB][2 * DIM][DIM];
  double J[EPB];
};

// Geometry loads.  Every load is unconditional (indices clamped into the
// arrays, results selected afterwards) so that the in-order vector-memory
// counter lets a wave wait for exactly the loads it needs; they are issued
// in two dependent waves (neighbour indices, then the neighbours' adjugate
// rows) and written to LDS by meta_store once the other loads are issued.
template <int DIM, int EPB, int TPB>
struct MetaRegs {
  static constexpr int NA = (EPB * DIM * DIM + EPB + TPB - 1) / TPB;
  static constexpr int NN = (EPB * 2 * DIM * DIM + TPB - 1) / TPB;
  double va[NA];
  double vn[NN];
  int nb[NN];
};

template <int DIM, int EPB, int TPB>
__device__ __forceinline__ void meta_load_index(const RHSCore &a, int ebase_blk, MetaRegs<DIM, EPB, TPB> &m) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int r = 0; r < MetaRegs<DIM, EPB, TPB>::NN; ++r) {
    const int w = min(tid + r * TPB, EPB * 2 * DIM * DIM - 1);
    const int pe = w / (2 * DIM * DIM);
    const int f = (w - pe * 2 * DIM * DIM) / DIM;
    const int ee = min(ebase_blk + pe, a.ne - 1);
    m.nb[r] = a.nbr[(size_t)ee * 2 * DIM + f];
  }
}

template <int DIM, int EPB, int TPB>
__device__ __forceinline__ void meta_load_own(const RHSCore &a, int ebase_blk, MetaRegs<DIM, EPB, TPB> &m) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int r = 0; r < MetaRegs<DIM, EPB, TPB>::NA; ++r) {
    const int w = min(tid + r * TPB, EPB * DIM * DIM + EPB - 1);
    const bool is_adj = w < EPB * DIM * DIM;
    const int pe = is_adj ? w / (DIM * DIM) : w - EPB * DIM * DIM;
    const int ee = min(ebase_blk + pe, a.ne - 1);
    const double *src = is_adj ? a.adj + (size_t)ee * DIM * DIM + (w - pe * DIM * DIM) : a.J + ee;
    m.va[r] = *src;
  }
}

template <int DIM, int EPB, int TPB>
__device__ __forceinline__ void meta_load_normals(const RHSCore &a, int ebase_blk, MetaRegs<DIM, EPB, TPB> &m) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int r = 0; r < MetaRegs<DIM, EPB, TPB>::NN; ++r) {
    const int w = min(tid + r * TPB, EPB * 2 * DIM * DIM - 1);
    const int pe = w / (2 * DIM * DIM);
    const int rem = w - pe * 2 * DIM * DIM;
    const int f = rem / DIM, dd = rem - f * DIM;
    const int d = f >> 1, s = f & 1;
    const int ee = min(ebase_blk + pe, a.ne - 1);
    const int nb = m.nb[r];
    const int src = (nb >= 0 && !s) ? nb : ee;
    m.vn[r] = a.adj[(size_t)src * DIM * DIM + d * DIM + dd];
  }
}

template <int DIM, int EPB, int TPB>
__device__ __forceinline__ void meta_store(const MetaRegs<DIM, EPB, TPB> &m, ElemMeta<DIM, EPB> &s) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int r = 0; r < MetaRegs<DIM, EPB, TPB>::NA; ++r) {
    const int w = tid + r * TPB;
    if (w < EPB * DIM * DIM) s.adj[w / (DIM * DIM)][w % (DIM * DIM)] = m.va[r];
    else if (w < EPB * DIM * DIM + EPB) s.J[w - EPB * DIM * DIM] = m.va[r];
  }
#pragma unroll
  for (int r = 0; r < MetaRegs<DIM, EPB, TPB>::NN; ++r) {
    const int w = tid + r * TPB;
    if (w < EPB * 2 * DIM * DIM) {
      const int f = (w % (2 * DIM * DIM)) / DIM;
      // a boundary face takes the outward area vector of the element itself
      (&s.nrm[0][0][0])[w] = (m.nb[r] < 0 && !(f & 1)) ? -m.vn[r] : m.vn[r];
    }
  }
}

// Scratch rows of one element in k_rhs: the larger of the flux buffers (in
// the compute type R) and the indicator's FP64 sweeps, in doubles.
template <typename R, int DIM, int P>
struct RhsScratch {
  using S = Shape<DIM, P>;
  static constexpr int Np = S::Np, ND = S::NDOF, NQ = S::NEQ;
  static constexpr int NF = S::NFP;
  static constexpr int NFACE = 2 * DIM * NF;
  static constexpr int NLINE = ND / Np;
  static constexpr int NPL = Np * (Np - 1) / 2;
  static constexpr int NPAIR = NLINE * NPL;
  static constexpr int NI = (Np - 1) * NLINE;
  static constexpr int NS0 = 2 * NPAIR > NFACE ? 2 * NPAIR : NFACE;  // pair fluxes: two directions in flight
  static constexpr int ROWS = NS0 > NI ? NS0 : NI;
  static constexpr int FLUX_D = (ROWS * NQ * (int)sizeof(R) + 7) / 8;
  static constexpr int IND_D = 2 * ND + 3 * S::NPOW2;  // indicator sweeps + reduction
  static constexpr int DOUBLES = FLUX_D > IND_D ? FLUX_D : IND_D;
};

// Workgroup scratch of k_rhs, in doubles: the volume flux rows of VolPlan,
// the face rows (EPB * NFACE), the subcell FV rows (EPB * NI), all in the
// compute type R, or the indicator's FP64 sweeps (EPB * IND_D), reused in
// turn.  The indicator's input (pressure * density per node) is written
// into its first sweep buffer while the state is staged.
template <typename R, int DIM, int P, int EPBX>
struct RhsWgScratch {
  using S = Shape<DIM, P, EPBX>;
  using VP = VolPlan<DIM, P, EPBX>;
  static constexpr int ND = S::NDOF, NQ = S::NEQ, EPB = S::EPB, Np = S::Np;
  static constexpr int NFACE = 2 * DIM * S::NFP;
  static constexpr int NI = (Np - 1) * (ND / Np);
  static constexpr int IND_D = 2 * ND + 3 * S::NPOW2;
  static constexpr int rdoubles(int rows) { return (rows * NQ * (int)sizeof(R) + 7) / 8; }
  static constexpr int mx(int a, int b) { return a > b ? a : b; }
  static constexpr int DOUBLES = mx(mx(rdoubles(VP::ROWS), rdoubles(EPB * NFACE)), mx(rdoubles(EPB * NI), EPB * IND_D));
};

// Occupancy the register allocator is asked to reach for k_rhs (waves per
// SIMD).  At FP64 3-D p = 3 (two elements per workgroup) the LDS (19.9 KB per
// workgroup) allows 8 workgroups, i.e. 4 waves per SIMD, if the registers fit
// 128 VGPRs; they do without spilling for the central and Ranocha volume
// fluxes with the HLL or Lax-Friedrichs face flux (not for Navier-Stokes
// with boundary conditions), which get the request; elsewhere it is not made
// (1), since forcing it spills to scratch memory.
template <typename R, int DIM, int P, int VOL, int SURF, bool VISC, int EPBX, bool BND>
struct RhsWaves {
  static constexpr bool fits = sizeof(R) == sizeof(double) && DIM == 3 && P == 3 && EPBX == 0 &&
                               (VOL == 0 || VOL == 2) && (SURF == 4 || SURF == 5) && !(VISC && BND);
  static constexpr int value = fits ? 4 : 1;
};

// Workgroup LDS of the fused kernel.  RhsLds holds it as one object (the
// multi-order kernel overlays the orders' in a union); RhsLdsRef names the
// arrays wherever they live, so the single-order kernel keeps them as
// separate LDS variables.
template <typename R, int DIM, int P, int EPBX>
struct RhsLdsDims {
  using S = Shape<DIM, P, EPBX>;
  using SC = RhsWgScratch<R, DIM, P, EPBX>;
  static constexpr int Np = S::Np, NQ = S::NEQ, EPB = S::EPB;
  static constexpr int NDS = StagedIndex<Np, S::NDOF>::size;
  static constexpr bool FP64 = sizeof(R) == sizeof(double);
  static constexpr int PE = FP64 ? 1 : EPB, PN = FP64 ? 1 : NDS;
};

template <typename R, int DIM, int P, int EPBX>
struct RhsLds {
  using T = RhsLdsDims<R, DIM, P, EPBX>;
  R sq[T::EPB][T::NQ][T::NDS];
  R sp[T::EPB][T::NDS];
  R sir[T::EPB][T::NDS];
  R spd_[T::PE][T::PN];
  double sraw[T::SC::DOUBLES];
  R sD2[T::Np * T::Np];
  double sVi[T::Np * T::Np];
  ElemMeta<DIM, T::EPB> sm;
  int sblend;
  double sprefe[T::PE];
};

template <typename R, int DIM, int P, int EPBX>
struct RhsLdsRef {
  using T = RhsLdsDims<R, DIM, P, EPBX>;
  R (&sq)[T::EPB][T::NQ][T::NDS];
  R (&sp)[T::EPB][T::NDS];
  R (&sir)[T::EPB][T::NDS];
  R (&spd_)[T::PE][T::PN];
  double (&sraw)[T::SC::DOUBLES];
  R (&sD2)[T::Np * T::Np];
  double (&sVi)[T::Np * T::Np];
  ElemMeta<DIM, T::EPB> &sm;
  int &sblend;
  double (&sprefe)[T::PE];
  __device__ static RhsLdsRef of(RhsLds<R, DIM, P, EPBX> &l) {
    return RhsLdsRef{l.sq, l.sp, l.sir, l.spd_, l.sraw, l.sD2, l.sVi, l.sm, l.sblend, l.sprefe};
  }
};

// Fused volume + face kernel body: the Euler right-hand side, or the second
// Navier-Stokes pass when VISC.  R is the type of the flux arithmetic and of
// the LDS staging (double, or float in the mixed mode); the state, the RK
// combination and the blending indicator are FP64 in both.  BND: the mesh
// has boundary faces (a periodic mesh runs the variant without the
// boundary-condition code).  `blk` is the workgroup's logical block (its
// first element is blk * EPB); the body runs on TPB lanes.  B is the basis
// (a.basis, passed apart so that a caller may hand over a copy of `a`
// without it).
template <typename R, int DIM, int P, int VOL, int SURF, bool VISC, int EPBX, bool BND>
__device__ __forceinline__ void rhs_body(const RHSCore &a, const BasisArgs &B, const int blk,
                                         RhsLdsRef<R, DIM, P, EPBX> L) {
  using S = Shape<DIM, P, EPBX>;
  using SC = RhsWgScratch<R, DIM, P, EPBX>;
  using VP = VolPlan<DIM, P, EPBX>;
  constexpr int Np = S::Np, ND = S::NDOF, NF = S::NFP, NQ = S::NEQ, EPB = S::EPB, TPB = S::TPB;
  constexpr int NVD = DIM + 1;                // viscous terms: momentum and energy
  constexpr int NFACE = 2 * DIM * NF;
  constexpr int NFR = (EPB * NFACE + TPB - 1) / TPB;  // face-point rounds per lane
  constexpr int NPL = VP::NPL, NPAIR = VP::NPAIR, NI = SC::NI;
  constexpr bool FP64 = sizeof(R) == sizeof(double);

  using SX = StagedIndex<Np, ND>;
  constexpr int NDS = SX::size;
  auto &sq = L.sq;
  auto &sp = L.sp;
  auto &sir = L.sir;
  auto &spd_ = L.spd_;  // p - p_ref (mixed precision)
  auto &sraw = L.sraw;  // flux rows (as R) or indicator sweeps (FP64), see RhsWgScratch
  auto &sD2 = L.sD2;
  auto &sVi = L.sVi;
  auto &sm = L.sm;
  auto &sblend = L.sblend;
  // each element's background pressure (mixed mode): lanes evaluate face
  // points of the other elements of the workgroup, which in the serving
  // pool may belong to another case
  auto &sprefe = L.sprefe;
  auto sscr = [&](int row, int q) -> R & { return reinterpret_cast<R *>(sraw)[row * NQ + q]; };
  double *const sind = sraw + (lane_ok_pre(threadIdx.x, ND, EPB) ? (threadIdx.x / ND) * SC::IND_D : 0);
  // pressure for the momentum fluxes: p itself in FP64, p - p_ref in the mixed mode
  auto spd = [&](int pe) -> const R * {
    if constexpr (FP64) return sp[pe];
    else return spd_[pe];
  };

  // variable-step clock: a stage of a step of length 0 (the end time was
  // reached) does nothing, so the state is left exactly as it is
  if (a.mode != 0 && a.case_dt && !a.elem_case && a.case_dt[0] == 0.0) return;
  const GasC<R> &g = gas_of<R>(a);
  const int tid = threadIdx.x;
  const int el = tid / ND;
  const int node = tid - el * ND;
  const bool lane_ok = el < EPB;
  const int elc = lane_ok ? el : 0;
  const int ebase_blk = blk * EPB;
  const int e = ebase_blk + el;
  // serving pool: the case owning this element (free slots are skipped)
  const int ecase = a.elem_case ? a.elem_case[min(ebase_blk + elc, a.ne - 1)] : 0;
  if (EPB == 1 && ecase < 0) return;  // uniform over the workgroup
  const bool valid = lane_ok && e < a.ne && ecase >= 0;
  const long long Ntot = a.Ntot;
  const long long Gtot = (long long)a.ne * NFACE;
  const R *Vd = static_cast<const R *>(a.Vd);
  const R *Gf = static_cast<const R *>(a.Gf);
  const double pref = FP64 ? 0.0 : (a.case_pref ? a.case_pref[max(ecase, 0)] : *a.pref);
  const double dt = a.case_dt ? a.case_dt[max(ecase, 0)] : a.dt;

  // ---- global loads, all before the first barrier and unconditional:
  //      neighbour indices first, then the element's own data, then what
  //      depends on the indices (the neighbours' face traces and adjugate rows)
  constexpr int NFT = EPB * NFACE;
  constexpr int NPW = DIM * EPB * NPAIR;  // volume pairs of the workgroup
  // this lane's first volume pair, from the compile-time pair table (later
  // rounds' entries are fetched one round ahead in the pair loop)
  uint64_t pcode = kPairTable<DIM, P, EPBX>.v[min(tid, NPW - 1)];
  int fnb[NFR];
#pragma unroll
  for (int r = 0; r < NFR; ++r) {
    const int fi = min(tid + r * TPB, NFT - 1);
    const int pe = fi / NFACE;
    const int ee = min(ebase_blk + pe, a.ne - 1);
    fnb[r] = a.nbr[(size_t)ee * 2 * DIM + (fi - pe * NFACE) / NF];
  }
  MetaRegs<DIM, EPB, TPB> mr;
  meta_load_index<DIM, EPB, TPB>(a, ebase_blk, mr);
  meta_load_own<DIM, EPB, TPB>(a, ebase_blk, mr);

  const size_t ebase = (size_t)(valid ? e : 0) * ND + node;
  double qown[NQ];
  R fsum[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    qown[q] = a.U[q * Ntot + ebase];
    fsum[q] = R(0);
  }
  static_assert(Np * Np <= TPB, "one basis entry per lane");
  const int kb = min(tid, Np * Np - 1);
  const double bD2 = B.Dhat2[kb], bVi = B.Vinv[kb];

  meta_load_normals<DIM, EPB, TPB>(a, ebase_blk, mr);
  if (tid < Np * Np) {
    sD2[tid] = R(bD2);
    sVi[tid] = bVi;
  }

  // ---- stage the element's state (coalesced: nodes are contiguous per equation)
  meta_store<DIM, EPB, TPB>(mr, sm);
  if (tid == 0) sblend = 0;
  if constexpr (!FP64)
    if (lane_ok && node == 0) sprefe[el] = pref;
  if (valid) {
    Pt<R, DIM> me;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      me.q[q] = R(qown[q]);
      sq[el][q][SX::at(node)] = me.q[q];
    }
    derive(g, me);
    sir[el][SX::at(node)] = me.ir;
    if constexpr (FP64) {
      sp[el][SX::at(node)] = me.p;
      if (a.blending) sind[node] = me.p * me.q[0];
    } else {
      // pressure from the FP64 state: p rounded once, p - p_ref formed before rounding
      Pt<double, DIM> m64;
#pragma unroll
      for (int q = 0; q < NQ; ++q) m64.q[q] = qown[q];
      derive(a.gas, m64);
      sp[el][SX::at(node)] = R(m64.p);
      spd_[el][SX::at(node)] = R(m64.p - pref);
      if (a.blending) sind[node] = m64.p * m64.q[0];
    }
  } else if (lane_ok && a.blending) {
    sind[node] = 0.0;
  }
  lds_sync();

  const int i = node % Np;
  const int j = DIM > 1 ? (node / Np) % Np : 0;
  const int k = DIM > 2 ? node / (Np * Np) : 0;
  const int idx[3] = {i, j, k};
  constexpr int stride[3] = {1, Np, Np * Np};
  const double Je = sm.J[elc];

  // ---- blending coefficient (needs only the staged state; FP64)
  double alpha = 0.0;
  if (a.blending) {
    alpha = blending_alpha<DIM, P>(a, sind, sVi, node, i, j, k, lane_ok, valid);
    if (valid && node == 0) a.alpha[e] = alpha;
    if (valid && alpha > 1e-16) sblend = 1;
    lds_sync();
  }

  // ---- volume: split-form flux differencing, each pair evaluated once
  //      (DGSEMIntegrator.hpp:124-218: dU_i = -(1/J) sum_dir sum_{m != i} Dhat2(i,m) F#(u_lo,u_hi)).
  //      The pairs of all directions form one list processed in full
  //      rounds; direction dir's fluxes go to scratch slot dir & 1 and are
  //      gathered as soon as the round holding its last pair is done.
  constexpr int NPT = VP::NPT;    // pairs of one direction
  constexpr int NRND = VP::NRND;
  R rate[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) rate[q] = R(0);
#pragma unroll 1
  for (int rr = 0; rr < NRND; ++rr) {
    const int w = rr * TPB + tid;
    const uint64_t code = pcode;
    if (rr + 1 < NRND) pcode = kPairTable<DIM, P, EPBX>.v[min(w + TPB, NPW - 1)];
    if (w < DIM * NPT) {
      const int pe = (int)((code >> 44) & 15u);
      if (ebase_blk + pe < a.ne) {
        const int sa = (int)(code & 0xFFFFu), sb = (int)((code >> 16) & 0xFFFFu);
        const int row = (int)((code >> 32) & 0xFFFu), dir = (int)((code >> 48) & 3u);
        Pt<R, DIM> lo, hi;
        load_pt<R, DIM, NDS, NQ>(sq[pe], sp[pe], sir[pe], spd(pe), sa, lo);
        load_pt<R, DIM, NDS, NQ>(sq[pe], sp[pe], sir[pe], spd(pe), sb, hi);
        R met[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) met[d] = R(sm.adj[pe][dir * DIM + d]);
        R F[NQ];
        volume_flux<VOL, R, DIM>(g, lo, hi, met, F);
#pragma unroll
        for (int q = 0; q < NQ; ++q) sscr(row, q) = F[q];
      }
    }
#pragma unroll
    for (int dir = 0; dir < DIM; ++dir) {
      if (VP::last(dir) != rr) continue;
      lds_sync();
      if (valid) {
        const int mi = dir == 0 ? i : (dir == 1 ? j : k);
        const int line = face_index<DIM, Np>(dir, i, j, k);
#pragma unroll
        for (int m = 0; m < Np; ++m) {
          if (m == mi) continue;
          const int pr = m < mi ? pair_index<Np>(m, mi) : pair_index<Np>(mi, m);
          const R coef = sD2[mi * Np + m];
          const int prow = VP::phys(dir, el * NPAIR + line * NPL + pr);
#pragma unroll
          for (int q = 0; q < NQ; ++q) rate[q] += coef * sscr(prow, q);
        }
      }
      lds_sync();
    }
  }
  R du[NQ];
  {
    const R scale = R(-1.0 / Je);
#pragma unroll
    for (int q = 0; q < NQ; ++q) du[q] = scale * rate[q];
  }

  // ---- subcell FV residual and blend (EulerOperator_impl.hpp:120-157, DGSEMIntegrator.hpp:338-483)
  if (a.blending && sblend) {
    R dfv[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) dfv[q] = R(0);
#pragma unroll
    for (int dir = 0; dir < DIM; ++dir) {
      for (int w = tid; w < EPB * NI; w += TPB) {
        const int pe = w / NI;
        const int ii = w - pe * NI;
        if (ebase_blk + pe < a.ne) {
          const int L = ii / (Np - 1), pos = ii - L * (Np - 1);
          const int ln = face_node<DIM, Np>(dir, 0, L) + pos * stride[dir];
          Pt<R, DIM> sl, sr;
          load_pt<R, DIM, NDS, NQ>(sq[pe], sp[pe], sir[pe], spd(pe), SX::at(ln), sl);
          load_pt<R, DIM, NDS, NQ>(sq[pe], sp[pe], sir[pe], spd(pe), SX::at(ln + stride[dir]), sr);
          R nor[DIM];
#pragma unroll
          for (int d = 0; d < DIM; ++d) nor[d] = R(sm.adj[pe][dir * DIM + d]);
          R F[NQ];
          surface_flux<SURF, R, DIM>(g, sl, sr, nor, F);
#pragma unroll
          for (int q = 0; q < NQ; ++q) sscr(pe * NI + ii, q) = F[q];
        }
      }
      lds_sync();
      if (valid) {
        const int L = face_index<DIM, Np>(dir, i, j, k);
        const int pos = dir == 0 ? i : (dir == 1 ? j : k);
        // one reciprocal per direction instead of a division per field
        const R rden = drcp(R(Je * B.w[pos]));
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const R fm = pos > 0 ? sscr(el * NI + L * (Np - 1) + pos - 1, q) : R(0);
          const R fp = pos < Np - 1 ? sscr(el * NI + L * (Np - 1) + pos, q) : R(0);
          const R val = (fm - fp) * rden;
          dfv[q] = dir == 0 ? val : dfv[q] + val;
        }
      }
      lds_sync();
    }
    if (valid && alpha > 1e-16) {
      const R a_dg = R(1.0 - alpha), a_fv = R(alpha);
#pragma unroll
      for (int q = 0; q < NQ; ++q) du[q] = a_dg * du[q] + a_fv * dfv[q];
    }
  }

  // the neighbours' face traces (state and, for Navier-Stokes, n.Fv),
  // issued once the volume and subcell terms no longer hold registers
  double qn[NFR][NQ];
  R gn[NFR][VISC ? NVD : 1];
#pragma unroll
  for (int r = 0; r < NFR; ++r) {
    const int fi = min(tid + r * TPB, NFT - 1);
    const int pe = fi / NFACE;
    const int fl = fi - pe * NFACE;
    const int f = fl / NF, t = fl - f * NF;
    const int d = f >> 1, s = f & 1;
    const int nbs = max(fnb[r], 0);
    size_t nbase = (size_t)nbs * ND + face_node<DIM, Np>(d, 1 - s, t);
    // a boundary point loads the element's own trace (FP64), the input of
    // its exterior state and wall pressure
    if constexpr (BND)
      if (fnb[r] < 0) nbase = (size_t)min(ebase_blk + pe, a.ne - 1) * ND + face_node<DIM, Np>(d, s, t);
#pragma unroll
    for (int q = 0; q < NQ; ++q) qn[r][q] = a.U[q * Ntot + nbase];
    if constexpr (VISC) {
      // the neighbour's n.Fv at this point (the own side's half is already in Vd)
      const size_t gb_n = (size_t)nbs * NFACE + (f ^ 1) * NF + t;
#pragma unroll
      for (int c = 0; c < NVD; ++c) gn[r][c] = Gf[c * Gtot + gb_n];
    }
  }
  // ---- faces: Riemann flux at every point of the element's own faces
  //      (EulerOperator_impl.hpp:211-392; NSOperator_impl.hpp:579-713),
  //      neighbour traces from the registers loaded at the top, which
  //      arrived while the volume terms were computed
  {
    const double w0 = B.w[0];
#pragma unroll
    for (int r = 0; r < NFR; ++r) {
      const int fi = tid + r * TPB;
      const int pe = fi / NFACE;
      const int fl = fi - pe * NFACE;
      if (fi < EPB * NFACE && ebase_blk + pe < a.ne) {
        const int f = fl / NF, t = fl - f * NF;
        const int d = f >> 1, s = f & 1;
        Pt<R, DIM> own;
        load_pt<R, DIM, NDS, NQ>(sq[pe], sp[pe], sir[pe], spd(pe), SX::at(face_node<DIM, Np>(d, s, t)), own);
        const R wown = R(1.0 / (w0 * sm.J[pe]));
        R n[DIM];
#pragma unroll
        for (int dd = 0; dd < DIM; ++dd) n[dd] = R(sm.nrm[pe][f][dd]);
        // One flux evaluation per face point, operands put in canonical
        // (low side first) order by selects, so that lanes on the two sides
        // of the element do not diverge.  A boundary point evaluates
        // F*(q, q_ext, n_out) (ApplyBoundaryConditionInviscid,
        // bc_kernels.hpp:219-287) with the interior trace first, or the
        // wall flux (0, p* n_out, 0).
        const bool interior = !BND || fnb[r] >= 0;
        R F[NQ];
        if constexpr (!BND) {
          Pt<R, DIM> oth;  // the neighbour's trace (loaded at the top)
#pragma unroll
          for (int q = 0; q < NQ; ++q) oth.q[q] = R(qn[r][q]);
          derive(g, oth);
          if constexpr (!FP64) {
            // pressure from the FP64 trace, as for the staged nodes
            Pt<double, DIM> o64;
#pragma unroll
            for (int q = 0; q < NQ; ++q) o64.q[q] = qn[r][q];
            derive(a.gas, o64);
            oth.p = R(o64.p);
            oth.pd = R(o64.p - sprefe[FP64 ? 0 : pe]);
          }
          // Lax-Friedrichs without ordering the operands (face_llf_signed):
          // 78 selects fewer per lane; the FP64 Euler kernel keeps the
          // ordered form, with which its 4-wave register budget holds
          if constexpr (SURF == 5 && (VISC || !FP64)) {
            R nn = R(0);
#pragma unroll
            for (int dd = 0; dd < DIM; ++dd) nn += n[dd] * n[dd];
            face_llf_signed<R, DIM>(g, own, oth, n, dsqrt(nn), s != 0 ? R(1) : R(-1), F);
          } else {
            Pt<R, DIM> pl, pr;
            pt_select(s != 0, own, oth, pl);
            pt_select(s != 0, oth, own, pr);
            surface_flux<SURF, R, DIM>(g, pl, pr, n, F);
          }
        } else {
          double qx[NQ];  // FP64 exterior state: the neighbour's trace, or the boundary's
#pragma unroll
          for (int q = 0; q < NQ; ++q) qx[q] = qn[r][q];
          int bct = BC_SUPERSONIC_OUTFLOW;
          if (!interior) {
            const int b = -(fnb[r] + 1);
            bct = a.bc_type[b];
            bc_exterior<DIM>(bct, a.bc_data + (size_t)b * BC_STRIDE, sm.nrm[pe][f], qx);
          }
          Pt<R, DIM> oth;
#pragma unroll
          for (int q = 0; q < NQ; ++q) oth.q[q] = R(qx[q]);
          derive(g, oth);
          if constexpr (!FP64) {
            // pressure from the FP64 state, as for the staged nodes
            Pt<double, DIM> o64;
#pragma unroll
            for (int q = 0; q < NQ; ++q) o64.q[q] = qx[q];
            derive(a.gas, o64);
            oth.p = R(o64.p);
            oth.pd = R(o64.p - sprefe[FP64 ? 0 : pe]);
          }
          Pt<R, DIM> pl, pr;
          const bool own_left = s != 0 || !interior;
          pt_select(own_left, own, oth, pl);
          pt_select(own_left, oth, own, pr);
          surface_flux<SURF, R, DIM>(g, pl, pr, n, F);
          if (!interior && (bc_wall_pressure(bct, VISC) || bc_zero_flux(bct, VISC))) {
            // SlipWallInviscidFluxKernel (bc_kernels.hpp:138-152) and the
            // walls' pressure term (:181-183, :212-214), from the FP64 trace
            const bool zero = bc_zero_flux(bct, VISC);
            const double pw = zero ? 0.0 : slipwall_pstar<DIM>(a.gas, qn[r], sm.nrm[pe][f]) - (FP64 ? 0.0 : sprefe[FP64 ? 0 : pe]);
            F[0] = R(0);
#pragma unroll
            for (int dd = 0; dd < DIM; ++dd) F[1 + dd] = zero ? R(0) : R(pw * sm.nrm[pe][f][dd]);
            F[DIM + 1] = R(0);
          }
        }
        if constexpr (VISC) {
          // F* - 0.5 n.(Fv_L + Fv_R)   (DGSEMIntegrator.hpp:261-316); the own
          // side's -0.5 n.Fv was folded into Vd by the gradient kernel
#pragma unroll
          for (int c = 0; c < NVD; ++c) F[1 + c] -= interior ? R(0.5) * gn[r][c] : R(0);
        }
        const R cs = (interior && !s) ? wown : -wown;
#pragma unroll
        for (int q = 0; q < NQ; ++q) sscr(pe * NFACE + fl, q) = cs * F[q];
      }
    }
  }
  // RK base state (mode 2) and, for Navier-Stokes, the viscous volume
  // divergence with the own half of the face term (from the gradient
  // kernel): needed only by the face gather and the epilogue, so issued
  // after the face fluxes to keep them out of that phase's registers.
  // The base state is loaded unconditionally, from the stage state itself
  // in the other modes (lines already in cache), so that waits stay exact;
  // the mixed mode also re-reads the FP64 stage state (its LDS copy is FP32).
  double xb[NQ], q64[FP64 ? 1 : NQ];
  {
    const double *Xp = a.mode == 2 ? a.X : a.U;
#pragma unroll
    for (int q = 0; q < NQ; ++q) xb[q] = Xp[q * Ntot + ebase];
    if constexpr (!FP64) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) q64[q] = a.U[q * Ntot + ebase];
    }
  }
  if constexpr (VISC) {
#pragma unroll
    for (int c = 0; c < NVD; ++c) fsum[1 + c] = Vd[c * Ntot + ebase];
  }
  lds_sync();
  if (valid) {
    // a node lies on at most one face per direction (Np >= 2)
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const bool hi = idx[d] == Np - 1;
      const bool on = hi || idx[d] == 0;
      const int row = (2 * d + (hi ? 1 : 0)) * NF + face_index<DIM, Np>(d, i, j, k);
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const R v = sscr(el * NFACE + row, q);
        fsum[q] += on ? v : R(0);
      }
    }
  }

  // ---- epilogue: dU, or the fused SSP-RK3 stage combination (FP64)
  if (valid) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const double r = FP64 ? double(du[q] + fsum[q]) : double(du[q]) + double(fsum[q]);
      double qo;
      if constexpr (FP64) qo = sq[el][q][SX::at(node)];
      else qo = q64[q];
      double o;
      if (a.mode == 0) o = r;
      else if (a.mode == 1) o = qo + dt * r;
      else o = a.cx * xb[q] + a.cy * (qo + dt * r);
      a.out[q * Ntot + ebase] = o;
    }
  }
}

template <typename R, int DIM, int P, int VOL, int SURF, bool VISC, int EPBX = 0, bool BND = false>
__global__ void __launch_bounds__((Shape<DIM, P, EPBX>::TPB))
__attribute__((amdgpu_waves_per_eu(RhsWaves<R, DIM, P, VOL, SURF, VISC, EPBX, BND>::value))) k_rhs(RHSArgs a) {
  using T = RhsLdsDims<R, DIM, P, EPBX>;
  __shared__ R sq[T::EPB][T::NQ][T::NDS];
  __shared__ R sp[T::EPB][T::NDS];
  __shared__ R sir[T::EPB][T::NDS];
  __shared__ R spd_[T::PE][T::PN];
  __shared__ double sraw[T::SC::DOUBLES];
  __shared__ R sD2[T::Np * T::Np];
  __shared__ double sVi[T::Np * T::Np];
  __shared__ ElemMeta<DIM, T::EPB> sm;
  __shared__ int sblend;
  __shared__ double sprefe[T::PE];
  rhs_body<R, DIM, P, VOL, SURF, VISC, EPBX, BND>(
      a, a.basis, logical_block(a.xcd_remap),
      RhsLdsRef<R, DIM, P, EPBX>{sq, sp, sir, spd_, sraw, sD2, sVi, sm, sblend, sprefe});
}

// Several orders in one launch (3-D Euler, 128 lanes per workgroup: p = 2
// with four elements per workgroup, p = 3 with two, p = 4 with one): blocks
// [0, off[0]) run order 2 on args[0], [off[0], off[1]) order 3 on args[1],
// [off[1], gridDim.x) order 4 on args[2]; each order's range starts on a
// multiple of 8 blocks so its XCD-aware block order stays within the range,
// and its trailing padding blocks return at once.  The three orders' LDS
// share one union; their arguments are kernel arguments, the bases
// (basis[0..2]) device memory.  Its register count is the largest of the
// three orders' code paths: 128 in FP64, as the single-order kernels; 96 in
// the mixed mode (5 waves per SIMD, where the single-order kernels take 80
// and run 6) and 128 with boundary conditions -- the smallest counts that
// need no scratch memory.
template <typename R, int VOL, int SURF, bool BND>
__global__ void __launch_bounds__(128)
__attribute__((amdgpu_waves_per_eu(sizeof(R) == sizeof(double) || BND ? 4 : 5))) k_rhs_orders(RHSCore a2, RHSCore a3, RHSCore a4, const BasisArgs *basis,
                                                    int off0, int off1, int nb0, int nb1, int nb2) {
  static_assert(Shape<3, 2, 4>::TPB == 128 && Shape<3, 3>::TPB == 128 && Shape<3, 4>::TPB == 128, "128 lanes");
  __shared__ union U {
    RhsLds<R, 3, 2, 4> p2;
    RhsLds<R, 3, 3, 0> p3;
    RhsLds<R, 3, 4, 0> p4;
  } lds;
  const int b = blockIdx.x;
  auto remap = [](int lb, int nb, int remap_on) {
    return (remap_on && (nb & 7) == 0) ? (lb & 7) * (nb >> 3) + (lb >> 3) : lb;
  };
  if (b < off0) {
    if (b >= nb0) return;
    rhs_body<R, 3, 2, VOL, SURF, false, 4, BND>(a2, basis[0], remap(b, nb0, a2.xcd_remap),
                                                RhsLdsRef<R, 3, 2, 4>::of(lds.p2));
  } else if (b < off1) {
    const int lb = b - off0;
    if (lb >= nb1) return;
    rhs_body<R, 3, 3, VOL, SURF, false, 0, BND>(a3, basis[1], remap(lb, nb1, a3.xcd_remap),
                                                RhsLdsRef<R, 3, 3, 0>::of(lds.p3));
  } else {
    const int lb = b - off1;
    if (lb >= nb2) return;
    rhs_body<R, 3, 4, VOL, SURF, false, 0, BND>(a4, basis[2], remap(lb, nb2, a4.xcd_remap),
                                                RhsLdsRef<R, 3, 4, 0>::of(lds.p4));
  }
}

// ---------------------------------------------------------------------------
// Navier-Stokes pass 1 (BR1): entropy variables, their gradient with the
// interior-face lift, conversion to primitive gradients and the physical
// viscous flux (NSOperator_impl.hpp:10-577, DGSEMIntegrator.hpp:764-1044,
// EOS.hpp:206-263, NavierStokesFlux.hpp:389-519).  It stores what the
// second kernel consumes: per node the viscous volume divergence
// (1/J) sum_dir sum_l Dhat(i,l) (adj_dir . Fv)(l)  (DGSEMIntegrator.hpp:486-622)
// plus the element's own half of the central viscous face flux,
// -0.5 n.Fv lifted with the face weight, and n.Fv at every point of the
// element's faces for the neighbours' halves (n the canonical area vector).
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Matrix-core line contractions of the gradient pass (FP32, 3-D, p = 3, one
// element per wave, lane = node i + 4j + 16k).  v_mfma_f32_4x4x1f32 computes
// 16 independent 4x4 outer products per wave: lane t receives
// out[r] += a(lane 4(t/4) + r) * b(lane t).  With b(t) = Op[t%4][l] at step
// l (l = 0..3) and a(4b + r) = field r along line b at position l, the four
// steps give lane t the contraction sum_l Op[t%4][l] f_r(line b, l) for four
// fields r at position t%4 of line b = t/4 -- in the same order of fused
// multiply-adds as the per-lane loop, so bitwise the same values.  Lines in x
// are numbered so that lane t holds node t; lines in y and z put the node of
// lane t elsewhere, and one ds_bpermute per value brings it home.
// ---------------------------------------------------------------------------
template <typename R, int DIM, int P, int EPBX>
struct GradMfma {
  static constexpr bool on = sizeof(R) == sizeof(float) && DIM == 3 && P == 3 && Shape<DIM, P, EPBX>::NDOF == 64 &&
                             Shape<DIM, P, EPBX>::TPB == 64 * Shape<DIM, P, EPBX>::EPB;
};

typedef float mfma_f4 __attribute__((ext_vector_type(4)));

// lane holding node t's value after a y / z line contraction
__device__ __forceinline__ int mfma_src_y(int t) { return 4 * (t & 3) + 16 * (t >> 4) + ((t >> 2) & 3); }
__device__ __forceinline__ int mfma_src_z(int t) { return 4 * (t & 15) + (t >> 4); }
__device__ __forceinline__ float lane_fetch(int src, float v) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src * 4, __float_as_int(v)));
}

// Lines of direction DIR: lane t's base node and stride (block b = t/4 is the
// line, numbered so that in x lane t holds node t).
template <int DIR>
__device__ __forceinline__ int mfma_line_base(int b) {
  return DIR == 0 ? 4 * b : (DIR == 1 ? (b & 3) + 16 * (b >> 2) : b);
}
constexpr int mfma_line_stride(int dir) { return dir == 0 ? 1 : (dir == 1 ? 4 : 16); }

// sum_l Op[pos][l] * F[r][node at position l of the line] for the fields
// r = 0..3 of direction DIR, returned at the lane of each node.
template <int DIR>
__device__ __forceinline__ void line_contract_dir(const float (*F)[64], const float op[4], int t, float out[4]) {
  const int b = t >> 2, r = t & 3, base = mfma_line_base<DIR>(b);
  mfma_f4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int l = 0; l < 4; ++l)
    c = __builtin_amdgcn_mfma_f32_4x4x1f32(F[r][base + mfma_line_stride(DIR) * l], op[l], c, 0, 0, 0);
  const int src = DIR == 0 ? t : (DIR == 1 ? mfma_src_y(t) : mfma_src_z(t));
#pragma unroll
  for (int q = 0; q < 4; ++q) out[q] = DIR == 0 ? c[q] : lane_fetch(src, c[q]);
}

// All three directions of five fields: out[d][r]; the fifth field's three
// directions share one product set (rows 0, 1, 2 of its blocks).  Every
// lane of the wave must call it.
__device__ __forceinline__ void line_contract_mfma5(const float (*F)[64], const float op[4], int t, float (*out)[5]) {
  float o[4];
  line_contract_dir<0>(F, op, t, o);
#pragma unroll
  for (int q = 0; q < 4; ++q) out[0][q] = o[q];
  line_contract_dir<1>(F, op, t, o);
#pragma unroll
  for (int q = 0; q < 4; ++q) out[1][q] = o[q];
  line_contract_dir<2>(F, op, t, o);
#pragma unroll
  for (int q = 0; q < 4; ++q) out[2][q] = o[q];
  const int b = t >> 2, r = t & 3;
  const int b4 = r == 0 ? mfma_line_base<0>(b) : (r == 1 ? mfma_line_base<1>(b) : mfma_line_base<2>(b));
  const int s4 = r == 0 ? 1 : (r == 1 ? 4 : 16);
  mfma_f4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int l = 0; l < 4; ++l) c = __builtin_amdgcn_mfma_f32_4x4x1f32(r < 3 ? F[4][b4 + s4 * l] : 0.f, op[l], c, 0, 0, 0);
  out[0][4] = c[0];
  out[1][4] = lane_fetch(mfma_src_y(t), c[1]);
  out[2][4] = lane_fetch(mfma_src_z(t), c[2]);
}

// Occupancy requested for the gradient pass: 5 waves per SIMD for FP64 3-D
// p = 3 on periodic meshes (at most 96 VGPRs; the LDS allows 5.5), where
// the allocator would otherwise take 116 and run 4; 8 for its FP32 form
// with the matrix-core contractions (64 VGPRs, as without them); elsewhere
// the default.  With the boundary code either would spill.
template <typename R, int DIM, int P, int EPBX, bool BND, bool LTE = false>
struct GradWaves {
  static constexpr bool fits = sizeof(R) == sizeof(double) && DIM == 3 && P == 3 && EPBX == 0 && !BND && !LTE;
  static constexpr int value = fits ? 5 : (GradMfma<R, DIM, P, EPBX>::on && !BND && !LTE ? 8 : 1);  // 64 VGPRs
};

// The tabulated gas on the BR1 path: the primitive gradient (0, grad u, grad T)
// from the entropy-variable gradient (LTEHelmholtzGasEOS.hpp:159-171) and the
// viscous flux with the node's tabulated viscosity and conductivity, grad T
// taken from the energy slot (grad_temperature, :133-142)
template <typename R, int DIM>
__device__ __forceinline__ void grad_entropy_to_grad_prim(const lte::ViscNode<DIM, R> &t, const R *, const R *dE,
                                                          R *dP) {
  const R T = t.T;
  dP[0] = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) dP[1 + d] = T * dE[1 + d] + T * t.u[d] * dE[DIM + 1];
  dP[DIM + 1] = T * T * dE[DIM + 1];
}
template <typename R, int DIM>
__device__ __forceinline__ void viscous_flux(const GasC<R> &g, const lte::ViscNode<DIM, R> &t, const R (*dP)[DIM + 2],
                                             R *out) {
  R gt[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) gt[d] = dP[d][DIM + 1];
  viscous_flux_core<R, DIM>(t.mu, t.kappa, g.mu_bulk, t.u, gt, dP, out);
}

// LTE: the tabulated gas (periodic meshes); its nodes' temperatures and
// cells come from k_lte_nodes of the same state (FP64)
template <typename R, int DIM, int P, int EPBX = 0, bool BND = false, bool LTE = false>
__global__ void __launch_bounds__((Shape<DIM, P, EPBX>::TPB))
__attribute__((amdgpu_waves_per_eu(GradWaves<R, DIM, P, EPBX, BND, LTE>::value))) k_ns_grad(RHSArgs a) {
  static_assert(!LTE || !BND, "the tabulated gas: periodic meshes");
  using S = Shape<DIM, P, EPBX>;
  constexpr int Np = S::Np, ND = S::NDOF, NF = S::NFP, NQ = S::NEQ, EPB = S::EPB, TPB = S::TPB;
  constexpr int NV = VisComp<DIM>::N;
  constexpr int NVD = DIM + 1;
  constexpr int NFACE = 2 * DIM * NF;
  constexpr int NFR = (EPB * NFACE + TPB - 1) / TPB;
  constexpr int NA = EPB * NQ * ND + EPB * NFACE * NQ;  // entropy variables + face jumps
  constexpr int NB = EPB * DIM * NVD * ND;              // contravariant viscous fluxes
  constexpr int NBUF = NA > NB ? NA : NB;

  __shared__ R sbuf[NBUF];
  __shared__ R sD[Np * Np];
  __shared__ R sDh[Np * Np];
  __shared__ ElemMeta<DIM, EPB> sm;
  __shared__ int sfnb[BND ? EPB : 1][BND ? 2 * DIM : 1];  // face neighbours (boundary codes < 0)
  R (*sS)[NQ][ND] = reinterpret_cast<R (*)[NQ][ND]>(sbuf);
  R (*sjump)[NFACE][NQ] = reinterpret_cast<R (*)[NFACE][NQ]>(sbuf + EPB * NQ * ND);
  R (*sG)[DIM][NVD][ND] = reinterpret_cast<R (*)[DIM][NVD][ND]>(sbuf);

  const GasC<R> &g = gas_of<R>(a);
  R *Vd_out = static_cast<R *>(a.Vd_out);
  R *Gf_out = static_cast<R *>(a.Gf_out);
  const int tid = threadIdx.x;
  const int el = tid / ND;
  const int node = tid - el * ND;
  const bool lane_ok = el < EPB;
  const int elc = lane_ok ? el : 0;
  const int ebase_blk = logical_block(a.xcd_remap) * EPB;
  const int e = ebase_blk + el;
  // serving pool: the case owning this element (free slots are skipped)
  const int ecase = a.elem_case ? a.elem_case[min(ebase_blk + elc, a.ne - 1)] : 0;
  if (EPB == 1 && ecase < 0) return;  // uniform over the workgroup
  const bool valid = lane_ok && e < a.ne && ecase >= 0;
  const long long Ntot = a.Ntot;
  const long long Gtot = (long long)a.ne * NFACE;

  // ---- loads (unconditional, clamped): neighbour indices, own data, then
  //      the neighbours' face traces
  constexpr int NFT = EPB * NFACE;
  int fnb[NFR];
#pragma unroll
  for (int r = 0; r < NFR; ++r) {
    const uint32_t fc = kFaceTable<DIM, P, EPBX>.v[min(tid + r * TPB, NFT - 1)];
    const int pe = (int)((fc >> 24) & 15u), f = (int)(fc >> 28);
    const int ee = min(ebase_blk + pe, a.ne - 1);
    fnb[r] = a.nbr[(size_t)ee * 2 * DIM + f];
    if constexpr (BND) {
      const int fu = tid + r * TPB;  // unclamped: one writer per face
      if (fu < NFT && ((fc >> 16) & 255u) % NF == 0) sfnb[pe][f] = fnb[r];
    }
  }
  MetaRegs<DIM, EPB, TPB> mr;
  meta_load_index<DIM, EPB, TPB>(a, ebase_blk, mr);
  meta_load_own<DIM, EPB, TPB>(a, ebase_blk, mr);
  const size_t ebase = (size_t)(valid ? e : 0) * ND + node;
  double q64[NQ];
#pragma unroll
  for (int c = 0; c < NQ; ++c) q64[c] = a.U[c * Ntot + ebase];
  double Town = 0.0;
  int cown = 0;
  if constexpr (LTE) {
    Town = a.lte.node_T[ebase];
    cown = a.lte.node_cell[ebase];
  }
  static_assert(Np * Np <= TPB, "one basis entry per lane");
  const int kb = min(tid, Np * Np - 1);
  const double bD = a.basis.D[kb], bDh = a.basis.Dhat[kb];
  meta_load_normals<DIM, EPB, TPB>(a, ebase_blk, mr);
  double qn[NFR][NQ];
  double Tn[LTE ? NFR : 1];
  int cn[LTE ? NFR : 1];
#pragma unroll
  for (int r = 0; r < NFR; ++r) {
    const uint32_t fc = kFaceTable<DIM, P, EPBX>.v[min(tid + r * TPB, NFT - 1)];
    const size_t nbase = (size_t)max(fnb[r], 0) * ND + ((fc >> 8) & 255u);
#pragma unroll
    for (int c = 0; c < NQ; ++c) qn[r][c] = a.U[c * Ntot + nbase];
    if constexpr (LTE) {
      Tn[r] = a.lte.node_T[nbase];
      cn[r] = a.lte.node_cell[nbase];
    }
  }
  if (tid < Np * Np) {
    sD[tid] = R(bD);
    sDh[tid] = R(bDh);
  }

  // entropy variables of the element's nodes (EOS.hpp:206-234)
  meta_store<DIM, EPB, TPB>(mr, sm);
  R q[NQ];
#pragma unroll
  for (int c = 0; c < NQ; ++c) q[c] = R(q64[c]);
  if (valid) {
    R Sv[NQ];
    if constexpr (LTE) lte::entropy_state<DIM>(a.lte, q, R(Town), cown, Sv);
    else entropy_state<R, DIM>(g, q, Sv);
#pragma unroll
    for (int c = 0; c < NQ; ++c) sS[el][c][node] = Sv[c];
  }
  lds_sync();

  // canonical interior-face jumps 0.5*(S_R - S_L) (AssembleGradInteriorFacePointKernel);
  // at a boundary point the jump S_bc - S along the outward normal
  // (ComputeBdrFaceGradFlux, bc_kernels.hpp:32-135)
#pragma unroll
  for (int r = 0; r < NFR; ++r) {
    const int fi = tid + r * TPB;
    const uint32_t fc = kFaceTable<DIM, P, EPBX>.v[min(fi, NFT - 1)];
    const int pe = (int)((fc >> 24) & 15u);
    const int fl = (int)((fc >> 16) & 255u);
    if (fi < NFT && ebase_blk + pe < a.ne && (BND || fnb[r] >= 0)) {
      const int s = (int)(fc >> 28) & 1;
      const int on = (int)(fc & 255u);
      if (!BND || fnb[r] >= 0) {
        R qr[NQ], Sn[NQ];
#pragma unroll
        for (int c = 0; c < NQ; ++c) qr[c] = R(qn[r][c]);
        if constexpr (LTE) lte::entropy_state<DIM>(a.lte, qr, R(Tn[LTE ? r : 0]), cn[LTE ? r : 0], Sn);
        else entropy_state<R, DIM>(g, qr, Sn);
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
          const R so = sS[pe][c][on];
          sjump[pe][fl][c] = s ? R(0.5) * (Sn[c] - so) : R(0.5) * (so - Sn[c]);
        }
      } else {
        const int b = -(fnb[r] + 1);
        R So[NQ], jb[NQ];
#pragma unroll
        for (int c = 0; c < NQ; ++c) So[c] = sS[pe][c][on];
        bc_grad_jump<R, DIM>(g, a.bc_type[b], a.bc_data + (size_t)b * BC_STRIDE, So, jb);
#pragma unroll
        for (int c = 0; c < NQ; ++c) sjump[pe][fl][c] = jb[c];
      }
    }
  }
  lds_sync();

  const int i = node % Np;
  const int j = DIM > 1 ? (node / Np) % Np : 0;
  const int k = DIM > 2 ? node / (Np * Np) : 0;
  const int idx[3] = {i, j, k};
  const double Je = sm.J[elc];
  const R invJ = R(1.0 / Je);
  const R wown = R(1.0 / (a.basis.w[0] * Je));

  R vf[NV];
  R hv[NVD];  // the own side's share -0.5 n.Fv of the central viscous face flux, lifted
#pragma unroll
  for (int c = 0; c < NVD; ++c) hv[c] = R(0);
  constexpr bool MF = GradMfma<R, DIM, P, EPBX>::on && !BND;  // with the boundary code it would cost a wave
  // the line contractions D.S of all directions on the matrix cores (every lane of the wave)
  float drefm[MF ? 3 : 1][MF ? 5 : 1];
  float opD[4];  // matrix-core operand: row node%4 of D (then of Dhat)
  if constexpr (MF) {
#pragma unroll
    for (int l = 0; l < 4; ++l) opD[l] = sD[(node & 3) * 4 + l];
    line_contract_mfma5(reinterpret_cast<const float (*)[64]>(sS[elc]), opD, node, drefm);
  }
  if (valid) {
    // volume gradient (AssembleGradVolumePointKernel): acc[dd] = sum_ref (D.S)_ref adj(ref,dd),
    // one reference direction at a time, in the reference's ref = 0,1,2 order
    R acc[DIM][NQ];
#pragma unroll MF ? DIM : 1
    for (int dir = 0; dir < DIM; ++dir) {
      R dref[NQ];
      if constexpr (MF) {
#pragma unroll
        for (int c = 0; c < NQ; ++c) dref[c] = drefm[MF ? dir : 0][MF ? c : 0];
      } else {
#pragma unroll
        for (int c = 0; c < NQ; ++c) dref[c] = R(0);
        const int mi = dir == 0 ? i : (dir == 1 ? j : k);
        const int st = dir == 0 ? 1 : (dir == 1 ? Np : Np * Np);
        const int base = node - mi * st;
#pragma unroll
        for (int l = 0; l < Np; ++l) {
          const R coef = sD[mi * Np + l];
          const int sn_ = base + l * st;
#pragma unroll
          for (int c = 0; c < NQ; ++c) dref[c] += sS[el][c][sn_] * coef;
        }
      }
#pragma unroll
      for (int dd = 0; dd < DIM; ++dd) {
        const R ad = R(sm.adj[el][dir * DIM + dd]);
#pragma unroll
        for (int c = 0; c < NQ; ++c) acc[dd][c] = dir == 0 ? dref[c] * ad : acc[dd][c] + dref[c] * ad;
      }
    }
    using Thermo = std::conditional_t<LTE, lte::ViscNode<DIM, R>, NodeThermo<R, DIM>>;
    Thermo th;
    if constexpr (LTE) th = lte::visc_node<DIM>(a.lte, q, R(Town), cown);
    else th = node_thermo<R, DIM>(g, q);

    // gradient = volume part / J + w * jump * n from the face through the
    // node in each direction (a node lies on at most one face per direction)
#pragma unroll
    for (int dd = 0; dd < DIM; ++dd)
#pragma unroll
      for (int c = 0; c < NQ; ++c) acc[dd][c] *= invJ;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const bool hi = idx[d] == Np - 1;
      const bool on = hi || idx[d] == 0;
      const int f = 2 * d + (hi ? 1 : 0);
      const int row = f * NF + face_index<DIM, Np>(d, i, j, k);
      R jw[NQ];
#pragma unroll
      for (int c = 0; c < NQ; ++c) {
        const R v = wown * sjump[el][row][c];
        jw[c] = on ? v : R(0);
      }
#pragma unroll
      for (int dd = 0; dd < DIM; ++dd) {
        const R nd = R(sm.nrm[el][f][dd]);
#pragma unroll
        for (int c = 0; c < NQ; ++c) acc[dd][c] += jw[c] * nd;
      }
    }
    R dP[DIM][NQ];
#pragma unroll
    for (int dd = 0; dd < DIM; ++dd) grad_entropy_to_grad_prim<R, DIM>(th, q, acc[dd], dP[dd]);
    viscous_flux<R, DIM>(g, th, dP, vf);

    // n.Fv at the element's own face points, n the canonical area vector
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const bool hi = idx[d] == Np - 1;
      if (hi || idx[d] == 0) {
        const int f = 2 * d + (hi ? 1 : 0);
        const size_t gi = (size_t)e * NFACE + f * NF + face_index<DIM, Np>(d, i, j, k);
        R n[DIM];
#pragma unroll
        for (int dd = 0; dd < DIM; ++dd) n[dd] = R(sm.nrm[el][f][dd]);
        R vn[NVD];
#pragma unroll
        for (int c = 0; c < NVD; ++c) {
          R v = n[0] * vf[vis_index<DIM>(1 + c, 0)];
#pragma unroll
          for (int dd = 1; dd < DIM; ++dd) v += n[dd] * vf[vis_index<DIM>(1 + c, dd)];
          Gf_out[c * Gtot + gi] = v;
          vn[c] = v;
        }
        const int fc = BND ? sfnb[el][f] : 0;
        if (!BND || fc >= 0) {
          // face term scale: -w on the element's high face, +w on its low face
#pragma unroll
          for (int c = 0; c < NVD; ++c) hv[c] += (hi ? wown : -wown) * (R(0.5) * vn[c]);
        } else {
          // the boundary's viscous share, -w times the viscous part of its
          // flux (n is the outward area vector here)
          const int b = -(fc + 1);
          R bs[NVD];
          bc_viscous_share<R, DIM>(a.bc_type[b], a.bc_data + (size_t)b * BC_STRIDE, vn, th.u, n, bs);
#pragma unroll
          for (int c = 0; c < NVD; ++c) hv[c] += wown * bs[c];
        }
      }
    }
  }
  lds_sync();  // the entropy/jump buffers are reused below

  // contravariant viscous fluxes adj_dir . Fv per node
  if (valid) {
#pragma unroll
    for (int dir = 0; dir < DIM; ++dir) {
      R row[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) row[d] = R(sm.adj[el][dir * DIM + d]);
#pragma unroll
      for (int c = 0; c < NVD; ++c) {
        R fr = row[0] * vf[vis_index<DIM>(1 + c, 0)];
#pragma unroll
        for (int ph = 1; ph < DIM; ++ph) fr += row[ph] * vf[vis_index<DIM>(1 + c, ph)];
        sG[el][dir][c][node] = fr;
      }
    }
  }
  lds_sync();
  // the divergence's line contractions on the matrix cores, one direction at
  // a time (then summed in x, y, z order)
  float dvm[MF ? 3 : 1][MF ? 4 : 1];
  if constexpr (MF) {
#pragma unroll
    for (int l = 0; l < 4; ++l) opD[l] = sDh[(node & 3) * 4 + l];
    const float(*G)[4][64] = reinterpret_cast<const float(*)[4][64]>(sG[elc]);
    float gx[4], gy[4], gz[4];
    line_contract_dir<0>(G[0], opD, node, gx);
    line_contract_dir<1>(G[1], opD, node, gy);
    line_contract_dir<2>(G[2], opD, node, gz);
#pragma unroll
    for (int c = 0; c < 4; ++c) dvm[0][c] = gx[c] + gy[c] + gz[c];
  }
  if (valid) {
    R dv[NVD];
    if constexpr (MF) {
#pragma unroll
      for (int c = 0; c < NVD; ++c) dv[c] = dvm[0][MF ? c : 0];
    } else {
#pragma unroll
      for (int c = 0; c < NVD; ++c) dv[c] = R(0);
#pragma unroll
      for (int dir = 0; dir < DIM; ++dir) {
        const int mi = dir == 0 ? i : (dir == 1 ? j : k);
        const int st = dir == 0 ? 1 : (dir == 1 ? Np : Np * Np);
        const int base = node - mi * st;
#pragma unroll
        for (int l = 0; l < Np; ++l) {
          const R coef = sDh[mi * Np + l];
#pragma unroll
          for (int c = 0; c < NVD; ++c) dv[c] += coef * sG[el][dir][c][base + l * st];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < NVD; ++c) Vd_out[c * Ntot + ebase] = invJ * dv[c] + hv[c];
  }
}

// ---------------------------------------------------------------------------
// Stability estimate (RHSOperator_impl.hpp:10-270): per-node and per-face
// rates reduced on the device; non-negative doubles are max-reduced through
// their bit patterns with 64-bit integer atomics.
// ---------------------------------------------------------------------------
// Stability rates buffer: [0..2] the advective, surface and diffusive rate
// (bit patterns of non-negative doubles), then kRateSlots partial maxima of
// each, zeroed before and folded after each estimate.
constexpr int kRateSlots = 64;
constexpr int kRateWords = 3 + 3 * kRateSlots;

struct StabArgs {
  const double *U;
  unsigned long long *rates;  // kRateWords: [3] results, then the slots (bit patterns)
  long long Ntot;
  int ne;             // elements estimated: [e0, e0 + ne)
  int e0;
  int viscous;
  double adv_scale, diff_scale, surf_scale;
  const double *J;
  const double *adj;
  const int *nbr;
  const int *bc_type;     // [nbc]
  const double *bc_data;  // [nbc][BC_STRIDE]
  GasC<double> gas;
  double w0;
  lte::View lte;  // tabulated LTE gas: sound speed from the table when lte.nx > 0
  int lte_nodes;  // the state's temperatures and cells are in lte.node_T / node_cell (k_lte_nodes)
};

__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

template <int DIM, int P>
__global__ void __launch_bounds__((Shape<DIM, P>::TPB)) k_stability(StabArgs a) {
  using S = Shape<DIM, P>;
  constexpr int Np = S::Np, ND = S::NDOF, NF = S::NFP, NQ = S::NEQ, EPB = S::EPB;
  const GasC<double> &g = a.gas;
  const int tid = threadIdx.x;
  const int el = tid / ND;
  const int node = tid - el * ND;
  const int e = a.e0 + blockIdx.x * EPB + el;
  const bool valid = el < EPB && e < a.e0 + a.ne;
  double adv = 0.0, surf = 0.0, diff = 0.0;
  if (valid) {
    const long long Ntot = a.Ntot;
    Pt<double, DIM> s;
    const size_t ebase = (size_t)e * ND + node;
#pragma unroll
    for (int c = 0; c < NQ; ++c) s.q[c] = a.U[c * Ntot + ebase];
    derive(g, s);
    const bool tab = a.lte.nx > 0;
    lte::Eval<double> ev;
    if (tab)
      ev = a.lte_nodes ? lte::at_packed<double>(a.lte, s.q[0], a.lte.node_T[ebase], a.lte.node_cell[ebase])
                       : lte::state_eval<DIM>(a.lte, s.q);
    const double c_s = tab ? sqrt(ev.c2()) : sound_speed(g, s);
    const double invJ = 1.0 / a.J[e];
    const double *adj_e = a.adj + (size_t)e * DIM * DIM;
    double dsum = 0.0, msum = 0.0;
#pragma unroll
    for (int rd = 0; rd < DIM; ++rd) {
      double vdm = 0.0, mn2 = 0.0;
#pragma unroll
      for (int pd = 0; pd < DIM; ++pd) {
        const double val = adj_e[rd * DIM + pd];
        vdm += s.v[pd] * val;
        mn2 += val * val;
      }
      dsum += (fabs(vdm) + c_s * sqrt(mn2)) * invJ;
      msum += mn2 * invJ * invJ;
    }
    adv = a.adv_scale * dsum;
    if (a.viscous) {
      // StabilityEstimate.hpp:23-38; the tabulated gas takes viscosity and
      // conductivity from the table and gamma = cp/cv from the interpolant
      const double rho = s.q[0];
      double shear = g.mu, kappa = g.mu * g.cp * g.PrInverse, gam = g.gamma, cp = g.cp;
      if (tab) {
        shear = a.lte.sutherland ? lte::sutherland_viscosity(a.lte, ev.T) : ev.transport<4>();
        kappa = ev.transport<5>();
        cp = ev.cp();
        gam = cp / ev.cv();
      }
      const double longitudinal = (2.0 - g.mu_bulk) * shear;
      const double md = fmax(shear, longitudinal) / rho;
      const double td = kappa * gam / (rho * cp);
      diff = a.diff_scale * fmax(md, td) * msum;
    }
    // face rates: each node on a face takes the max over the faces it touches
    const int i = node % Np;
    const int j = DIM > 1 ? (node / Np) % Np : 0;
    const int k = DIM > 2 ? node / (Np * Np) : 0;
    const int idx[3] = {i, j, k};
    const double wown = 1.0 / (a.w0 * a.J[e]);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
      for (int sd = 0; sd < 2; ++sd) {
        if (idx[d] != (sd ? Np - 1 : 0)) continue;
        const int nb = a.nbr[(size_t)e * 2 * DIM + 2 * d + sd];
        double n[DIM];
        const double *adj_n = (nb >= 0 && !sd) ? a.adj + (size_t)nb * DIM * DIM : adj_e;
        double n2 = 0.0, vn = 0.0;
#pragma unroll
        for (int dd = 0; dd < DIM; ++dd) {
          n[dd] = (nb < 0 && !sd) ? -adj_e[d * DIM + dd] : adj_n[d * DIM + dd];
          n2 += n[dd] * n[dd];
          vn += s.v[dd] * n[dd];
        }
        const double nm = sqrt(n2);
        double ws = fabs(vn) + c_s * nm;
        double wmax = fabs(wown);
        if (nb >= 0) {
          const int t = face_index<DIM, Np>(d, i, j, k);
          const int nn = face_node<DIM, Np>(d, 1 - sd, t);
          Pt<double, DIM> o;
#pragma unroll
          for (int c = 0; c < NQ; ++c) o.q[c] = a.U[c * Ntot + (size_t)nb * ND + nn];
          derive(g, o);
          double vo = 0.0;
#pragma unroll
          for (int dd = 0; dd < DIM; ++dd) vo += o.v[dd] * n[dd];
          double c_o;
          if (a.lte.nx > 0) {
            const size_t nbase = (size_t)nb * ND + nn;
            const lte::Eval<double> eo =
                a.lte_nodes ? lte::at_packed<double>(a.lte, o.q[0], a.lte.node_T[nbase], a.lte.node_cell[nbase])
                            : lte::state_eval<DIM>(a.lte, o.q);
            c_o = sqrt(eo.c2());
          } else {
            c_o = sound_speed(g, o);
          }
          ws = fmax(ws, fabs(vo) + c_o * nm);
          wmax = fmax(wmax, fabs(1.0 / (a.w0 * a.J[nb])));
        } else {
          // an inflow or prescribed exterior state adds its wave speed and,
          // when viscous, its diffusivity with the element's metric sum
          // (RHSOperator_impl.hpp:176-267)
          const int b = -(nb + 1);
          const int bt = a.bc_type[b];
          if (bt == BC_SUPERSONIC_INFLOW || bt == BC_PRESCRIBED_STATE) {
            Pt<double, DIM> o;
#pragma unroll
            for (int c = 0; c < NQ; ++c) o.q[c] = a.bc_data[(size_t)b * BC_STRIDE + c];
            derive(g, o);
            double vo = 0.0;
#pragma unroll
            for (int dd = 0; dd < DIM; ++dd) vo += o.v[dd] * n[dd];
            ws = fmax(ws, fabs(vo) + sound_speed(g, o) * nm);
            if (a.viscous) {
              const double rho_o = o.q[0];
              const double longitudinal = (2.0 - g.mu_bulk) * g.mu;
              const double md = fmax(g.mu, longitudinal) / rho_o;
              const double td = g.mu * g.cp * g.PrInverse * g.gamma / (rho_o * g.cp);
              diff = fmax(diff, a.diff_scale * fmax(md, td) * msum);
            }
          }
        }
        surf = fmax(surf, ws * wmax * a.surf_scale);
      }
  }
  adv = wave_max(adv);
  surf = wave_max(surf);
  diff = wave_max(diff);
  if ((threadIdx.x & 63) == 0) {
    // one of kRateSlots slots per rate (by block), so that the grid's
    // atomics do not all queue on three addresses; k_rates_fold takes the max
    unsigned long long *slot = a.rates + 3 + (blockIdx.x % kRateSlots);
    atomicMax(slot, (unsigned long long)__double_as_longlong(adv));
    atomicMax(slot + kRateSlots, (unsigned long long)__double_as_longlong(surf));
    atomicMax(slot + 2 * kRateSlots, (unsigned long long)__double_as_longlong(diff));
  }
}


// ---------------------------------------------------------------------------
// Integral measures (RHSOperator_impl.hpp:530-717): node-quadrature integrals
// of rho, 0.5|m|^2/rho and rhoE (weights J w_i w_j w_k) and the extremes of
// pressure, temperature and density.  Each workgroup writes its partials with
// plain stores; k_measures_final combines them in a fixed order, so the
// result does not depend on scheduling.
// ---------------------------------------------------------------------------
constexpr int NMEAS = 9;  // mass, ke, en, max_p, min_p, max_T, min_T, max_rho, min_rho

struct MeasArgs {
  const double *U;
  double *partial;  // [gridDim.x][NMEAS]
  long long Ntot;
  int ne, e0;       // elements measured: [e0, e0 + ne)
  const double *J;
  GasC<double> gas;
  double w[MAXNP];  // GLL weights on [0,1]
  lte::View lte;
};

__device__ __forceinline__ void meas_identity(double *m) {
  m[0] = m[1] = m[2] = 0.0;
  m[3] = m[5] = m[7] = 0.0;  // maxima start at 0
  m[4] = m[6] = m[8] = 1e32;  // minima start at 1e32
}

__device__ __forceinline__ void meas_combine(double *m, const double *o) {
#pragma unroll
  for (int k = 0; k < 3; ++k) m[k] += o[k];
#pragma unroll
  for (int k = 3; k < NMEAS; k += 2) {
    m[k] = fmax(m[k], o[k]);
    m[k + 1] = fmin(m[k + 1], o[k + 1]);
  }
}

__device__ __forceinline__ void meas_wave(double *m) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double o[NMEAS];
#pragma unroll
    for (int k = 0; k < NMEAS; ++k) o[k] = __shfl_xor(m[k], off, 64);
    meas_combine(m, o);
  }
}

template <int DIM, int P>
__global__ void __launch_bounds__((Shape<DIM, P>::TPB)) k_measures(MeasArgs a) {
  using S = Shape<DIM, P>;
  constexpr int Np = S::Np, ND = S::NDOF, NQ = S::NEQ, EPB = S::EPB, TPB = S::TPB;
  constexpr int NW = (TPB + 63) / 64;
  __shared__ double red[NW][NMEAS];
  const int tid = threadIdx.x;
  const int el = tid / ND;
  const int node = tid - el * ND;
  const int e = a.e0 + blockIdx.x * EPB + el;
  double m[NMEAS];
  meas_identity(m);
  if (el < EPB && e < a.e0 + a.ne) {
    double q[NQ];
    const size_t ebase = (size_t)e * ND + node;
#pragma unroll
    for (int c = 0; c < NQ; ++c) q[c] = a.U[c * a.Ntot + ebase];
    const int i = node % Np;
    const int j = DIM > 1 ? (node / Np) % Np : 0;
    const int k = DIM > 2 ? node / (Np * Np) : 0;
    double w = a.w[i];
    if (DIM > 1) w *= a.w[j];
    if (DIM > 2) w *= a.w[k];
    const double rho = q[0];
    double m2 = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) m2 += q[1 + d] * q[1 + d];
    const double ke = 0.5 * m2 / rho;
    const double rhoE = q[DIM + 1];
    double press, temper;
    if (a.lte.nx > 0) {
      const lte::Eval<double> st = lte::state_eval<DIM>(a.lte, q);
      press = st.pressure();
      temper = st.T;
    } else {
      press = a.gas.gammaM1 * (rhoE - ke);
      temper = press / (rho * a.gas.R_gas);
    }
    const double measure = a.J[e] * w;
    m[0] = rho * measure;
    m[1] = ke * measure;
    m[2] = rhoE * measure;
    m[3] = fmax(0.0, press);
    m[4] = fmin(1e32, press);
    m[5] = fmax(0.0, temper);
    m[6] = fmin(1e32, temper);
    m[7] = fmax(0.0, rho);
    m[8] = fmin(1e32, rho);
  }
  meas_wave(m);
  if ((tid & 63) == 0) {
#pragma unroll
    for (int c = 0; c < NMEAS; ++c) red[tid >> 6][c] = m[c];
  }
  __syncthreads();
  if (tid == 0) {
    double r[NMEAS];
#pragma unroll
    for (int c = 0; c < NMEAS; ++c) r[c] = red[0][c];
    for (int wv = 1; wv < NW; ++wv) meas_combine(r, red[wv]);
#pragma unroll
    for (int c = 0; c < NMEAS; ++c) a.partial[(size_t)blockIdx.x * NMEAS + c] = r[c];
  }
}

}  // namespace thx
