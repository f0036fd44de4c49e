// Euler right-hand side of the tabulated LTE gas (EulerOperator_impl.hpp:
// 28-421 with the LTEHelmholtzGas of LTEGasModel.hpp:339), FP64, with one of
// the flux pairings the reference's factory builds for that gas
// (SimFactory_impl.hpp:140-245; FX, LteFlux): LTEEC (LTEECFlux.hpp:93-107),
// Pirozzoli (PirozzoliFlux.hpp:19-120), or the central volume flux with an
// HLL (HLLFlux.hpp:19-116) or LLF (LaxFriedrichsFlux.hpp:20-77) face; the
// last three take the gas's pressure, sound speed and cp/cv at each node
// from the table at its recovered temperature.  Same element-per-
// workgroup structure as k_rhs: the element's temperatures are recovered
// once per node per stage by k_lte_nodes (inverse-table guess + Newton)
// and staged in LDS with the table cell, so the pair fluxes only evaluate
// the interpolant; each GLL-line pair is evaluated once; the neighbours' face
// traces come with their recovered temperatures and cells.  With VISC it is
// the Navier-Stokes pass 2 of the tabulated gas (periodic meshes): the
// viscous volume divergence with the own half of the central viscous face
// flux, and the neighbours' n.Fv, come from k_ns_grad<..., LTE> as for the
// ideal gas (DGSEMIntegrator.hpp:261-316, 486-622).
//
// R = float is the mixed-precision form: the staged states and the pair and
// face fluxes (the Itoh-Abe means of the interpolant) in FP32; each node's
// temperature (Newton, k_lte_nodes), pressure and sound speed in FP64 once
// per node, the momentum flux's pressure taken less a constant background
// pressure (formed in FP64, as for the ideal gas), and the stage
// combination in FP64 from the FP64 state.
#pragma once

#include "rhs_kernels.hpp"

namespace thx {

// Flux pairings of the tabulated gas (SimFactory_impl.hpp:140-245)
enum LteFlux { LTE_FLUX_LTEEC = 0, LTE_FLUX_PIROZZOLI = 1, LTE_FLUX_HLL = 2, LTE_FLUX_LLF = 3, LTE_NFLUX = 4 };

template <typename R, int DIM, int P, bool VISC = false, int FX = LTE_FLUX_LTEEC>
__global__ void __launch_bounds__((Shape<DIM, P>::TPB)) k_rhs_lte(RHSArgs a) {
  using S = Shape<DIM, P>;
  using SC = RhsScratch<R, DIM, P>;
  constexpr bool FP64 = sizeof(R) == sizeof(double);
  constexpr int Np = S::Np, ND = S::NDOF, NF = S::NFP, NQ = S::NEQ, EPB = S::EPB, TPB = S::TPB;
  constexpr int NFACE = 2 * DIM * NF;
  constexpr int NFR = (EPB * NFACE + TPB - 1) / TPB;
  constexpr int NPL = SC::NPL, NPAIR = SC::NPAIR;
  constexpr int NVD = DIM + 1;  // viscous terms: momentum and energy

  constexpr int SCR = FP64 ? SC::DOUBLES : SC::ROWS * NQ;  // flux rows, in R

  __shared__ R sq[EPB][NQ][ND];
  __shared__ R sT[EPB][ND];     // recovered temperature
  __shared__ R sp[EPB][ND];     // pressure
  __shared__ R sc[EPB][ND];     // sound speed
  __shared__ R spd[FP64 ? 1 : EPB][FP64 ? 1 : ND];  // mixed: pressure less the background
  __shared__ int scell[EPB][ND];     // table cell j*nx + i
  constexpr bool GAM = FX == LTE_FLUX_PIROZZOLI;  // the Roe dissipation's gamma = cp/cv
  constexpr bool EC0 = FX == LTE_FLUX_LTEEC;
  __shared__ R sg[GAM ? EPB : 1][GAM ? ND : 1];
  __shared__ R sraw[EPB][SCR];
  __shared__ R sD2[Np * Np];
  __shared__ ElemMeta<DIM, EPB> sm;
  auto sscr = [&](int pe, int row, int q) -> R & { return sraw[pe][row * NQ + q]; };
  auto spm = [&](int pe, int n) -> R {  // the momentum flux's pressure
    if constexpr (FP64) return sp[pe][n];
    else return spd[pe][n];
  };

  // a stage of a variable step of length 0 (end time reached) does nothing
  if (a.mode != 0 && a.case_dt && a.case_dt[0] == 0.0) return;
  const lte::View &tv = a.lte;
  const int tid = threadIdx.x;
  const int el = tid / ND;
  const int node = tid - el * ND;
  const bool lane_ok = el < EPB;
  const int elc = lane_ok ? el : 0;
  const int ebase_blk = logical_block(a.xcd_remap) * EPB;
  const int e = ebase_blk + el;
  const bool valid = lane_ok && e < a.ne;
  const long long Ntot = a.Ntot;
  const long long Gtot = (long long)a.ne * NFACE;
  const R *Vd = static_cast<const R *>(a.Vd);
  const R *Gf = static_cast<const R *>(a.Gf);
  const double pref = (FP64 || !a.pref) ? 0.0 : a.pref[0];

  // ---- loads (unconditional, clamped), as in k_rhs
  constexpr int NFT = EPB * NFACE;
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
#pragma unroll
  for (int q = 0; q < NQ; ++q) qown[q] = a.U[q * Ntot + ebase];
  static_assert(Np * Np <= TPB, "one basis entry per lane");
  const double bD2 = a.basis.Dhat2[min(tid, Np * Np - 1)];
  meta_load_normals<DIM, EPB, TPB>(a, ebase_blk, mr);
  const double Town = tv.node_T[ebase];
  const int cown = tv.node_cell[ebase];
  if (tid < Np * Np) sD2[tid] = R(bD2);

  // ---- stage: each node's temperature and cell, recovered once per node by
  //      k_lte_nodes; pressure and sound speed evaluated in FP64 (the
  //      other pairs: evaluated by k_lte_nodes too, read here)
  meta_store<DIM, EPB, TPB>(mr, sm);
  if (valid) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) sq[el][q][node] = R(qown[q]);
    if constexpr (FX == LTE_FLUX_LTEEC) {
      const lte::Eval<double> ev = lte::at_packed<double>(tv, qown[0], Town, cown);
      sT[el][node] = R(ev.T);
      const double pn = ev.pressure();
      sp[el][node] = R(pn);
      if constexpr (!FP64) spd[el][node] = R(pn - pref);
      sc[el][node] = R(sqrt(ev.c2()));
      scell[el][node] = (int)(ev.cell.H - tv.planes);
    } else {
      const double pn = tv.node_p[ebase];
      sp[el][node] = R(pn);
      if constexpr (!FP64) spd[el][node] = R(pn - pref);
      sc[el][node] = R(tv.node_c[ebase]);
      if constexpr (GAM) sg[el][node] = R(tv.node_g[ebase]);
    }
  }
  lds_sync();

  const int i = node % Np;
  const int j = DIM > 1 ? (node / Np) % Np : 0;
  const int k = DIM > 2 ? node / (Np * Np) : 0;
  const int idx[3] = {i, j, k};
  const double Je = sm.J[elc];

  // a staged node's evaluation (cell from LDS, no search)
  auto node_eval = [&](int pe, int n) -> lte::Eval<R> {
    const int off = scell[pe][n];
    const int ci = off % tv.nx, cj = off / tv.nx;
    return lte::at_cell<R>(tv, sq[pe][0][n], sT[pe][n], ci, cj);
  };
  // the ideal-gas flux point of a state with the table's pressure p (pd:
  // less the background in the mixed form)
  auto make_pt = [&](const R *q, R p, R pd) {
    Pt<R, DIM> s;
#pragma unroll
    for (int e2 = 0; e2 < NQ; ++e2) s.q[e2] = q[e2];
    s.ir = drcp(q[0]);
    s.p = p;
    s.pd = pd;
#pragma unroll
    for (int d = 0; d < DIM; ++d) s.v[d] = q[1 + d] * s.ir;
    return s;
  };

  // ---- volume: every GLL-line pair once (DGSEMIntegrator.hpp:124-218)
  constexpr int NPT = EPB * NPAIR;
  constexpr int NRND = (DIM * NPT + TPB - 1) / TPB;
  static_assert(DIM < 3 || (2 * NPT) / TPB > (NPT - 1) / TPB, "slot reuse needs a round boundary");
  R rate[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) rate[q] = R(0);
#pragma unroll 1
  for (int rr = 0; rr < NRND; ++rr) {
    const int w = rr * TPB + tid;
    if (w < DIM * NPT) {
      const int dir = w / NPT;
      const int rem = w - dir * NPT;
      const int pe = rem / NPAIR;
      const int r = rem - pe * NPAIR;
      if (ebase_blk + pe < a.ne) {
        const int sdir = dir == 0 ? 1 : (dir == 1 ? Np : Np * Np);
        const int line = r / NPL;
        int pa, pb;
        pair_nodes<Np>(r - line * NPL, pa, pb);
        const int base = face_node<DIM, Np>(dir, 0, line);
        const int na = base + pa * sdir, nb = base + pb * sdir;
        R q1[NQ], q2[NQ], met[DIM];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          q1[q] = sq[pe][q][na];
          q2[q] = sq[pe][q][nb];
        }
#pragma unroll
        for (int d = 0; d < DIM; ++d) met[d] = R(sm.adj[pe][dir * DIM + d]);
        R F[NQ];
        if constexpr (FX == LTE_FLUX_LTEEC) {
          const lte::Eval<R> s1 = node_eval(pe, na), s2 = node_eval(pe, nb);
          lte::lteec_flux<DIM, false>(tv, q1, q2, s1, s2, sp[pe][na], sp[pe][nb], R(0), R(0), met, F, spm(pe, na),
                                      spm(pe, nb));
        } else {
          const Pt<R, DIM> A = make_pt(q1, sp[pe][na], spm(pe, na)), B = make_pt(q2, sp[pe][nb], spm(pe, nb));
          if constexpr (FX == LTE_FLUX_PIROZZOLI) vol_pirozzoli<R, DIM>(GasC<R>{}, A, B, met, F);
          else vol_central<R, DIM>(GasC<R>{}, A, B, met, F);
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) sscr(pe, (dir & 1) * NPAIR + r, q) = F[q];
      }
    }
#pragma unroll
    for (int dir = 0; dir < DIM; ++dir) {
      if (((dir + 1) * NPT - 1) / TPB != rr) continue;
      lds_sync();
      if (valid) {
        const int mi = dir == 0 ? i : (dir == 1 ? j : k);
        const int line = face_index<DIM, Np>(dir, i, j, k);
#pragma unroll
        for (int m = 0; m < Np; ++m) {
          if (m == mi) continue;
          const int pr = m < mi ? pair_index<Np>(m, mi) : pair_index<Np>(mi, m);
          const R coef = sD2[mi * Np + m];
#pragma unroll
          for (int q = 0; q < NQ; ++q) rate[q] += coef * sscr(el, (dir & 1) * NPAIR + line * NPL + pr, q);
        }
      }
      lds_sync();
    }
  }
  R du[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) du[q] = R(-1.0 / Je) * rate[q];

  // the neighbours' face traces (state, temperature, cell) and the RK base
  // state, loaded once the pair fluxes no longer hold registers
  double qn[NFR][NQ], Tn[NFR];  // Tn: the other pairs' pressure
  double pcn[EC0 ? 1 : NFR], gcn[GAM ? NFR : 1];
  int cn[NFR];
  R gn[NFR][VISC ? NVD : 1];
#pragma unroll
  for (int r = 0; r < NFR; ++r) {
    const int fi = min(tid + r * TPB, NFT - 1);
    const int pe = fi / NFACE;
    const int fl = fi - pe * NFACE;
    const int f = fl / NF, t = fl - f * NF;
    const int d = f >> 1, s = f & 1;
    const size_t nbase = (size_t)max(fnb[r], 0) * ND + face_node<DIM, Np>(d, 1 - s, t);
#pragma unroll
    for (int q = 0; q < NQ; ++q) qn[r][q] = a.U[q * Ntot + nbase];
    if constexpr (FX == LTE_FLUX_LTEEC) {
      Tn[r] = tv.node_T[nbase];
      cn[r] = tv.node_cell[nbase];
    } else {  // the neighbour's pressure, sound speed (and gamma) from k_lte_nodes
      Tn[r] = tv.node_p[nbase];
      cn[r] = 0;
      pcn[r] = tv.node_c[nbase];
      if constexpr (GAM) gcn[r] = tv.node_g[nbase];
    }
    if constexpr (VISC) {
      // the neighbour's n.Fv at this point (the own side's half is in Vd)
      const size_t gb_n = (size_t)max(fnb[r], 0) * NFACE + (f ^ 1) * NF + t;
#pragma unroll
      for (int c = 0; c < NVD; ++c) gn[r][c] = Gf[c * Gtot + gb_n];
    }
  }
  double xb[NQ];
  {
    const double *Xp = a.mode == 2 ? a.X : a.U;
#pragma unroll
    for (int q = 0; q < NQ; ++q) xb[q] = Xp[q * Ntot + ebase];
  }

  // ---- faces (EulerOperator_impl.hpp:211-392)
  {
    const double w0 = a.basis.w[0];
#pragma unroll
    for (int r = 0; r < NFR; ++r) {
      const int fi = tid + r * TPB;
      const int pe = fi / NFACE;
      const int fl = fi - pe * NFACE;
      if (fi < EPB * NFACE && ebase_blk + pe < a.ne) {
        const int f = fl / NF, t = fl - f * NF;
        const int d = f >> 1, s = f & 1;
        const int on = face_node<DIM, Np>(d, s, t);
        const bool interior = fnb[r] >= 0;
        R qo[NQ], qt[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          qo[q] = sq[pe][q][on];
          qt[q] = interior ? R(qn[r][q]) : qo[q];
        }
        constexpr bool EC = FX == LTE_FLUX_LTEEC;
        const lte::Eval<R> so = EC ? node_eval(pe, on) : lte::Eval<R>{};
        const R po = sp[pe][on], co = sc[pe][on], pdo = spm(pe, on);
        R go = R(0);
        if constexpr (GAM) go = sg[pe][on];
        lte::Eval<R> st = so;
        R pt = po, ct = co, pdt = pdo, gt = go;
        if (interior) {
          // the neighbour's trace: pressure and sound speed in FP64
          if constexpr (EC) {
            const lte::Eval<double> st64 = lte::at_packed<double>(tv, qn[r][0], Tn[r], cn[r]);
            if constexpr (FP64) st = st64;
            else st = lte::at_packed<R>(tv, qt[0], R(Tn[r]), cn[r]);
            const double pn = st64.pressure();
            pt = R(pn);
            pdt = FP64 ? pt : R(pn - pref);
            ct = R(sqrt(st64.c2()));
          } else {
            const double pn = Tn[r];
            pt = R(pn);
            pdt = FP64 ? pt : R(pn - pref);
            ct = R(pcn[r]);
            if constexpr (GAM) gt = R(gcn[r]);
          }
        }
        R n[DIM];
#pragma unroll
        for (int dd = 0; dd < DIM; ++dd) n[dd] = R(sm.nrm[pe][f][dd]);
        // canonical order (the low-side element first) by selects: one
        // evaluation per lane whichever side of the element it is on
        const bool own_low = s != 0;
        R ql[NQ], qh[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          ql[q] = own_low ? qo[q] : qt[q];
          qh[q] = own_low ? qt[q] : qo[q];
        }
        R F[NQ];
        if constexpr (EC) {
          const lte::Eval<R> sl = own_low ? so : st, sh = own_low ? st : so;
          lte::lteec_flux<DIM, true>(tv, ql, qh, sl, sh, own_low ? po : pt, own_low ? pt : po, own_low ? co : ct,
                                     own_low ? ct : co, n, F, own_low ? pdo : pdt, own_low ? pdt : pdo);
        } else {
          const Pt<R, DIM> L = make_pt(ql, own_low ? po : pt, own_low ? pdo : pdt);
          const Pt<R, DIM> H = make_pt(qh, own_low ? pt : po, own_low ? pdt : pdo);
          const R cl = own_low ? co : ct, ch = own_low ? ct : co;
          auto cs = [&](const Pt<R, DIM> &, int side) { return side ? ch : cl; };
          if constexpr (FX == LTE_FLUX_HLL) face_hll_core(L, H, cs, n, F);
          else if constexpr (FX == LTE_FLUX_LLF) face_llf_core(L, H, cs, n, F);
          else {
            const R gl = own_low ? go : gt;  // gamma of the first state (DissipationOperator.hpp:109)
            face_pirozzoli_core(L, H, [&] { return gl; }, cs, n, F);
          }
        }
        if constexpr (VISC) {
          // F* - 0.5 n.(Fv_L + Fv_R); the own side's -0.5 n.Fv is in Vd
#pragma unroll
          for (int c = 0; c < NVD; ++c) F[1 + c] -= interior ? R(0.5) * gn[r][c] : R(0);
        }
        const double wown = 1.0 / (w0 * sm.J[pe]);
        const R cs = R((interior && !s) ? wown : -wown);
#pragma unroll
        for (int q = 0; q < NQ; ++q) sscr(pe, fl, q) = cs * F[q];
      }
    }
  }
  lds_sync();
  R fsum[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) fsum[q] = R(0);
  if constexpr (VISC) {
    // the viscous volume divergence with the own half of the face term
#pragma unroll
    for (int c = 0; c < NVD; ++c) fsum[1 + c] = Vd[c * Ntot + ebase];
  }
  if (valid) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const bool hi = idx[d] == Np - 1;
      const bool on = hi || idx[d] == 0;
      const int row = (2 * d + (hi ? 1 : 0)) * NF + face_index<DIM, Np>(d, i, j, k);
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const R v = sscr(el, row, q);
        fsum[q] += on ? v : R(0);
      }
    }
    const double dt = a.case_dt ? a.case_dt[0] : a.dt;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const double r = FP64 ? double(du[q] + fsum[q]) : double(du[q]) + double(fsum[q]);
      // the stage combination from the FP64 state (its LDS copy is FP32 in the mixed form)
      const double qo = FP64 ? double(sq[el][q][node]) : a.U[q * Ntot + ebase];
      double o;
      if (a.mode == 0) o = r;
      else if (a.mode == 1) o = qo + dt * r;
      else o = a.cx * xb[q] + a.cy * (qo + dt * r);
      a.out[q * Ntot + ebase] = o;
    }
  }
}

}  // namespace thx
