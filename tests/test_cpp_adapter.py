"""The C++ operator adapter (include/theseus_rhs_operator.hpp) compiles
against the C-ABI with the host compiler alone, links the library, and maps
status codes to the reference's exception types (RHSOperator_impl.hpp:291-306)
before any device work: an invalid descriptor raises std::invalid_argument,
an unsupported configuration too, and the last-error text comes through."""
import subprocess
from pathlib import Path

import pytest

from theseus_amd import _abi as A

ROOT = Path(__file__).resolve().parents[1]

SRC = r"""
#include <cstdio>
#include <stdexcept>
#include "theseus_rhs_operator.hpp"

static theseus_case_desc make(int dim, int order, int flow, int precision) {
  static double J[2] = {0.125, 0.125};
  static double adj[2][9] = {{0.25, 0, 0, 0, 0.25, 0, 0, 0, 0.25}, {0.25, 0, 0, 0, 0.25, 0, 0, 0, 0.25}};
  static int nbr[2][6] = {{1, 1, 0, 0, 0, 0}, {0, 0, 1, 1, 1, 1}};
  theseus_case_desc d{};
  d.dim = dim; d.order = order; d.num_elements = 2; d.flow_model = flow;
  d.volume_flux = THESEUS_VOLUME_RANOCHA; d.surface_flux = THESEUS_SURFACE_LLF;
  d.precision = precision; d.blending = 0;
  d.gamma = 1.4; d.Pr = 0.72; d.R_gas = 287.05; d.mu = 0.02; d.mu_bulk = 2.0 / 3.0;
  d.alpha_max = 0.5; d.alpha_min = 0.001;
  d.elem_jac = J; d.elem_adj = &adj[0][0]; d.face_neighbor = &nbr[0][0];
  return d;
}

int main() {
  int seen = 0;
  try {
    theseus::RHSOperator op(make(4, 3, THESEUS_FLOW_EULER, THESEUS_PRECISION_FP64));
  } catch (const std::invalid_argument &e) {
    std::printf("invalid_argument: %s\n", e.what());
    seen |= 1;
  }
  try {
    theseus::RHSOperator op(make(3, 0, THESEUS_FLOW_EULER, THESEUS_PRECISION_FP64));
  } catch (const std::invalid_argument &e) {
    std::printf("invalid_argument: %s\n", e.what());
    seen |= 2;
  }
  try {
    theseus::RHSOperator op(make(3, 3, THESEUS_FLOW_EULER, 7));
  } catch (const std::invalid_argument &e) {
    std::printf("invalid_argument: %s\n", e.what());
    seen |= 4;
  }
  // the tabulated gas: blending is rejected before any device work
  try {
    static double grid[2] = {1.0, 2.0}, planes[24] = {0}, inv[4] = {1.5, 1.5, 1.5, 1.5};
    theseus_lte_table t{};
    t.nx = 2; t.ny = 2; t.rho_grid = grid; t.T_grid = grid; t.e_grid = grid; t.planes = planes; t.inverse = inv;
    theseus_case_desc d = make(3, 3, THESEUS_FLOW_NAVIER_STOKES, THESEUS_PRECISION_FP64);
    d.volume_flux = THESEUS_VOLUME_LTEEC; d.surface_flux = THESEUS_SURFACE_LTEEC; d.blending = 1;
    theseus::RHSOperator op(d, t);
  } catch (const std::invalid_argument &e) {
    std::printf("lte: %s\n", e.what());
    seen |= 256;
  }
  // the measure, device-clock and entropy members compile against the C-ABI (no device work here)
  auto m1 = &theseus::RHSOperator::ComputeIntegralMeasures;
  auto m2 = &theseus::RHSOperator::GetIntegralMeasuresBaseline;
  auto m3 = &theseus::RHSOperator::StepCFL;
  auto m4 = &theseus::RHSOperator::EntropyProduction;
  (void)m1; (void)m2; (void)m3; (void)m4;
  static_assert(sizeof(theseus_clock) == 64, "clock layout");
  static_assert(sizeof(theseus_integral_measures) == 72, "measures layout");
  theseus::StabilityEstimate s;
  s.advective_rate = 2.0; s.surface_rate = 3.0; s.diffusive_rate = 0.5;
  if (s.TotalRate() == 3.5) seen |= 8;
  // deferred creation (the adapter builds its engine in Finalize)
  theseus::RHSOperator later;
  if (!later) seen |= 16;
  // orientation helpers: two quads sharing an edge, the second stored rotated
  // by a quarter turn; aligned, their shared face is (0, 1) of the first
  const long long vids[8] = {0, 1, 3, 4, 2, 5, 1, 4};
  const double xyz[16] = {0, 0, 1, 0, 0, 1, 1, 1, 2, 0, 2, 1, 1, 0, 1, 1};
  unsigned char code[2];
  int nbr[8];
  double jac[2], adj[8];
  theseus::OrientElements(2, 2, vids, xyz, code, nbr, jac, adj);
  if (code[0] == 0 && code[1] != 0 && nbr[1] == 1 && nbr[4 + 0] == 0 && jac[0] == 1.0 && jac[1] == 1.0) seen |= 32;
  if (theseus::StorageCorner(2, 0, 3) == 3) seen |= 64;
  try {
    theseus::OrientElements(4, 2, vids, xyz, code, nbr, jac, adj);
  } catch (const std::invalid_argument &) {
    seen |= 128;
  }
  std::printf("seen=%d code=%d,%d nbr=%d,%d,%d,%d|%d,%d,%d,%d\n", seen, code[0], code[1], nbr[0], nbr[1], nbr[2], nbr[3],
              nbr[4], nbr[5], nbr[6], nbr[7]);
  return seen == 511 ? 0 : 1;
}
"""


def test_cpp_adapter_compiles_links_and_raises(tmp_path):
    lib = A.load()  # noqa: F841  (the library must be built)
    src = tmp_path / "adapter.cpp"
    src.write_text(SRC)
    exe = tmp_path / "adapter"
    libdir = A.LIB_PATH.parent
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT / 'include'}", str(src), f"-L{libdir}", "-ltheseus_rhs",
                    f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.count("invalid_argument:") == 3
    assert "dim" in out.stdout and "order" in out.stdout
    assert "lte: " in out.stdout and "blending" in out.stdout
