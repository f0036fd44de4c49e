#!/usr/bin/env python3
"""Benchmark: DOF*RHS-evals/s of the DGSEM engine on the 3-D compressible
Navier-Stokes Taylor-Green vortex, 32^3 hexes, p=3 (10,485,760 DOF), BR1
viscous terms, Ranocha volume + Lax-Friedrichs face flux, FP64.

A "step" is one SSP-RK3 time step (3 RHS evaluations with the stage
updates fused into the RHS kernels) over the whole mesh, inputs resident in
HBM.  ``value`` = DOF * 3 * steps / elapsed.  One JSON line on stdout.

  python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBPS = 8000.0  # MI355X HBM3E


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# Algorithmic bytes per DOF of one NS-BR1 FP64 RHS split over the two
# passes, the two-pass minimum: pass 1 reads U (8) and writes the gradient
# (24); pass 2 reads U (8) and the gradient (24) and writes dU (8).
ALG_B_PER_DOF = {0: 40.0, 1: 32.0}


def algorithmic_bytes(mesh_ne, order, dim, kind):
    """Algorithmic bytes per launch of one kernel kind (pass 2 = 0, pass 1 = 1):
    the per-DOF figure above times the DOFs one launch processes."""
    dofs = mesh_ne * (order + 1) ** dim * (dim + 2)
    return dofs * ALG_B_PER_DOF[kind], ALG_B_PER_DOF[kind]


def moved_bytes_per_dof(order, dim, kind, stage_modes):
    """This engine's own compulsory bytes per DOF, averaged over a step's
    launches.  Per node: the conservative state (dim+2 doubles), the viscous
    volume divergence (dim+1), n.Fv at the element's face points (dim+1 per
    face point, 2*dim*(p+1)^(dim-1) face points per element), the output
    state and, for SSP-RK3 stages 2 and 3, the step's base state."""
    ndof = (order + 1) ** dim
    neq = dim + 2
    nvd = dim + 1
    nfp = 2 * dim * (order + 1) ** (dim - 1)
    visc = nvd + nvd * nfp / ndof
    if kind == 1:
        return 8 * (neq + visc) / neq
    per = [8 * (neq + visc + neq + (neq if mode == 2 else 0)) for mode in stage_modes]
    return sum(per) / len(per) / neq


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(rt, mesh, U0, seconds):
    """The CPU FP64 oracle (a restatement of the reference's RHS path) on
    the host cores: SSP-RK3 steps of the workload, each trial timed with a
    wall clock after an untimed step, the median of 5 trials reported, on all
    host cores (OpenMP over elements) and on one core.  The all-core trials
    run the full grid; the one-core trials a 12^3 grid of the same physics
    (the cost per DOF does not depend on the grid size) so that the whole
    leg stays near `seconds`."""
    import statistics

    from oracle import pyoracle
    from theseus_amd.cases import c2_tgv3d
    from theseus_amd.operator import build_descriptor

    if not pyoracle.LIB.exists():
        pyoracle.build()
    lib = pyoracle.load()

    def trials(rt_, mesh_, U_, n_trials=5):
        desc, keep = build_descriptor(rt_, mesh_)
        case = pyoracle.OracleCase(desc)
        rates = case.stability(U_)
        dt = 0.5 / (max(rates[0], rates[1]) + rates[2])
        U = case.step(U_.copy(), dt)  # untimed
        t0 = time.perf_counter()
        U = case.step(U, dt)
        one = time.perf_counter() - t0
        steps = max(1, int(seconds / 2 / n_trials / max(one, 1e-9)))
        rates_ = []
        for _ in range(n_trials):
            t0 = time.perf_counter()
            for _ in range(steps):
                U = case.step(U, dt)
            rates_.append(U_.size * 3 * steps / (time.perf_counter() - t0))
        return statistics.median(rates_), steps

    cores = int(lib.oracle_num_threads())
    all_rate, all_steps = trials(rt, mesh, U0)
    rt1, mesh1, U1 = c2_tgv3d(n=12, order=rt.order)
    lib.oracle_set_num_threads(1)
    try:
        one_rate, one_steps = trials(rt1, mesh1, U1)
    finally:
        lib.oracle_set_num_threads(cores)
    return {
        "value": all_rate,
        "unit": "DOF*RHS-evals/s",
        "cores": cores,
        "kind": "port",
        "cpu_model": _cpu_model(),
        "nproc": os.cpu_count(),
        "one_core_value": one_rate,
        "sample": f"median of 5 trials of {all_steps} SSP-RK3 step(s) ({3 * all_steps} RHS evals) on the full "
                  f"{mesh.n[0]}^3 p={rt.order} grid with {cores} OpenMP threads; one core: median of 5 trials of "
                  f"{one_steps} step(s) on a 12^3 grid of the same physics; each after 1 untimed step",
    }


def serving_leg(num_cases, log, fp64=False):
    """Served-case stream: 3-D periodic Euler Taylor-Green-type cases
    (n^3 hexes, n in 8..24, p in {2,3,4}, amplitude in [0.5,5], random
    phases, 50 SSP-RK3 steps each, mixed precision), continuous batching over
    one device pool per order under a 64 GiB HBM budget.  A saturated burst
    (all cases at t = 0) measures the case throughput X; the Poisson stream
    then arrives at 0.7 X and reports p50/p99 time to solution.  Two ways of
    advancing the pools run the same burst and the same stream: a launch per
    order on a stream each, and one launch per stage for all orders."""
    from theseus_amd.serving import Server, make_requests

    burst = make_requests(num=64, rate_per_s=1e9)

    def variant(single, rate=None):
        srv = Server(hbm_budget_bytes=64 << 30, max_cases=256, single_launch=single)
        srv.run(make_requests(num=6, rate_per_s=1e9, n_range=(8, 10)))  # warm the graphs and kernels
        sat = srv.run(burst)
        x = sat["cases_done"] / sat["makespan_s"]
        rate = 0.7 * x if rate is None else rate
        res = srv.run(make_requests(num=num_cases, rate_per_s=rate))
        info = {p: pool.info() for p, pool in srv.pools.items()}
        srv.close()
        name = "one launch for all orders" if single else "a launch per order"
        log(f"serving ({name}): saturated {x:.1f} cases/s, {sat['throughput_dof_rhs_per_s']:.3e} DOF*RHS/s; "
            f"stream at {rate:.1f}/s: p50 {res['tts_p50_s']*1e3:.1f} ms, p99 {res['tts_p99_s']*1e3:.1f} ms")
        return {
            "cases": res["cases_done"],
            "arrival_rate_per_s": rate,
            "saturated_cases_per_s": x,
            "saturated_dof_rhs_per_s": sat["throughput_dof_rhs_per_s"],
            "tts_p50_ms": res["tts_p50_s"] * 1e3,
            "tts_p99_ms": res["tts_p99_s"] * 1e3,
            "throughput_dof_rhs_per_s": res["throughput_dof_rhs_per_s"],
            "makespan_s": res["makespan_s"],
            "pools": {str(k): v for k, v in info.items()},
        }

    per_order = variant(False)
    single = variant(True, rate=per_order["arrival_rate_per_s"])
    fp64_rate = None
    if fp64:
        from theseus_amd import _abi as A

        s64 = Server(hbm_budget_bytes=64 << 30, max_cases=256, precision=A.PRECISION_FP64)
        s64.run(make_requests(num=6, rate_per_s=1e9, n_range=(8, 10)))
        fp64_rate = s64.run(burst)["throughput_dof_rhs_per_s"]
        s64.close()
        log(f"serving: FP64 saturated {fp64_rate:.3e} DOF*RHS/s")
    out = dict(per_order)
    out["fp64_throughput_dof_rhs_per_s"] = fp64_rate
    out["single_launch"] = single
    out["workload"] = ("Euler TGV-type cases, n^3 hexes n in [8,24], p in {2,3,4}, amplitude U(0.5,5), 50 SSP-RK3 "
                       "steps each, mixed precision, Poisson arrivals, 64 GiB budget; top-level figures: a launch "
                       "per order (pools on their own streams); single_launch: one launch per stage for all orders, "
                       "the same arrival rate")
    return out


def _lte_run(viscous, steps, mixed=False, total_steps=50, flux="lteec"):
    """SSP-RK3 steps of the LTE Taylor-Green vortex (24^3, p=3) with the flux
    pair `flux`: elapsed ms of `steps` timed steps, finiteness, the
    per-kernel means (slot 0: the RHS pass; slot 1: the temperature recovery
    for Euler, the gradient pass for Navier-Stokes), and the state after
    `total_steps` steps (device tensor)."""
    import dataclasses

    import torch

    from theseus_amd import _abi as A
    from theseus_amd.cases import c4_lte_tgv
    from theseus_amd.operator import MakeRHSOperator

    rt, mesh, U0, table = c4_lte_tgv(n=24, order=3, viscous=viscous, flux=flux)
    if mixed:
        rt = dataclasses.replace(rt, precision=A.PRECISION_MIXED)
    op = MakeRHSOperator(rt, mesh, lte_table=table)
    u = torch.from_numpy(U0).cuda()
    s = op.EstimateStability(u)
    dt = 0.5 / s.TotalRate()
    t = 0.0
    t = op.Step(u, t, dt)
    torch.cuda.synchronize()
    op.enable_kernel_timing(3 * steps)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        t = op.Step(u, t, dt)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    k_ms, _ = op.kernel_timing(0)
    r_ms, _ = op.kernel_timing(1)
    for _ in range(total_steps - 1 - steps):
        t = op.Step(u, t, dt)
    torch.cuda.synchronize()
    finite = bool(torch.isfinite(u).all())
    op.close()
    return ms, finite, k_ms, r_ms, U0.size, u


def _lte_entropy(u):
    """Entropy production (FP64 check) of the FP64 and the mixed-precision
    LTE right-hand sides at the state u (24^3, p=3)."""
    import dataclasses

    import torch

    from theseus_amd import _abi as A
    from theseus_amd.cases import c4_lte_tgv
    from theseus_amd.operator import MakeRHSOperator

    rt, mesh, _, table = c4_lte_tgv(n=24, order=3)
    out = {}
    du = torch.empty_like(u)
    for name, r in (("fp64", rt), ("mixed", dataclasses.replace(rt, precision=A.PRECISION_MIXED))):
        op = MakeRHSOperator(r, mesh, lte_table=table)
        op.Mult(u, du)
        P, M = op.EntropyProduction(u, du)
        out[name] = P
        if name == "fp64":
            out["magnitude"] = M
        op.close()
    out["mixed_minus_fp64_over_magnitude"] = (out["mixed"] - out["fp64"]) / out["magnitude"]
    return out


def lte_leg(log, steps=10):
    """LTE Taylor-Green vortex, 24^3 hexes, p=3, LTEEC flux, the synthetic
    256 x 256 Helmholtz table, FP64: DOF*RHS-evals/s over SSP-RK3 steps; the
    mixed-precision form on the same workload (speed and per-field state
    error after 50 steps against FP64); then the workload as Navier-Stokes
    (BR1, viscosity and conductivity from the table's mu and lambda planes)."""
    import torch

    ms, finite, k_ms, r_ms, dof, u64 = _lte_run(False, steps)
    value = dof * 3 * steps / (ms * 1e-3)
    log(f"lte: {ms/steps:.3f} ms/step, {value:.3e} DOF*RHS/s, kernel {k_ms*1e3:.1f} us, "
        f"temperature recovery {r_ms*1e3:.1f} us")
    mx_ms, mx_finite, mk_ms, _, _, umx = _lte_run(False, steps, mixed=True)
    mx_value = dof * 3 * steps / (mx_ms * 1e-3)
    a = umx.reshape(5, -1)
    b = u64.reshape(5, -1)
    mx_err = [float(torch.linalg.norm(a[q] - b[q]) / torch.linalg.norm(b[q])) for q in range(5)]
    # the FP64 entropy-production check of both right-hand sides at the FP64 state
    entropy = _lte_entropy(u64)
    del u64, umx, a, b
    log(f"lte mixed: {mx_ms/steps:.3f} ms/step ({mx_value/value:.2f}x), kernel {mk_ms*1e3:.1f} us, "
        f"state err after 50 steps vs FP64 {mx_err}, entropy production fp64 {entropy['fp64']:.6e} "
        f"mixed {entropy['mixed']:.6e} (magnitude {entropy['magnitude']:.3e})")
    ns_ms, ns_finite, p2_ms, p1_ms, _, _ = _lte_run(True, steps)
    ns_value = dof * 3 * steps / (ns_ms * 1e-3)
    log(f"lte navier-stokes: {ns_ms/steps:.3f} ms/step, {ns_value:.3e} DOF*RHS/s, pass 2 {p2_ms*1e3:.1f} us, "
        f"pass 1 {p1_ms*1e3:.1f} us")
    # the other flux pairs of the tabulated gas, Euler, FP64 and mixed
    pairs = {}
    for fx in ("pirozzoli", "hll", "llf"):
        f_ms, f_fin, fk_ms, fr_ms, _, _ = _lte_run(False, steps, flux=fx)
        m_ms, m_fin, mk2_ms, _, _, _ = _lte_run(False, steps, mixed=True, flux=fx)
        pairs[fx] = {"value": dof * 3 * steps / (f_ms * 1e-3), "ms_per_step": f_ms / steps, "kernel_mean_ms": fk_ms,
                     "recovery_mean_ms": fr_ms, "mixed_ms_per_step": m_ms / steps, "mixed_kernel_mean_ms": mk2_ms,
                     "finite": f_fin and m_fin}
        log(f"lte {fx}: {f_ms/steps:.3f} ms/step, kernel {fk_ms*1e3:.1f} us, recovery + p, c "
            f"{fr_ms*1e3:.1f} us; mixed {m_ms/steps:.3f} ms/step, kernel {mk2_ms*1e3:.1f} us")
    return {"value": value, "unit": "DOF*RHS-evals/s", "ms_per_step": ms / steps, "steps": steps,
            "kernel_mean_ms": k_ms, "recovery_mean_ms": r_ms, "finite": finite, "dof": dof,
            "workload": "3-D LTE Taylor-Green vortex, 24^3 hexes, p=3, LTEEC, synthetic 256x256 Helmholtz table, "
                        "rho0 0.1, T0 3500 K, Mach 0.3, FP64",
            "mixed": {"value": mx_value, "ms_per_step": mx_ms / steps, "speedup_vs_fp64": mx_value / value,
                      "kernel_mean_ms": mk_ms, "finite": mx_finite,
                      "state_rel_l2_vs_fp64_per_field": mx_err, "state_after_steps": 50,
                      "entropy_production": entropy,
                      "arithmetic": "FP32 pair and face fluxes (Itoh-Abe means on an FP32 copy of the table); "
                                    "FP64 temperature recovery, nodal pressure and sound speed, stage combination"},
            "navier_stokes": {"value": ns_value, "ms_per_step": ns_ms / steps, "pass2_mean_ms": p2_ms,
                              "pass1_mean_ms": p1_ms, "finite": ns_finite,
                              "workload": "the same, Navier-Stokes BR1 with the table's mu and lambda planes"},
            "flux_pairs": {"workload": "the same (Euler) with the Pirozzoli, central/HLL and central/LLF pairs",
                           **pairs}}


# Whole-RHS bytes per DOF (node x equation) of the workload table's
# roofline fraction: the state read and the RHS written (Euler, 16), plus the
# gradient written and read (Navier-Stokes BR1, 72); half in the mixed mode.
TABLE_B_PER_DOF = {("euler", 0): 16.0, ("euler", 1): 8.0, ("ns", 0): 72.0, ("ns", 1): 36.0}


def _gpu_rate(rt, mesh, U0, steps, warmup=5, lte_table=None):
    """DOF*RHS-evals/s over `steps` SSP-RK3 steps at CFL 0.5 (device-resident)."""
    import torch

    from theseus_amd.operator import MakeRHSOperator

    op = MakeRHSOperator(rt, mesh, lte_table=lte_table)
    u = torch.from_numpy(U0).cuda()
    dt = 0.5 / op.EstimateStability(u).TotalRate()
    t = 0.0
    for _ in range(warmup):
        t = op.Step(u, t, dt)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        t = op.Step(u, t, dt)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    ok = bool(torch.isfinite(u).all())
    op.close()
    return U0.size * 3 * steps / (ms * 1e-3), ok


def _cpu_rate(rt, mesh, U0, seconds, lte_table=None):
    """The CPU oracle's SSP-RK3 steps on the host cores (bounded sample)."""
    from oracle import pyoracle
    from theseus_amd.operator import build_descriptor

    if not pyoracle.LIB.exists():
        pyoracle.build()
    lib = pyoracle.load()
    desc, keep = build_descriptor(rt, mesh)
    case = pyoracle.OracleCase(desc)
    tab = None
    if lte_table is not None:
        t = lte_table
        tab = pyoracle.LteTable(arrays=(t.rho, t.T, t.e, t.planes.reshape(-1), t.inv.reshape(-1)))
        tab.attach(case)
    r = case.stability(U0)
    dt = 0.5 / (max(r[0], r[1]) + r[2])
    U = case.step(U0, dt)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds or n == 0:
        U = case.step(U, dt)
        n += 1
    el = time.perf_counter() - t0
    return U0.size * 3 * n / el, int(lib.oracle_num_threads()), n


def config_table(log, c2_row, lte_row, serving_row, cpu_seconds=8.0):
    """The workload table (C1-C5): GPU FP64 and mixed DOF*RHS-evals/s over
    SSP-RK3 steps, the CPU oracle on the host cores (bounded samples) and the
    whole-RHS roofline fraction value * B/DOF / 8 TB/s."""
    import dataclasses

    from theseus_amd import _abi as A
    from theseus_amd.cases import c1_vortex2d, c3_riemann2d, c4_lte_tgv
    from theseus_amd.serving import euler_runtime, tgv_case

    rows = []

    def frac(v, kind, prec):
        return v * TABLE_B_PER_DOF[(kind, prec)] / (PEAK_HBM_GBPS * 1e9)

    for name, make, steps in (("C1 2-D Euler vortex 16^2, p=3, blending", c1_vortex2d, 300),
                              ("C3 2-D Riemann 128^2, p=4, blending, transmissive", c3_riemann2d, 60)):
        rt, mesh, U0 = make()
        g64, ok64 = _gpu_rate(rt, mesh, U0, steps)
        gmx, okmx = _gpu_rate(dataclasses.replace(rt, precision=A.PRECISION_MIXED), mesh, U0, steps)
        cpu, cores, n = _cpu_rate(rt, mesh, U0, cpu_seconds)
        rows.append({"config": name, "dof": U0.size, "gpu_fp64": g64, "gpu_mixed": gmx, "cpu_fp64": cpu,
                     "cpu_cores": cores, "cpu_steps": n, "roofline_frac_fp64": frac(g64, "euler", 0),
                     "roofline_frac_mixed": frac(gmx, "euler", 1), "finite": ok64 and okmx})
        log(f"table: {name}: GPU {g64:.3e} / mixed {gmx:.3e}, CPU {cpu:.3e} ({cores} cores)")
    from theseus_amd.cases import c2_tgv3d

    rt, mesh, U0 = c2_tgv3d(n=16)
    cpu, cores, n = _cpu_rate(rt, mesh, U0, cpu_seconds)
    rows.append({"config": "C2 3-D NS TGV 32^3, p=3", "dof": 10485760, "gpu_fp64": c2_row["value"],
                 "gpu_mixed": c2_row["mixed"]["value"], "cpu_fp64": cpu, "cpu_cores": cores, "cpu_steps": n,
                 "cpu_sample": "16^3 p=3 (same rate per DOF: element-local work)",
                 "roofline_frac_fp64": frac(c2_row["value"], "ns", 0),
                 "roofline_frac_mixed": frac(c2_row["mixed"]["value"], "ns", 1)})
    if lte_row is not None:
        rt, mesh, U0, table = c4_lte_tgv(n=8, order=3)
        cpu, cores, n = _cpu_rate(rt, mesh, U0, cpu_seconds, lte_table=table)
        mxr = lte_row.get("mixed")
        rows.append({"config": "C4 3-D LTE TGV 24^3, p=3, LTEEC", "dof": lte_row["dof"], "gpu_fp64": lte_row["value"],
                     "gpu_mixed": mxr["value"] if mxr else None, "cpu_fp64": cpu, "cpu_cores": cores,
                     "cpu_steps": n, "cpu_sample": "8^3 p=3 (same rate per DOF)",
                     "roofline_frac_fp64": frac(lte_row["value"], "euler", 0),
                     "roofline_frac_mixed": frac(mxr["value"], "euler", 1) if mxr else None})
        log(f"table: C4: GPU {lte_row['value']:.3e}, CPU {cpu:.3e}")
        nsr = lte_row.get("navier_stokes")
        if nsr is not None:
            rt, mesh, U0, table = c4_lte_tgv(n=8, order=3, viscous=True)
            cpu, cores, n = _cpu_rate(rt, mesh, U0, cpu_seconds, lte_table=table)
            rows.append({"config": "C4 as Navier-Stokes (BR1, tabulated mu and lambda)", "dof": lte_row["dof"],
                         "gpu_fp64": nsr["value"], "gpu_mixed": None, "cpu_fp64": cpu, "cpu_cores": cores,
                         "cpu_steps": n, "cpu_sample": "8^3 p=3 (same rate per DOF)",
                         "roofline_frac_fp64": frac(nsr["value"], "ns", 0), "roofline_frac_mixed": None})
            log(f"table: C4 Navier-Stokes: GPU {nsr['value']:.3e}, CPU {cpu:.3e}")
    if serving_row is not None:
        rt = euler_runtime(3, A.PRECISION_FP64)
        mesh, U0 = tgv_case(rt, 16, 2.0, (0.1, 0.2, 0.3))
        cpu, cores, n = _cpu_rate(rt, mesh, U0, cpu_seconds)
        rows.append({"config": "C5 serving, 512 cases, 50 steps", "gpu_fp64": serving_row.get("fp64_throughput_dof_rhs_per_s"),
                     "gpu_mixed": serving_row["saturated_dof_rhs_per_s"], "cpu_fp64": cpu, "cpu_cores": cores,
                     "cpu_steps": n, "cpu_sample": "one 16^3 p=3 served case",
                     "tts_p50_ms": serving_row["tts_p50_ms"], "tts_p99_ms": serving_row["tts_p99_ms"],
                     "roofline_frac_fp64": frac(serving_row["fp64_throughput_dof_rhs_per_s"], "euler", 0)
                     if serving_row.get("fp64_throughput_dof_rhs_per_s") else None,
                     "roofline_frac_mixed": frac(serving_row["saturated_dof_rhs_per_s"], "euler", 1),
                     "note": "GPU columns: saturated burst of 64 cases (all arrivals at t=0)"})
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=32, help="elements per direction")
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="skip the served-case stream")
    ap.add_argument("--no-lte", action="store_true", help="skip the LTE Taylor-Green leg")
    ap.add_argument("--serve-cases", type=int, default=512)
    ap.add_argument("--table", action="store_true",
                    help="also measure every workload (C1-C5: GPU FP64 / mixed, CPU oracle samples)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy 10 times (scales the HBM byte counters)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch

        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read_per_launch": src.numel() * 8, "copy_bytes_written_per_launch": src.numel() * 8}))
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this engine runs on one GPU; sharding across GPUs (element partitions with "
                                   "a face-trace halo exchange) is not built or measured", "n_gpus": args.gpus}))
        sys.exit(2)

    import dataclasses

    import numpy as np
    import torch

    from theseus_amd import _abi as A
    from theseus_amd.cases import c2_tgv3d
    from theseus_amd.operator import MakeRHSOperator

    rt, mesh, U0 = c2_tgv3d(n=args.n, order=args.order)
    dof = U0.size
    op = MakeRHSOperator(rt, mesh)
    stab = op.EstimateStability(torch.from_numpy(U0).cuda())
    dt = args.cfl / stab.TotalRate()
    log(f"workload: TGV {args.n}^3 p={args.order} NS, DOF={dof}, dt={dt:.4e} (CFL {args.cfl})")

    def timed_run(op):
        """warmup + timed SSP-RK3 steps from U0, then the same number of steps
        again with every kernel launch bracketed by HIP events on its stream;
        returns ms of the first run, the per-kernel means of the second, the
        final state.  The step time carries no per-launch events: recording
        them adds 3-5 % to a step (measured alternating on one box)."""
        u = torch.from_numpy(U0).cuda()
        t = 0.0
        for _ in range(args.warmup):
            t = op.Step(u, t, dt)
        stream = torch.cuda.current_stream()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(args.steps):
            t = op.Step(u, t, dt)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        op.enable_kernel_timing(3 * args.steps)
        for _ in range(args.steps):
            t = op.Step(u, t, dt)
        torch.cuda.synchronize()
        k0, k1 = op.kernel_timing(0), op.kernel_timing(1)
        op.enable_kernel_timing(0)
        if not torch.isfinite(u).all():
            raise RuntimeError("non-finite state after the timed steps")
        return ms, k0, k1, u

    def per_field(a, b):
        a = a.reshape(5, -1)
        b = b.reshape(5, -1)
        return [float(torch.linalg.norm(a[q] - b[q]) / torch.linalg.norm(b[q])) for q in range(5)]

    ms, (k2_ms, k2_n), (k1_ms, k1_n), u64 = timed_run(op)
    value = dof * 3 * args.steps / (ms * 1e-3)

    # mixed precision on the same workload, dt and step count
    rtm = dataclasses.replace(rt, precision=A.PRECISION_MIXED)
    opm = MakeRHSOperator(rtm, mesh)
    msm, (m2_ms, _), (m1_ms, _), u32 = timed_run(opm)
    value_m = dof * 3 * args.steps / (msm * 1e-3)
    # accuracy over the workload's full run (200 SSP-RK3 steps from the
    # initial condition at the same dt, untimed), whatever --steps is: the
    # z-momentum starts at zero, so its relative error early in a run is
    # measured against a field of nearly zero norm
    acc_steps = 200
    ua = torch.from_numpy(U0).cuda()
    uam = ua.clone()
    for _ in range(acc_steps):
        op.Step(ua, 0.0, dt)
        opm.Step(uam, 0.0, dt)
    # the RHS of both paths on the evolved FP64 state: at t = 0 the density
    # RHS of the Taylor-Green vortex vanishes (divergence-free velocity,
    # uniform density), so a relative error there measures nothing
    r64 = torch.empty_like(u64)
    r32 = torch.empty_like(u64)
    op.Mult(ua, r64)
    opm.Mult(ua, r32)
    torch.cuda.synchronize()
    # the FP64 entropy-production check of both right-hand sides there
    P64, M64 = op.EntropyProduction(ua, r64)
    P32, _ = opm.EntropyProduction(ua, r32)
    mixed = {
        "value": value_m,
        "ms_per_step": msm / args.steps,
        "speedup_vs_fp64": value_m / value,
        "rhs_rel_l2_vs_fp64_per_field": per_field(r32, r64),
        "rhs_state": f"FP64 state after {acc_steps} steps",
        "state_rel_l2_vs_fp64_per_field": per_field(uam, ua),
        "state_rel_l2_vs_fp64": float(torch.linalg.norm(uam - ua) / torch.linalg.norm(ua)),
        "state_after_steps": acc_steps,
        "kernels_mean_ms": {"ns_pass2_volume_face": m2_ms, "ns_pass1_br1_gradient": m1_ms},
        "entropy_production": {"fp64": P64, "mixed": P32, "magnitude": M64,
                               "mixed_minus_fp64_over_magnitude": (P32 - P64) / M64,
                               "check": "sum w J S(u).dudt in FP64 at the same state; <= 0 is entropy-stable"},
        "arithmetic": "FP32 fluxes, gradients and LDS staging; FP64 state, stage combination and indicator",
    }
    log(f"mixed: {msm/args.steps:.3f} ms/step, {value_m:.4e} DOF*RHS/s ({value_m/value:.2f}x), "
        f"state err after {acc_steps} steps {mixed['state_rel_l2_vs_fp64_per_field']}, "
        f"RHS err {mixed['rhs_rel_l2_vs_fp64_per_field']}, entropy production fp64 {P64:.6e} mixed {P32:.6e} "
        f"(magnitude {M64:.3e})")

    modes = [1, 2, 2]
    b2, b2_dof = algorithmic_bytes(mesh.num_elements, args.order, 3, 0)
    b1, b1_dof = algorithmic_bytes(mesh.num_elements, args.order, 3, 1)
    kernels = {
        "ns_pass2_volume_face": {"mean_ms": k2_ms, "launches": k2_n, "alg_bytes": b2, "alg_B_per_dof": b2_dof,
                                 "moved_B_per_dof": moved_bytes_per_dof(args.order, 3, 0, modes)},
        "ns_pass1_br1_gradient": {"mean_ms": k1_ms, "launches": k1_n, "alg_bytes": b1, "alg_B_per_dof": b1_dof,
                                  "moved_B_per_dof": moved_bytes_per_dof(args.order, 3, 1, modes)},
    }
    dom_name = max(kernels, key=lambda k: kernels[k]["mean_ms"])
    dom = kernels[dom_name]
    achieved = dom["alg_bytes"] / (dom["mean_ms"] * 1e-3) / 1e9
    for k, v in kernels.items():
        gbps = v["alg_bytes"] / (v["mean_ms"] * 1e-3) / 1e9 if v["mean_ms"] > 0 else 0.0
        log(f"kernel {k}: {v['mean_ms']*1e3:.1f} us x {v['launches']}, {gbps:.0f} GB/s algorithmic "
            f"({v['alg_B_per_dof']:.2f} B/DOF)")
    log(f"step: {ms/args.steps:.3f} ms, {value:.4e} DOF*RHS/s")

    traffic = None
    tr_file = ROOT / "profiles" / "traffic_tgv32_p3_fp64.txt"
    if tr_file.exists():
        for line in tr_file.read_text().splitlines():
            if line.startswith(f"{dom_name}_hbm_bytes_per_launch:"):
                traffic = float(line.split(":", 1)[1])

    out = {
        "metric": "DOF*RHS-evals/s (3-D TGV p=3, FP64)",
        "value": value,
        "unit": "DOF*RHS-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Taylor-Green initial condition)",
        "config": {
            "workload": f"3-D compressible Navier-Stokes Taylor-Green vortex, {args.n}^3 hexes, p={args.order}, "
                        "BR1, Ranocha volume + Lax-Friedrichs face, SSP-RK3, Ma 0.1, Re 1600, Pr 0.71",
            "elements": mesh.num_elements,
            "order": args.order,
            "dof": dof,
            "rhs_evals_per_step": 3,
            "dt": dt,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": achieved,
            "peak": PEAK_HBM_GBPS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBPS,
            "traffic": traffic,
            "traffic_source": "profiles/traffic_tgv32_p3_fp64.txt (rocprofv3 FETCH_SIZE/WRITE_SIZE runs of this "
                              "workload), not measured in this run",
            "duration_source": "mean of per-launch HIP events on the launch stream over a second run of the same "
                               "steps (the timed steps carry no per-launch events)",
        },
        "kernels": kernels,
        "mixed": mixed,
    }
    del op, opm
    torch.cuda.empty_cache()
    if not args.no_serving:
        out["serving"] = serving_leg(args.serve_cases, log, fp64=args.table)
    if not args.no_lte:
        out["lte"] = lte_leg(log)
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(rt, mesh, U0, args.cpu_seconds)
    if args.table:
        out["table"] = config_table(log, out, out.get("lte"), out.get("serving"))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
