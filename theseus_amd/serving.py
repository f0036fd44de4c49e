"""Continuous-batching server for many independent DGSEM cases.

Cases of the same (order, flux pairing, precision) share one device pool
(``theseus_pool_*`` in include/theseus_rhs.h): one element array holding
every admitted case, one kernel launch per SSP-RK3 stage for all of them,
the step captured once as a hipGraph and replayed.  Requests that arrive
while others run are admitted as soon as a free element range and a case
slot exist (the pool's capacity comes from an HBM budget); finished cases
are read back and their slots reused.  Pools of different orders run on
their own streams, so their launches overlap on the device.

The reference has no counterpart: its driver runs one case per process
(src/Simulation.cpp:1255).  ``Server.run`` reports per-case time to
solution (p50/p99) and aggregate throughput.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A
from .config import RunTime
from .mesh import box_mesh, gll_nodes
from .operator import build_descriptor


class CapacityError(A.TheseusError):
    pass


class Pool:
    """One device pool (theseus_pool)."""

    def __init__(self, rt: RunTime, hbm_budget_bytes: int, max_cases: int, tile: int = -1, xcd_remap: int = -1):
        self._lib = A.load()
        self.rt = rt
        d = A.PoolDesc()
        d.dim, d.order, d.flow_model = rt.dim, rt.order, rt.flow_model
        d.volume_flux, d.surface_flux, d.precision = rt.volume_flux, rt.surface_flux, rt.precision
        d.blending = int(rt.blending)
        d.gamma, d.Pr, d.R_gas, d.mu, d.mu_bulk = rt.gamma, rt.Pr, rt.R_gas, rt.mu, rt.mu_bulk
        d.alpha_max, d.alpha_min = rt.alpha_max, rt.alpha_min
        d.hbm_budget_bytes = int(hbm_budget_bytes)
        d.max_cases = int(max_cases)
        d.tile, d.xcd_remap = int(tile), int(xcd_remap)
        h = C.c_void_p()
        A.check(self._lib, self._lib.theseus_pool_create(C.byref(d), C.byref(h)))
        self._h = h
        self.ndof = (rt.order + 1) ** rt.dim
        self.neq = rt.dim + 2

    def info(self) -> dict:
        v = [C.c_int() for _ in range(6)]
        A.check(self._lib, self._lib.theseus_pool_info(self._h, *[C.byref(x) for x in v]))
        keys = ("capacity_elements", "free_elements", "largest_free_range", "active_cases", "tile", "xcd_remap")
        return {k: x.value for k, x in zip(keys, v)}

    def admit(self, mesh, u0, cfl: float, stream=None, want_dt: bool = True):
        """Admit one case; returns (case_id, dt).  Raises CapacityError when
        the pool has no free range or case slot for it.  With want_dt False
        the admission is only enqueued on `stream` (dt is computed on the
        device) and dt is returned as None."""
        desc, keep = build_descriptor(self.rt, mesh)
        cid, dt = C.c_int(), C.c_double()
        ptr = u0.data_ptr() if hasattr(u0, "data_ptr") else u0.ctypes.data
        rc = self._lib.theseus_pool_admit(self._h, C.byref(desc), C.c_void_p(ptr), C.c_double(cfl), C.byref(cid),
                                          C.byref(dt) if want_dt else None, _stream(stream))
        del keep
        if rc == A.ERR_CAPACITY:
            raise CapacityError(rc, self._lib.theseus_last_error().decode())
        A.check(self._lib, rc)
        return cid.value, (dt.value if want_dt else None)

    def step(self, stream=None) -> None:
        A.check(self._lib, self._lib.theseus_pool_step(self._h, _stream(stream)))

    def read(self, case_id: int, out, stream=None) -> None:
        ptr = out.data_ptr() if hasattr(out, "data_ptr") else out.ctypes.data
        A.check(self._lib, self._lib.theseus_pool_read(self._h, int(case_id), C.c_void_p(ptr), _stream(stream)))

    def integral_measures(self, case_id: int, stream=None) -> dict:
        """Integral measures of one case's current state (RHSOperator_impl.hpp:530-717)."""
        m = A.IntegralMeasures()
        A.check(self._lib, self._lib.theseus_pool_integral_measures(self._h, int(case_id), C.byref(m),
                                                                    _stream(stream)))
        return m.as_dict()

    def release(self, case_id: int, stream=None) -> None:
        A.check(self._lib, self._lib.theseus_pool_release(self._h, int(case_id), _stream(stream)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.theseus_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PoolGroup:
    """Pools of orders 2, 3 and 4 advanced by one kernel launch per stage
    (theseus_pool_group); each pool keeps its own admission and release."""

    def __init__(self, pools):
        self._lib = A.load()
        self.pools = list(pools)
        arr = (C.c_void_p * len(self.pools))(*[p._h.value for p in self.pools])
        h = C.c_void_p()
        A.check(self._lib, self._lib.theseus_pool_group_create(arr, len(self.pools), C.byref(h)))
        self._h = h

    def step(self, stream=None) -> None:
        A.check(self._lib, self._lib.theseus_pool_group_step(self._h, _stream(stream)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.theseus_pool_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stream(stream):
    if stream is None:
        return None
    return C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


# ---------------------------------------------------------------------------
# Requests: 3-D periodic Euler Taylor-Green-type vortices
# ---------------------------------------------------------------------------
@dataclass
class Request:
    rid: int
    arrival_s: float
    n: int
    order: int
    amplitude: float
    phases: tuple
    steps: int = 50
    # filled by the server
    admitted_s: float = math.nan
    done_s: float = math.nan
    dt: float = math.nan
    result: np.ndarray | None = field(default=None, repr=False)

    @property
    def dofs(self) -> int:
        return self.n ** 3 * (self.order + 1) ** 3 * 5


def _tgv_fields(xp, X, amplitude, phases, gamma, p0):
    x, y, z = X[0] + phases[0], X[1] + phases[1], X[2] + phases[2]
    u = amplitude * xp.sin(x) * xp.cos(y) * xp.cos(z)
    v = -amplitude * xp.cos(x) * xp.sin(y) * xp.cos(z)
    rho = xp.ones_like(u)
    p = p0 + amplitude ** 2 / 16.0 * (xp.cos(2 * x) + xp.cos(2 * y)) * (xp.cos(2 * z) + 2.0)
    E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v)
    return xp.concatenate([rho, rho * u, rho * v, xp.zeros_like(u), E])


def tgv_case(rt: RunTime, n: int, amplitude: float, phases, p_inf: float | None = None):
    """Mesh and state of one served case (numpy): periodic [0,2pi]^3, n^3 hexes,
    rho = 1, u = A sin(x+a) cos(y+b) cos(z+c), v = -A cos(x+a) sin(y+b) cos(z+c),
    w = 0, p = p0 + A^2/16 (cos 2(x+a) + cos 2(y+b)) (cos 2(z+c) + 2), p0 = 100/gamma."""
    mesh = box_mesh(3, n, 0.0, 2 * np.pi, periodic=True)
    X = mesh.node_coordinates(gll_nodes(rt.order)[0])
    p0 = 100.0 / rt.gamma if p_inf is None else p_inf
    return mesh, _tgv_fields(np, X, amplitude, phases, rt.gamma, p0).astype(np.float64)


class CaseFactory:
    """Builds served cases fast: one mesh per n (cached), states on the device."""

    def __init__(self, device="cuda"):
        self.device = device
        self._mesh = {}
        self._coords = {}

    def mesh(self, n: int):
        m = self._mesh.get(n)
        if m is None:
            m = self._mesh[n] = box_mesh(3, n, 0.0, 2 * np.pi, periodic=True)
        return m

    def coords(self, n: int, order: int):
        import torch

        key = (n, order)
        c = self._coords.get(key)
        if c is None:
            # node coordinates in the L-vector dof order e*ndof + (i + Np*(j + Np*k)), x fastest
            x01 = torch.tensor(gll_nodes(order)[0], dtype=torch.float64, device=self.device)
            Np = order + 1
            h = 2 * np.pi / n
            e = torch.arange(n ** 3, device=self.device)
            nd = torch.arange(Np ** 3, device=self.device)
            ei = [e % n, (e // n) % n, e // (n * n)]
            ni = [nd % Np, (nd // Np) % Np, nd // (Np * Np)]
            c = torch.stack([(h * (ei[d][:, None] + x01[ni[d]][None, :])).reshape(-1) for d in range(3)])
            self._coords[key] = c
        return c

    def state(self, rt: RunTime, n: int, amplitude: float, phases, p_inf: float | None = None):
        import torch

        p0 = 100.0 / rt.gamma if p_inf is None else p_inf
        return _tgv_fields(torch, self.coords(n, rt.order), amplitude, phases, rt.gamma, p0).contiguous()


def make_requests(num: int = 512, rate_per_s: float = 100.0, n_range=(8, 24), orders=(2, 3, 4),
                  amp_range=(0.5, 5.0), steps: int = 50, seed_cases: int = 20261202, seed_arrivals: int = 20261203):
    """Poisson arrivals at `rate_per_s`; case parameters drawn uniformly."""
    rc = np.random.default_rng(seed_cases)
    ra = np.random.default_rng(seed_arrivals)
    t = np.cumsum(ra.exponential(1.0 / rate_per_s, size=num))
    reqs = []
    for i in range(num):
        n = int(rc.integers(n_range[0], n_range[1] + 1))
        order = int(rc.choice(orders))
        amp = float(rc.uniform(*amp_range))
        ph = tuple(float(v) for v in rc.uniform(0, 2 * np.pi, size=3))
        reqs.append(Request(i, float(t[i]), n, order, amp, ph, steps))
    return reqs


def euler_runtime(order: int, precision=A.PRECISION_MIXED) -> RunTime:
    """The served physics: 3-D Euler, Ranocha volume + Lax-Friedrichs face."""
    return RunTime(dim=3, order=order, flow_model=A.FLOW_EULER, volume_flux=A.VOLUME_RANOCHA,
                   surface_flux=A.SURFACE_LLF, blending=False, precision=precision)


def percentile(values, q):
    v = np.sort(np.asarray(values, dtype=np.float64))
    if v.size == 0:
        return math.nan
    return float(np.percentile(v, q))


class Server:
    """Continuous batching over one pool per polynomial order.

    The pools advance either on a stream each (a launch per pool and stage,
    the launches of different orders overlapping on the device), or, with
    ``single_launch``, together: one launch per stage runs every order
    (PoolGroup) on one stream.  The host loop never blocks on the device
    while work is in flight: steps go out in batches of at most
    ``max_batch_steps`` (never past a case's last step), each followed by an
    event; the loop polls the events, admits arrivals the moment a range is
    free (admission is only enqueued on the stream, behind the batches
    already there), and keeps up to ``max_inflight`` batches queued so the
    device does not wait for the host between them.  A case joins the
    batches enqueued after its admission; a finished case is read and
    released once its last batch's event has completed."""

    def __init__(self, orders=(2, 3, 4), hbm_budget_bytes: int = 64 << 30, max_cases: int = 256,
                 precision=A.PRECISION_MIXED, cfl: float = 0.5, keep_results: bool = False, tile: int = -1,
                 xcd_remap: int = -1, max_batch_steps: int = 4, max_inflight: int = 2, single_launch: bool = False):
        import torch

        self.torch = torch
        self.cfl = cfl
        self.max_batch_steps = max(1, int(max_batch_steps))
        self.max_inflight = max(1, int(max_inflight))
        self.keep_results = keep_results
        budget = int(hbm_budget_bytes) // len(orders)
        self.pools = {p: Pool(euler_runtime(p, precision), budget, max_cases, tile, xcd_remap) for p in orders}
        self.group = None
        if single_launch:
            self.group = PoolGroup([self.pools[p] for p in orders])
            self.lanes = [tuple(orders)]
        else:
            self.lanes = [(p,) for p in orders]
        self.lane_of = {p: lane for lane in self.lanes for p in lane}
        self.lane_stream = {lane: torch.cuda.Stream() for lane in self.lanes}  # default priority
        self.streams = {p: self.lane_stream[self.lane_of[p]] for p in orders}
        self.factory = CaseFactory()

    def close(self):
        if self.group is not None:
            self.group.close()
        for pool in self.pools.values():
            pool.close()

    def _step(self, lane, stream):
        if self.group is not None:
            self.group.step(stream)
        else:
            self.pools[lane[0]].step(stream)

    def run(self, requests, time_limit_s: float = 600.0, log=None) -> dict:
        torch = self.torch
        waiting = deque(sorted(requests, key=lambda r: r.arrival_s))
        queued: list[Request] = []
        # lane -> {(order, case_id): [request, steps completed, steps enqueued]}
        active = {lane: {} for lane in self.lanes}
        inflight = {lane: deque() for lane in self.lanes}  # lane -> (event, steps, case keys) per batch
        done: list[Request] = []
        states = {}
        t0 = time.perf_counter()
        dof_rhs = 0.0
        while (waiting or queued or any(active.values())) and time.perf_counter() - t0 < time_limit_s:
            progressed = False
            # completed batches: count their steps, hand back finished cases
            for lane, q in inflight.items():
                st = self.lane_stream[lane]
                while q and q[0][0].query():
                    _, k, keys = q.popleft()
                    progressed = True
                    cases = active[lane]
                    for key in keys:
                        rec = cases[key]
                        rec[1] += k
                        dof_rhs += 3 * k * rec[0].dofs
                    for key in [c for c in keys if cases[c][1] >= cases[c][0].steps]:
                        p, cid = key
                        r = cases.pop(key)[0]
                        if self.keep_results:
                            out = torch.empty(r.dofs, dtype=torch.float64, device="cuda")
                            self.pools[p].read(cid, out, st)
                            r.result = out.cpu().numpy()
                        else:
                            out = states.get((p, r.dofs))
                            if out is None:
                                out = states[(p, r.dofs)] = torch.empty(r.dofs, dtype=torch.float64, device="cuda")
                            self.pools[p].read(cid, out, st)
                        self.pools[p].release(cid, st)
                        r.done_s = time.perf_counter() - t0
                        done.append(r)
                        if log:
                            log(f"case {r.rid} p={r.order} n={r.n} done in {r.done_s - r.arrival_s:.3f} s")
            now = time.perf_counter() - t0
            while waiting and waiting[0].arrival_s <= now:
                queued.append(waiting.popleft())
            # admit in arrival order; a full pool does not block the others
            still = []
            for r in queued:
                pool = self.pools[r.order]
                st = self.streams[r.order]
                with torch.cuda.stream(st):
                    u0 = self.factory.state(pool.rt, r.n, r.amplitude, r.phases)
                try:
                    cid, _ = pool.admit(self.factory.mesh(r.n), u0, self.cfl, st, want_dt=False)
                except CapacityError:
                    still.append(r)
                    continue
                r.admitted_s = time.perf_counter() - t0
                active[self.lane_of[r.order]][(r.order, cid)] = [r, 0, 0]
                progressed = True
            queued = still
            # keep each lane's queue of batches topped up
            for lane, cases in active.items():
                if not cases or len(inflight[lane]) >= self.max_inflight:
                    continue
                k = min(self.max_batch_steps, min(rec[0].steps - rec[2] for rec in cases.values()))
                if k <= 0:
                    continue  # a case ends with the batches in flight: wait for them
                st = self.lane_stream[lane]
                for _ in range(k):
                    self._step(lane, st)
                for rec in cases.values():
                    rec[2] += k
                ev = torch.cuda.Event()
                ev.record(st)
                inflight[lane].append((ev, k, list(cases.keys())))
                progressed = True
            if not progressed:
                if not any(active.values()) and waiting:
                    time.sleep(max(0.0, min(waiting[0].arrival_s - (time.perf_counter() - t0), 0.01)))
                else:
                    time.sleep(2e-4)
        makespan = time.perf_counter() - t0
        for st in self.lane_stream.values():
            st.synchronize()
        tts = [r.done_s - r.arrival_s for r in done]
        return {
            "cases_done": len(done),
            "cases_total": len(requests),
            "makespan_s": makespan,
            "throughput_dof_rhs_per_s": dof_rhs / makespan if makespan > 0 else 0.0,
            "tts_p50_s": percentile(tts, 50),
            "tts_p99_s": percentile(tts, 99),
            "tts_mean_s": float(np.mean(tts)) if tts else math.nan,
            "pools": {p: pool.info() for p, pool in self.pools.items()},
        }
