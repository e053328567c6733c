"""``compile_sparse_mpc_sqp(..., qdldl_backend="hip")``: fixed-pattern SQP for nonlinear MPC.

Mirrors ``warpmpc/jax_sqp/sparse_mpc.py``:

* ``SparseMPCProblem.from_stage_functions`` (``:71-102``): stage functions
  first, intermediate (repeated), terminal; ``horizon + 1`` stages;
* ``build_sparse_mpc_plan`` (``:532-691``): stage variable / parameter /
  constraint offsets, the global P pattern (every diagonal plus each stage's
  cost-Hessian entries) and A pattern (each stage's constraint Jacobian),
  per-stage scatter positions, representative bounds that keep each row's
  equality / lower / upper type (``:476-518``), and the OSQP plan;
* ``compile_sparse_mpc_sqp`` (``:703-1053``): linearization (cost, gradient,
  Hessian, constraint values and Jacobian of every stage, scattered into the
  QP ``P, A, q`` and ``l = lower - g``, ``u = upper - g``), the QP solved by the
  HIP OSQP from a warm start (``solve_warm_start``), the fixed step
  ``z + beta * d`` and the line-search step (the step ladder evaluated at once,
  then the filter of ``line_search.py``).

Stage functions are ``TorchStageFunction`` s; linearization, candidate
evaluation and line search run as batched torch operations on the GPU, the QP
in the fused HIP OSQP kernel.  With ``derivatives=True`` the whole step is
differentiable through torch autograd (the QP through the HIP adjoint).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, NamedTuple, Sequence

import numpy as np
import scipy.sparse as sp

from ..osqp import OSQPSettings, OSQPWarmStart, build_osqp_plan, compile_osqp
from ..osqp.types import FixedOSQPPlan
from .line_search import (
    FilterLineSearchSettings,
    FilterLineSearchResult,
    constraint_violation,
    filter_line_search_from_evaluations,
    make_step_lengths,
)
from .stage import TorchStageFunction, device_const

OSQP_INFTY = 1e30
OSQP_MIN_SCALING = 1e-4


def _torch():
    import torch

    return torch


@dataclass(frozen=True)
class SparseMPCProblem:
    """A stage-ordered nonlinear MPC problem with fixed local sparsity."""

    stages: tuple[TorchStageFunction, ...]

    @classmethod
    def from_stage_functions(cls, *, horizon: int, first: TorchStageFunction,
                             intermediate: TorchStageFunction | Sequence[TorchStageFunction],
                             terminal: TorchStageFunction) -> "SparseMPCProblem":
        if horizon < 1:
            raise ValueError("horizon must be at least 1")
        if isinstance(intermediate, TorchStageFunction):
            middle = [intermediate] * max(0, horizon - 1)
        else:
            middle = list(intermediate)
            if len(middle) != max(0, horizon - 1):
                raise ValueError("intermediate sequence must have length horizon - 1")
        return cls(stages=(first, *middle, terminal))


@dataclass(frozen=True)
class StageAssembly:
    z_offset: int
    next_z_offset: int
    param_offset: int
    constraint_offset: int
    q_cols: np.ndarray
    p_pos: np.ndarray
    a_pos: np.ndarray


@dataclass(frozen=True)
class SparseMPCPlan:
    stage_var_dims: tuple[int, ...]
    param_dims: tuple[int, ...]
    constraint_dims: tuple[int, ...]
    z_offsets: np.ndarray
    param_offsets: np.ndarray
    constraint_offsets: np.ndarray
    p_pattern: sp.csc_matrix
    a_pattern: sp.csc_matrix
    representative_l: np.ndarray
    representative_u: np.ndarray
    osqp_plan: FixedOSQPPlan
    assemblies: tuple[StageAssembly, ...]

    @property
    def n_variables(self) -> int:
        return int(self.z_offsets[-1])

    @property
    def n_parameters(self) -> int:
        return int(self.param_offsets[-1])

    @property
    def n_constraints(self) -> int:
        return int(self.constraint_offsets[-1])


class SQPLinearization(NamedTuple):
    p_values: object
    a_values: object
    q: object
    l: object
    u: object
    cost: object
    g: object
    constraint_l: object
    constraint_u: object


class SQPSolveResult(NamedTuple):
    direction: object
    z: object
    y: object
    prim_res: object
    dual_res: object
    obj_val: object
    linearization: SQPLinearization


class SQPStepResult(NamedTuple):
    z_next: object
    solve: SQPSolveResult


class SQPLineSearchStepResult(NamedTuple):
    z_next: object
    solve: SQPSolveResult
    line_search: FilterLineSearchResult
    is_finite: object


@dataclass(frozen=True)
class CompiledSparseMPCSQP:
    plan: SparseMPCPlan
    qp_solver: str
    osqp: object
    dtype: np.dtype
    evaluate_nlp: Callable
    linearize_nlp: Callable
    build_qp: Callable
    init_state: Callable
    solve_qp_data: Callable
    compute_direction: Callable
    fixed_step: Callable
    step: Callable
    line_search_step: Callable
    line_search_settings: FilterLineSearchSettings
    line_search_step_lengths: np.ndarray
    capture_step: Callable | None = None
    extras: dict = field(default_factory=dict, compare=False, repr=False)


class SQPGraph:
    """One line-search SQP iteration for a fixed batch, captured in a HIP graph.

    The graph holds the whole iteration -- linearization, the warm-started QP
    (the fused HIP OSQP launch), the candidate evaluations and the filter line
    search -- followed by device copies that feed its results back into its
    inputs: the iterate ``z`` and the QP warm start.  ``replay(k)`` therefore
    runs k SQP iterations with k graph launches and no host work in between.
    Inputs are loaded with ``load``; ``z``, ``state`` and ``last`` (the
    ``SQPLineSearchStepResult`` of the latest iteration) are the graph's own
    buffers, overwritten by every replay (clone what must be kept).  No
    counterpart in the reference, which traces the step with jax.jit
    (sparse_mpc.py:997-1053).
    """

    def __init__(self, graph, z, params, state, last, pool=None):
        self.graph = graph
        self.z = z
        self.params = params
        self.state = state
        self.last = last
        self.pool = pool

    @property
    def batch_size(self) -> int:
        return int(self.z.shape[0])

    def load(self, z, params, state: OSQPWarmStart | None = None) -> None:
        """Copies an iterate, parameters and a warm start (zeros if None) into the graph's inputs."""
        self.z.copy_(z)
        self.params.copy_(params)
        for dst, src in zip(self.state, state if state is not None else (None, None, None)):
            if src is None:
                dst.zero_()
            else:
                dst.copy_(src)

    def replay(self, iterations: int = 1):
        """Runs ``iterations`` SQP iterations on the current stream; returns ``z``."""
        for _ in range(int(iterations)):
            self.graph.replay()
        return self.z


def _offsets(widths) -> np.ndarray:
    out = np.zeros(len(widths) + 1, dtype=np.int64)
    np.cumsum(np.asarray(widths, dtype=np.int64), out=out[1:])
    return out


def _csc_pattern(shape, keys):
    """CSC pattern (values 1) of a set of (row, col) keys and each key's slot."""
    rows = np.fromiter((k[0] for k in keys), dtype=np.int64, count=len(keys))
    cols = np.fromiter((k[1] for k in keys), dtype=np.int64, count=len(keys))
    m = sp.csc_matrix((np.ones(len(keys)), (rows, cols)), shape=shape)
    m.sum_duplicates()
    m.sort_indices()
    m.data[:] = 1.0
    pos = {}
    for c in range(shape[1]):
        for k in range(m.indptr[c], m.indptr[c + 1]):
            pos[(int(m.indices[k]), c)] = k
    return m, pos


def _classify_bounds(problem: SparseMPCProblem):
    """Representative bounds keeping each row's type (sparse_mpc.py:476-518):
    bounds probed with all inputs 0, 1 and -1; equal at every probe -> [0, 0],
    finite lower -> 0, finite upper -> 1, else +-1e30."""
    th = _torch()
    big = OSQP_INFTY * OSQP_MIN_SCALING
    lo_parts, up_parts = [], []
    for st in problem.stages:
        los, ups = [], []
        for fill in (0.0, 1.0, -1.0):
            z = th.full((st.z_dim,), fill, dtype=th.float64)
            zn = th.full((st.next_z_dim,), fill, dtype=th.float64)
            p = th.full((st.param_dim,), fill, dtype=th.float64)
            _, _, lo, up = st.values(z, zn, p)
            los.append(lo.detach().numpy())
            ups.append(up.detach().numpy())
        los, ups = np.stack(los), np.stack(ups)
        equal = np.all(np.isclose(los, ups, rtol=1e-9, atol=1e-9), axis=0)
        lower = np.where(np.any(los > -big, axis=0), 0.0, -OSQP_INFTY)
        upper = np.where(np.any(ups < big, axis=0), 1.0, OSQP_INFTY)
        lower[equal] = 0.0
        upper[equal] = 0.0
        lo_parts.append(lower)
        up_parts.append(upper)
    return np.concatenate(lo_parts), np.concatenate(up_parts)


def build_sparse_mpc_plan(problem: SparseMPCProblem, *, osqp_settings: OSQPSettings | None = None,
                          derivatives: bool = False, representative_z=None,
                          representative_params=None) -> SparseMPCPlan:
    """All fixed index data of the SQP's QP assembly (sparse_mpc.py:532-691).

    With ``representative_z`` / ``representative_params`` the OSQP plan's
    scaling is computed from the QP linearized there (sparse_mpc.py:627-660);
    otherwise from the patterns (all values 1)."""
    stages = problem.stages
    if len(stages) < 2:
        raise ValueError("MPC SQP needs at least one transition and one terminal stage")
    for i, st in enumerate(stages[:-1]):
        if not st.has_next:
            raise ValueError(f"stage {i} must have a next-stage input")
    if stages[-1].has_next:
        raise ValueError("terminal stage must not have a next-stage input")
    var_dims = [stages[0].z_dim] + [st.next_z_dim for st in stages[:-1]]
    for i, (d, st) in enumerate(zip(var_dims, stages)):
        if st.z_dim != d:
            raise ValueError(f"stage {i} z_dim={st.z_dim} does not match expected {d}")
    z_off = _offsets(var_dims)
    p_off = _offsets([st.param_dim for st in stages])
    c_off = _offsets([st.constraint_dim for st in stages])
    n, m = int(z_off[-1]), int(c_off[-1])
    p_keys = {(i, i) for i in range(n)}
    a_keys = set()
    local = []
    for i, st in enumerate(stages):
        zo = int(z_off[i])
        glob = np.arange(zo, zo + st.z_dim)
        if st.has_next:
            glob = np.concatenate([glob, np.arange(int(z_off[i + 1]), int(z_off[i + 1]) + st.next_z_dim)])
        pk = []
        for r, c in zip(st.hess_rows, st.hess_cols):
            a, b = int(glob[r]), int(glob[c])
            key = (min(a, b), max(a, b))
            p_keys.add(key)
            pk.append(key)
        ak = []
        for r, c in zip(st.jac_rows, st.jac_cols):
            key = (int(c_off[i]) + int(r), int(glob[c]))
            a_keys.add(key)
            ak.append(key)
        local.append((glob[st.grad_cols], pk, ak))
    p_pattern, p_pos = _csc_pattern((n, n), p_keys)
    a_pattern, a_pos = _csc_pattern((m, n), a_keys)
    assemblies = tuple(
        StageAssembly(z_offset=int(z_off[i]), next_z_offset=int(z_off[i + 1]) if st.has_next else 0,
                      param_offset=int(p_off[i]), constraint_offset=int(c_off[i]),
                      q_cols=np.asarray(qc, dtype=np.int32),
                      p_pos=np.asarray([p_pos[k] for k in pk], dtype=np.int32),
                      a_pos=np.asarray([a_pos[k] for k in ak], dtype=np.int32))
        for i, (st, (qc, pk, ak)) in enumerate(zip(stages, local)))
    settings = osqp_settings or OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, scaling=0)
    lower0, upper0 = _classify_bounds(problem)
    p_for, a_for, scaling_q = p_pattern, a_pattern, None
    if (representative_z is None) != (representative_params is None):
        raise ValueError("representative_z and representative_params must be provided together")
    if representative_z is not None:
        p_for, a_for, scaling_q = _representative_qp(problem, assemblies, p_pattern, a_pattern,
                                                     representative_z, representative_params, n,
                                                     int(p_off[-1]))
    oplan = build_osqp_plan(p_for, a_for, lower0, upper0, settings, derivatives=derivatives, scaling_q=scaling_q)
    return SparseMPCPlan(
        stage_var_dims=tuple(int(v) for v in var_dims),
        param_dims=tuple(st.param_dim for st in stages),
        constraint_dims=tuple(st.constraint_dim for st in stages),
        z_offsets=z_off, param_offsets=p_off, constraint_offsets=c_off,
        p_pattern=p_pattern, a_pattern=a_pattern,
        representative_l=lower0, representative_u=upper0,
        osqp_plan=oplan, assemblies=assemblies,
    )


def _representative_qp(problem, assemblies, p_pattern, a_pattern, z_ref, p_ref, n, n_params):
    """P, A (pattern copies with linearized values) and q at one point, float64."""
    th = _torch()
    z_ref = np.asarray(z_ref, dtype=np.float64)
    p_ref = np.asarray(p_ref, dtype=np.float64)
    z_ref = z_ref[0] if z_ref.ndim == 2 else z_ref
    p_ref = p_ref[0] if p_ref.ndim == 2 else p_ref
    if z_ref.shape != (n,):
        raise ValueError(f"representative_z must have shape {(n,)}, got {z_ref.shape}")
    if p_ref.shape != (n_params,):
        raise ValueError(f"representative_params must have shape {(n_params,)}, got {p_ref.shape}")
    pv = np.zeros(p_pattern.nnz)
    av = np.zeros(a_pattern.nnz)
    q = np.zeros(n)
    zt, pt = th.as_tensor(z_ref), th.as_tensor(p_ref)
    for st, a in zip(problem.stages, assemblies):
        zs = zt[a.z_offset:a.z_offset + st.z_dim][None]
        zn = zt[a.next_z_offset:a.next_z_offset + st.next_z_dim][None]
        ps = pt[a.param_offset:a.param_offset + st.param_dim][None]
        out = [t[0].detach().numpy() for t in st.sparse_values(zs, zn, ps)]
        np.add.at(q, a.q_cols, out[4])
        np.add.at(pv, a.p_pos, out[5])
        np.add.at(av, a.a_pos, out[6])
    p_out, a_out = p_pattern.copy(), a_pattern.copy()
    p_out.data = pv
    a_out.data = av
    return p_out, a_out, q


@dataclass
class _Group:
    """Stages sharing one stage function, evaluated together."""

    stage: TorchStageFunction
    z_offsets: np.ndarray
    next_z_offsets: np.ndarray
    param_offsets: np.ndarray
    constraint_rows: np.ndarray  # (S, c)
    q_cols: np.ndarray           # (S, ng)
    p_pos: np.ndarray            # (S, nh)
    a_pos: np.ndarray            # (S, nj)


def _groups(problem: SparseMPCProblem, plan: SparseMPCPlan, group_repeated: bool):
    order: list[list[int]] = []
    seen: dict[int, int] = {}
    for i, st in enumerate(problem.stages):
        key = id(st) if group_repeated else i
        if key not in seen:
            seen[key] = len(order)
            order.append([])
        order[seen[key]].append(i)
    out = []
    for idx in order:
        st = problem.stages[idx[0]]
        asm = [plan.assemblies[i] for i in idx]
        out.append(_Group(
            stage=st,
            z_offsets=np.asarray([a.z_offset for a in asm]),
            next_z_offsets=np.asarray([a.next_z_offset for a in asm]),
            param_offsets=np.asarray([a.param_offset for a in asm]),
            constraint_rows=np.stack([a.constraint_offset + np.arange(st.constraint_dim) for a in asm]),
            q_cols=np.stack([a.q_cols for a in asm]),
            p_pos=np.stack([a.p_pos for a in asm]),
            a_pos=np.stack([a.a_pos for a in asm]),
        ))
    return out


def compile_sparse_mpc_sqp(problem: SparseMPCProblem, plan: SparseMPCPlan | None = None, *,
                           dtype=np.float32, qp_solver: str = "jax_osqp", osqp_settings: OSQPSettings | None = None,
                           transpose_work: bool = True, segmented: bool = True, segment_budget: int = 64,
                           segment_strategy: str = "optimal", level_scheduled_solve: bool = False,
                           level_scheduled_solve_threshold: int = 1, qdldl_backend: str = "hip",
                           qdldl_factor_backend: str | None = None, qdldl_solve_backend: str | None = None,
                           derivatives: bool = False, derivative_refinement_iters: int = 100,
                           line_search_settings: FilterLineSearchSettings | None = None,
                           group_repeated_stages: bool = True,
                           derivative_mode: str = "reference",
                           derivative_kkt_refinement_iters: int = 4,
                           derivative_kkt_delta: float = 3e-6) -> CompiledSparseMPCSQP:
    """Compile the SQP step functions on the HIP OSQP (sparse_mpc.py:703-1228).

    ``derivative_mode`` selects the QP adjoint (``compile_osqp``): the
    reference's algorithm, or ``"kkt"``, the limit of its refinement."""
    dtype = np.dtype(dtype)
    if dtype != np.dtype(np.float32):
        raise ValueError("the HIP backend computes in float32")
    if qp_solver != "jax_osqp":
        raise ValueError(f"unsupported qp_solver: {qp_solver!r} (this build provides the OSQP QP solver)")
    if plan is None:
        plan = build_sparse_mpc_plan(problem, osqp_settings=osqp_settings, derivatives=derivatives)
    elif derivatives and plan.osqp_plan.derivative_plan is None:
        raise ValueError("compile_sparse_mpc_sqp(..., derivatives=True) requires a plan built with "
                         "build_sparse_mpc_plan(..., derivatives=True), or no explicit plan.")
    qp = compile_osqp(plan.osqp_plan, dtype, qdldl_backend=qdldl_backend, qdldl_factor_backend=qdldl_factor_backend,
                      qdldl_solve_backend=qdldl_solve_backend, transpose_work=transpose_work, segmented=segmented,
                      segment_budget=segment_budget, segment_strategy=segment_strategy,
                      level_scheduled_solve=level_scheduled_solve,
                      level_scheduled_solve_threshold=level_scheduled_solve_threshold, derivatives=derivatives,
                      derivative_refinement_iters=derivative_refinement_iters, derivative_mode=derivative_mode,
                      derivative_kkt_refinement_iters=derivative_kkt_refinement_iters,
                      derivative_kkt_delta=derivative_kkt_delta)
    th = _torch()
    groups = _groups(problem, plan, group_repeated_stages)
    n, m = plan.n_variables, plan.n_constraints
    nnz_p, nnz_a = plan.p_pattern.nnz, plan.a_pattern.nnz
    ls = line_search_settings or FilterLineSearchSettings()
    steps_np = make_step_lengths(ls).astype(np.float32)
    f32 = th.float32
    cache: dict = {}

    def dev_idx(a, device):
        key = (id(a), str(device))
        if key not in cache:
            cache[key] = th.as_tensor(np.asarray(a, dtype=np.int64), device=device)
        return cache[key]

    def as_batch(x, width, name):
        t = th.as_tensor(x, dtype=f32) if not isinstance(x, th.Tensor) else x.to(f32)
        if not t.is_cuda:
            t = t.cuda()
        if t.dim() == 1:
            t = t[None]
        if t.dim() != 2 or t.shape[1] != width:
            raise ValueError(f"{name} must have shape (batch, {width}), got {tuple(t.shape)}")
        return t

    def stage_inputs(g: _Group, z, params):
        st = g.stage
        zc = th.stack([z[:, o:o + st.z_dim] for o in g.z_offsets])
        pc = th.stack([params[:, o:o + st.param_dim] for o in g.param_offsets])
        zn = th.stack([z[:, o:o + st.next_z_dim] for o in g.next_z_offsets]) if st.has_next else zc[..., :0]
        return zc, zn, pc

    def scatter_rows(dst, rows, val):
        # dst (B, m); rows (S, c); val (S, B, c)
        b = dst.shape[0]
        return dst.index_copy(1, dev_idx(rows, dst.device).reshape(-1), val.permute(1, 0, 2).reshape(b, -1))

    def scatter_add(dst, pos, val):
        b = dst.shape[0]
        return dst.index_add(1, dev_idx(pos, dst.device).reshape(-1), val.permute(1, 0, 2).reshape(b, -1))

    def evaluate_values(z, params):
        b = z.shape[0]
        cost = th.zeros((b,), dtype=f32, device=z.device)
        g_all = th.zeros((b, m), dtype=f32, device=z.device)
        lo_all, up_all = th.zeros_like(g_all), th.zeros_like(g_all)
        for g in groups:
            c, gv, lo, up = g.stage.values(*stage_inputs(g, z, params))
            cost = cost + c.sum(dim=0)
            g_all = scatter_rows(g_all, g.constraint_rows, gv)
            lo_all = scatter_rows(lo_all, g.constraint_rows, lo)
            up_all = scatter_rows(up_all, g.constraint_rows, up)
        return cost, g_all, lo_all, up_all

    def linearize(z, params) -> SQPLinearization:
        b = z.shape[0]
        kw = dict(dtype=f32, device=z.device)
        pv, av, q = th.zeros((b, nnz_p), **kw), th.zeros((b, nnz_a), **kw), th.zeros((b, n), **kw)
        g_all = th.zeros((b, m), **kw)
        lo_all, up_all = th.zeros_like(g_all), th.zeros_like(g_all)
        cost = th.zeros((b,), **kw)
        for g in groups:
            c, gv, lo, up, grad, hess, jac = g.stage.sparse_values(*stage_inputs(g, z, params))
            cost = cost + c.sum(dim=0)
            q = scatter_add(q, g.q_cols, grad)
            pv = scatter_add(pv, g.p_pos, hess)
            av = scatter_add(av, g.a_pos, jac)
            g_all = scatter_rows(g_all, g.constraint_rows, gv)
            lo_all = scatter_rows(lo_all, g.constraint_rows, lo)
            up_all = scatter_rows(up_all, g.constraint_rows, up)
        return SQPLinearization(p_values=pv, a_values=av, q=q, l=lo_all - g_all, u=up_all - g_all, cost=cost,
                                g=g_all, constraint_l=lo_all, constraint_u=up_all)

    def init_state(batch_size: int) -> OSQPWarmStart:
        return qp.init_warm_start(batch_size)

    def state_from(solve: SQPSolveResult) -> OSQPWarmStart:
        return OSQPWarmStart(x=solve.direction, z=solve.z, y=solve.y)

    def solve_qp_data(lin: SQPLinearization, *, state: OSQPWarmStart | None = None):
        b = int(lin.q.shape[0])
        state = init_state(b) if state is None else state
        d, qz, qy, prim, dual, obj = qp.solve_warm_start(lin.p_values, lin.a_values, lin.q, lin.l, lin.u,
                                                         state.x, state.z, state.y)
        solve = SQPSolveResult(direction=d, z=qz, y=qy, prim_res=prim, dual_res=dual, obj_val=obj,
                               linearization=lin)
        return solve, state_from(solve)

    def compute_direction(z, params, *, state: OSQPWarmStart | None = None):
        z = as_batch(z, n, "z")
        params = as_batch(params, plan.n_parameters, "params")
        return solve_qp_data(linearize(z, params), state=state)

    def fixed_step(z, params, beta=1.0, *, state: OSQPWarmStart | None = None):
        """z + beta * d (sparse_mpc.py:980-995); returns (SQPStepResult, next state)."""
        z = as_batch(z, n, "z")
        solve, nxt = compute_direction(z, params, state=state)
        return SQPStepResult(z_next=z + beta * solve.direction, solve=solve), nxt

    def line_search_step(z, params, solve: SQPSolveResult) -> SQPLineSearchStepResult:
        """The filter line search along an already solved direction
        (sparse_mpc.py:1011-1053)."""
        z = as_batch(z, n, "z")
        params = as_batch(params, plan.n_parameters, "params")
        base = solve.linearization
        d = solve.direction
        scale = ls.line_search_constraint_scale
        v0 = constraint_violation(base.constraint_l, base.g, base.constraint_u, scale)
        metric = (base.q * d).sum(dim=1)
        steps = device_const(steps_np, f32, z.device)
        k, b = steps.shape[0], z.shape[0]
        cand = (z[None] + steps[:, None, None] * d[None]).reshape(k * b, n)
        cost, g, lo, up = evaluate_values(cand, params.repeat(k, 1))
        viol = constraint_violation(lo, g, up, scale).reshape(k, b)
        res = filter_line_search_from_evaluations(
            settings=ls, step_lengths=steps, baseline_cost=base.cost, baseline_constraint_violation=v0,
            armijo_descent_metric=metric, candidate_costs=cost.reshape(k, b), candidate_constraint_violations=viol)
        z_next = z + res.step_length[:, None] * d
        return SQPLineSearchStepResult(z_next=z_next, solve=solve, line_search=res,
                                       is_finite=th.isfinite(z_next).all(dim=1))

    def step(z, params, *, state: OSQPWarmStart | None = None):
        """Line-search step (sparse_mpc.py:997-1053); returns (result, next state)."""
        z = as_batch(z, n, "z")
        params = as_batch(params, plan.n_parameters, "params")
        solve, nxt = solve_qp_data(linearize(z, params), state=state)
        return line_search_step(z, params, solve), nxt

    def capture_step(batch_size: int, warmup: int = 2) -> SQPGraph:
        """Captures one line-search iteration for ``batch_size`` instances in a
        HIP graph (see ``SQPGraph``)."""
        b = int(batch_size)
        if b <= 0:
            raise ValueError("batch_size must be positive")
        kw = dict(dtype=f32, device="cuda")
        zb = th.zeros((b, n), **kw)
        pb = th.zeros((b, plan.n_parameters), **kw)
        sb = tuple(th.zeros((b, w), **kw) for w in (n, m, m))

        def body():
            res, nxt = step(zb, pb, state=OSQPWarmStart(*sb))
            zb.copy_(res.z_next)
            for dst, src in zip(sb, nxt):
                dst.copy_(src)
            return res

        # warm-up on a side stream at the default priority: index tables and
        # constants reach the device, the caching allocator sees every size
        side = th.cuda.Stream()
        side.wait_stream(th.cuda.current_stream())
        with th.cuda.stream(side):
            for _ in range(max(1, int(warmup))):
                body()
        th.cuda.current_stream().wait_stream(side)
        th.cuda.synchronize()
        graph = th.cuda.CUDAGraph()
        with th.cuda.graph(graph):
            last = body()
        th.cuda.synchronize()
        return SQPGraph(graph, zb, pb, sb, last)

    def evaluate_nlp(z, params):
        return evaluate_values(as_batch(z, n, "z"), as_batch(params, plan.n_parameters, "params"))

    def linearize_nlp(z, params) -> SQPLinearization:
        return linearize(as_batch(z, n, "z"), as_batch(params, plan.n_parameters, "params"))

    return CompiledSparseMPCSQP(
        plan=plan, qp_solver=qp_solver, osqp=qp, dtype=dtype, evaluate_nlp=evaluate_nlp,
        linearize_nlp=linearize_nlp, build_qp=linearize_nlp, init_state=init_state, solve_qp_data=solve_qp_data,
        compute_direction=compute_direction, fixed_step=fixed_step, step=step, line_search_step=line_search_step,
        line_search_settings=ls,
        line_search_step_lengths=steps_np,
        capture_step=capture_step,
    )


__all__ = [
    "CompiledSparseMPCSQP",
    "SQPGraph",
    "SQPLineSearchStepResult",
    "SQPLinearization",
    "SQPSolveResult",
    "SQPStepResult",
    "SparseMPCPlan",
    "SparseMPCProblem",
    "StageAssembly",
    "build_sparse_mpc_plan",
    "compile_sparse_mpc_sqp",
]
