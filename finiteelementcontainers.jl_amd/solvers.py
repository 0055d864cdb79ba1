"""Newton with conjugate gradients on the GPU: the caller of the assembly path (src/Solvers.jl:95-220).

IterativeLinearSolver.solve(Uu, p) assembles the residual and the CSR Jacobian in one pass, solves K dU = R with the
library's FP64 CG on the device (Jacobi-preconditioned; the reference calls Krylov.jl's cg with the identity
preconditioner) and updates Uu -= dU, as solve!(::IterativeLinearSolver) does (src/Solvers.jl:128-153).
NewtonSolver.solve repeats that with the reference's convergence rule (src/Solvers.jl:186-220): stop when
|dU| < abs_increment_tol or |R| < abs_residual_tol or |R| / |R_1| < rel_residual_tol, at most max_iters times.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from .assembler import (_ptr, assemble_vector_and_stiffness_, create_unknowns)

EPS = float(np.finfo(np.float64).eps)


@dataclass
class IterativeLinearSolverSettings:
    rtol: float = EPS ** 0.5   # Krylov.jl cg defaults: ||r|| <= atol + rtol ||b||
    atol: float = EPS ** 0.5
    max_iters: int = 10000
    preconditioner: str = "jacobi"   # or "none" (the reference's `I`)


class IterativeLinearSolver:
    def __init__(self, assembler, solver_sym="cg", settings=None):
        if solver_sym != "cg":
            raise ValueError("only :cg runs on the GPU")
        self.assembler = assembler
        self.settings = settings or IterativeLinearSolverSettings()
        self.dUu = create_unknowns(assembler)
        self._x = create_unknowns(assembler)
        self.last_iters = 0
        self.last_resnorm = 0.0

    def cg(self, vals, b, x):
        """x <- K^-1 b by CG from the given x; returns (iterations, final residual norm)."""
        s = self.settings
        it, res = ctypes.c_int(0), ctypes.c_double(0.0)
        self.assembler.ctx.call("fecx_cg_solve", _ptr(vals), _ptr(b), _ptr(x), ctypes.c_double(s.rtol),
                                ctypes.c_double(s.atol), ctypes.c_int(s.max_iters),
                                ctypes.c_int(1 if s.preconditioner == "jacobi" else 0), ctypes.byref(it),
                                ctypes.byref(res))
        return it.value, res.value

    def solve(self, Uu, p):
        asm = self.assembler
        assemble_vector_and_stiffness_(asm, Uu, p)
        R = asm._residual()
        self._x.zero_()
        self.last_iters, self.last_resnorm = self.cg(asm.stiffness_storage, R, self._x)
        # dU = -K^-1 R ; Uu += dU
        self.dUu.zero_()
        asm.ctx.call("fecx_vector_axpy", ctypes.c_int64(self.dUu.numel()), ctypes.c_double(-1.0), _ptr(self._x),
                     _ptr(self.dUu))
        asm.ctx.call("fecx_vector_axpy", ctypes.c_int64(Uu.numel()), ctypes.c_double(1.0), _ptr(self.dUu), _ptr(Uu))


@dataclass
class NewtonSolverSettings:
    max_iters: int = 10
    abs_increment_tol: float = 1e-12
    abs_residual_tol: float = 1e-12
    rel_residual_tol: float = 1e-12
    log_callback: object = None   # (iter, norm_dU, norm_R, rel_R, converged) -> None


@dataclass
class NewtonRecord:
    iteration: int
    norm_dU: float
    norm_R: float
    rel_R: float
    converged: bool
    cg_iters: int


class NewtonSolver:
    def __init__(self, linear_solver, settings=None):
        self.linear_solver = linear_solver
        self.settings = settings or NewtonSolverSettings()
        self.history = []

    def _norm(self, v):
        out = ctypes.c_double(0.0)
        self.linear_solver.assembler.ctx.call("fecx_vector_norm", _ptr(v), ctypes.c_int64(v.numel()),
                                              ctypes.byref(out))
        return out.value

    def solve(self, Uu, p, before_iteration=None):
        """before_iteration(n, history), if given, runs before each linear solve (surrogate schedules use it)."""
        s = self.settings
        self.history = []
        initial = 0.0
        ls = self.linear_solver
        for n in range(1, s.max_iters + 1):
            if before_iteration is not None:
                before_iteration(n, self.history)
            ls.solve(Uu, p)
            norm_dU = self._norm(ls.dUu)
            norm_R = self._norm(ls.assembler.residual_unknowns)
            if n == 1:
                initial = norm_R
            rel_R = norm_R / initial if initial > 0.0 else norm_R
            converged = norm_dU < s.abs_increment_tol or norm_R < s.abs_residual_tol or rel_R < s.rel_residual_tol
            self.history.append(NewtonRecord(n, norm_dU, norm_R, rel_R, converged, ls.last_iters))
            if s.log_callback is not None:
                s.log_callback(n, norm_dU, norm_R, rel_R, converged)
            if converged:
                break
        return self.history


class SurrogateSchedule:
    """Newton with a surrogate material: iteration 1 checks every quadrature point (network or plain fallback), later
    iterations keep those decisions (a fixed residual function, so Newton can converge on it), and once the surrogate
    iterations stall (residual ratio above `stall`) or reach `handover`, the remaining iterations use the plain model,
    which carries the solve to the plain method's tolerance. Records per step how many iterations each phase took."""

    def __init__(self, net, handover=1e-3, stall=0.5):
        self.net, self.handover, self.stall = net, handover, stall
        self.phase = []       # per iteration of the current step: "surrogate" or "plain"
        self.fallbacks = []   # per surrogate iteration: quadrature points on the plain model

    def __call__(self, n, history):
        if n > 1 and self.phase[-1] == "surrogate":
            self.fallbacks.append(self.net.fallbacks()[0])
        if n == 1:
            self.phase, self.fallbacks = ["surrogate"], []
            self.net.enable(True)
            self.net.freeze(False)
            return
        if n == 2:
            self.net.freeze(True)
        if self.phase[-1] == "surrogate":
            r = history[-1].rel_R
            prev = history[-2].rel_R if len(history) >= 2 else 0.0
            if r < self.handover or (prev > 0.0 and r / prev > self.stall):
                self.net.enable(False)
                self.phase.append("plain")
            else:
                self.phase.append("surrogate")
        else:
            self.phase.append("plain")


class J2LoadPath:
    """The J2 workload as a quasi-static problem: n^3 hex8 elements, all boundary nodes driven by the macroscopic
    strain path (E(t) x plus the smooth perturbation, as the workload's displacement), each load step solved by
    Newton + CG from the previous step's solution plus the increment of the prescribed field (extended to the
    interior), state committed after each step."""

    def __init__(self, fec, n, seed=0, surrogate=False, newton_settings=None, linear_settings=None, device=0):
        from . import workloads
        self.fec, self.seed, self.workloads = fec, seed, workloads
        self.wl = wl = workloads.j2_hex8(n, seed=seed)
        self.path = wl.extra["path"]
        V = fec.FunctionSpace(wl.mesh, "H1Field", "Lagrange")
        var = fec.VectorFunction(V, "displ")
        self.asm = asm = fec.SparseMatrixAssembler(var, sparse_matrix_type="csr", device=device)
        wl.mesh.nodeset_nodes["_boundary"] = fec.boundary_nodes(wl.mesh)
        self.k = 1
        bcs = [fec.DirichletBC(name, self._bc(d), nset_name="_boundary") for d, name in enumerate(var.names)]
        physics = fec.J2Surrogate() if surrogate else fec.J2Plasticity()
        self.p = fec.create_parameters(wl.mesh, asm, physics, dirichlet_bcs=bcs)
        self.Uu = create_unknowns(asm)
        self.newton = NewtonSolver(IterativeLinearSolver(asm, "cg", linear_settings), newton_settings)
        self._unknown = asm.dof.unknown_dofs - 1
        self._prev = np.zeros(len(self._unknown))

    def _field(self, k):
        return self.workloads.j2_displacement(self.wl.mesh.coords, self.path[k - 1], k / len(self.path), self.seed)

    def _bc(self, d):
        def f(X, t):
            u = self.workloads.j2_displacement(X, self.path[self.k - 1], self.k / len(self.path), self.seed)
            return u.reshape(-1, 3)[:, d]
        return f

    def step(self, k, schedule=None):
        """Load step k (1-based): boundary values, predictor, Newton; then the state commit."""
        self.k = k
        self.p.t = k / len(self.path)
        self.p.update_bc_values()
        cur = self._field(k)[self._unknown]
        self.Uu += torch.from_numpy(cur - self._prev).to(self.Uu.device)
        self._prev = cur
        hist = self.newton.solve(self.Uu, self.p, schedule)
        self.asm.ctx.call("fecx_state_commit")
        return hist

    def state(self):
        """The committed quadrature state (ne, 8, 7) FP32 on the host."""
        ne = self.wl.mesh.num_elements
        st = np.zeros(ne * 8 * 7, dtype=np.float32)
        self.asm.ctx.call("fecx_get_state_new", st.ctypes.data_as(ctypes.c_void_p))
        return st.reshape(ne, 8, 7)


def j2_load_steps(fec, n, steps=20, seed=0, surrogate=False, newton_settings=None, linear_settings=None,
                  device=0, log=None, setup=None, schedule=None):
    """Runs `steps` load steps of J2LoadPath. setup(asm), if given, runs before the first step (it attaches the
    trained network of a surrogate run); schedule, if given, is the Newton iteration hook (SurrogateSchedule).
    Returns the assembler, the parameters, the unknowns, the per-step Newton histories and the Newton solver."""
    lp = J2LoadPath(fec, n, seed, surrogate, newton_settings, linear_settings, device)
    if setup is not None:
        setup(lp.asm)
    histories = []
    for k in range(1, steps + 1):
        hist = lp.step(k, schedule)
        histories.append(hist)
        if log:
            log(k, hist)
    return lp.asm, lp.p, lp.Uu, histories, lp.newton


def j2_collect_samples(fec, n, steps=20, seed=0, newton_settings=None, device=0):
    """Training samples of the plain J2 model on the workload itself: the load steps are solved with the plain model,
    and before every Newton iteration every quadrature point contributes its raw input (displacement gradient at the
    current iterate, state at the start of the step) labelled by the plain update. Returns X (m, 16) and T (m, 13) on
    the device, m = ne x 8 x (Newton iterations)."""
    from . import surrogate as S
    from .assembler import _update_for_assembly
    lp = J2LoadPath(fec, n, seed, False, newton_settings, None, device)
    state = torch.zeros((lp.wl.mesh.num_elements, 8, 7), dtype=torch.float32, device=lp.asm.device)
    xs = []

    def before(k_iter, history):
        _update_for_assembly(lp.p, lp.asm, lp.Uu)
        H = S.qp_gradients(lp.asm, lp.p.field)
        xs.append(torch.cat([H, state], dim=2).reshape(-1, 16))

    for k in range(1, steps + 1):
        lp.step(k, before)
        state.copy_(torch.from_numpy(lp.state()))
    X = torch.cat(xs, dim=0).contiguous()
    T = S.material_update(lp.asm, fec._lib.J2, X)
    return X, T
