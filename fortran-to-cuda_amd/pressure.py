"""ctypes bindings of the pressure provider and its ELMM adapter (include/ftc_poisson.h), and a small
host-side helper that mirrors the reference's resident_pressure object
(benchmarks/elmm/sources/resident_pressure_state.f90:29-101): create once per domain, solve many times,
destroy before the domain closes."""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_uint32, c_uint64, c_void_p

from . import scope

PERIODIC, NEUMANN = 0, 4
REAL32, MIXED = 1, 2


class SolverConfig(ctypes.Structure):
    _fields_ = [("version", c_uint32), ("flags", c_uint32), ("dimensions", c_int32), ("approximation", c_int32),
                ("sizes", c_int32 * 3), ("lengths", c_double * 3), ("boundary_conditions", c_int32 * 6),
                ("device_budget_bytes", c_uint64)]


class Workspace(ctypes.Structure):
    _fields_ = [("version", c_uint32), ("flags", c_uint32), ("persistent_device_bytes", c_uint64),
                ("peak_device_bytes", c_uint64), ("execution_extra_bytes", c_uint64)]


class Report(ctypes.Structure):
    _fields_ = [("version", c_uint32), ("flags", c_uint32), ("iterations", c_int32), ("converged", c_int32),
                ("relative_residual", c_double), ("rhs_norm", c_double)]


class AdapterConfig(ctypes.Structure):
    _fields_ = [("version", c_uint32), ("reserved", c_uint32), ("solver", SolverConfig),
                ("phi_lower_bounds", c_int64 * 3), ("rhs_lower_bounds", c_int64 * 3)]


SIGNATURES = {
    "poisson_solver_create_v1": (c_int, [POINTER(SolverConfig), POINTER(c_void_p)]),
    "poisson_solver_workspace_query_v1": (c_int, [c_void_p, POINTER(Workspace)]),
    "poisson_solver_execute_device_v1": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p, c_size_t,
                                                 POINTER(c_int32), POINTER(c_int32), c_void_p, POINTER(c_int)]),
    "poisson_solver_destroy_v1": (c_int, [POINTER(c_void_p)]),
    "poisson_solver_last_error_v1": (c_char_p, []),
    "poisson_solver_set_refinement_v1": (c_int, [c_void_p, c_double, c_int32, c_int32]),
    "poisson_solver_last_report_v1": (c_int, [c_void_p, POINTER(Report)]),
    "elmm_gpu_pressure_create_v1": (c_int, [c_uint64, POINTER(AdapterConfig), POINTER(c_void_p)]),
    "elmm_gpu_pressure_solve_v1": (c_int, [c_uint64, c_void_p, c_uint64, c_uint64, POINTER(scope.ExternalReport)]),
    "elmm_gpu_pressure_destroy_v1": (c_int, [POINTER(c_void_p)]),
    "elmm_gpu_pressure_error_v1": (c_char_p, []),
    "elmm_gpu_pressure_provider_v1": (c_void_p, [c_void_p]),
}


def bind(lib: ctypes.CDLL) -> ctypes.CDLL:
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def solver_config(n, lengths, mixed=False, bc_xy=(PERIODIC, PERIODIC), budget=1 << 40) -> SolverConfig:
    bcs = (bc_xy[0], bc_xy[0], bc_xy[1], bc_xy[1], NEUMANN, NEUMANN)
    return SolverConfig(1, MIXED if mixed else 0, 3, 2, (c_int32 * 3)(*n), (c_double * 3)(*lengths),
                        (c_int32 * 6)(*bcs), budget)


class PressureSolver:
    """The ELMM adapter on a scope.Domain: phi (n+4 extents, lower bounds -1) and rhs (n+2, lower bounds 0)."""

    def __init__(self, lib, domain: "scope.Domain", n, lengths, mixed=False, bc_xy=(PERIODIC, PERIODIC)):
        self.lib = bind(lib)
        self.domain = domain
        # The provider's workspace cap is what the domain has not claimed yet
        # (benchmarks/elmm/sources/resident_pressure_state.f90:29-50).
        budget, unclaimed = ctypes.c_size_t(), ctypes.c_size_t()
        scope.check(lib, lib.fort_scope_domain_budget_get_v1(domain.domain, ctypes.byref(budget),
                                                             ctypes.byref(unclaimed)))
        cfg = AdapterConfig(1, 0, solver_config(n, lengths, mixed, bc_xy, unclaimed.value),
                            (c_int64 * 3)(-1, -1, -1), (c_int64 * 3)(0, 0, 0))
        self.handle = c_void_p()
        status = lib.elmm_gpu_pressure_create_v1(domain.domain, ctypes.byref(cfg), ctypes.byref(self.handle))
        if status:
            raise scope.ScopeError(status, (lib.elmm_gpu_pressure_error_v1() or b"").decode())
        self.provider = c_void_p(lib.elmm_gpu_pressure_provider_v1(self.handle))

    def refinement(self, tol=1e-12, max_solves=6, measure=True) -> None:
        if self.lib.poisson_solver_set_refinement_v1(self.provider, tol, max_solves, 1 if measure else 0):
            raise RuntimeError(self.lib.poisson_solver_last_error_v1().decode())

    def solve(self, phi_handle: int, rhs_handle: int) -> scope.ExternalReport:
        report = scope.ExternalReport()
        status = self.lib.elmm_gpu_pressure_solve_v1(self.domain.ctx, self.handle, phi_handle, rhs_handle,
                                                     ctypes.byref(report))
        if status:
            raise scope.ScopeError(status, (self.lib.elmm_gpu_pressure_error_v1() or b"").decode())
        return report

    def report(self) -> Report:
        r = Report()
        self.lib.poisson_solver_last_report_v1(self.provider, ctypes.byref(r))
        return r

    def destroy(self) -> None:
        if self.handle:
            status = self.lib.elmm_gpu_pressure_destroy_v1(ctypes.byref(self.handle))
            if status:
                raise scope.ScopeError(status, (self.lib.elmm_gpu_pressure_error_v1() or b"").decode())
