"""The ISpdSolver backend slot filled by the MI355X supernodal Cholesky.

``HipSpdSolver`` mirrors src/ImpressionFEA.Core/Solver/Linear/ISpdSolver.cs:24-41 member for member (Name, Kind,
Analyze, Factorize, Solve, FactorNonZeros, Dispose) on top of the C ABI in include/ifea_spd.h, with the reference
backends' error behaviour: not-SPD -> MatrixNotPositiveDefiniteError, call-order violations -> RuntimeError
("Analyze must be called first."), other statuses -> SolverDivergenceError.
"""
from __future__ import annotations

import ctypes as C
import enum

import numpy as np

from . import _native as N
from .errors import raise_for


class LinearBackend(enum.IntEnum):
    """LinearBackend (ISpdSolver.cs:4-17); HipGpu takes the GPU place of CudaGpu."""
    Auto = 0
    HipGpu = 1
    MklCpu = 2
    Managed = 3


class HipSpdSolver:
    """Sparse SPD direct solver on one MI355X; full symmetric CSC in, like every ISpdSolver."""

    Kind = LinearBackend.HipGpu

    def __init__(self, precision: int = 0, refine_tol: float = 1e-10, refine_maxit: int = 30):
        L = N.lib()
        h = C.c_void_p()
        self._check(L.ifea_spd_create(C.byref(h)), "ifea_spd_create")
        self._h = h
        self.set_option(N.IFEA_OPT_PRECISION, precision)
        self.set_option(N.IFEA_OPT_REFINE_TOL, refine_tol)
        self.set_option(N.IFEA_OPT_REFINE_MAXIT, refine_maxit)
        self.n = 0

    @staticmethod
    def _check(rc, what):
        if rc:
            raise_for(rc, what, N.last_error())

    @property
    def Name(self) -> str:
        buf = C.create_string_buffer(256)
        self._check(N.lib().ifea_spd_name(self._h, buf, 256), "ifea_spd_name")
        return buf.value.decode("utf-8", errors="replace")

    def set_option(self, key: int, value: float):
        self._check(N.lib().ifea_spd_set_option(self._h, key, float(value)), "ifea_spd_set_option")

    def set_coordinates(self, coords):
        """Optional DOF coordinates [n, 3] for a geometric nested dissection (None clears them)."""
        if coords is None:
            self._check(N.lib().ifea_spd_set_coordinates(self._h, 0, None), "set_coordinates")
            return
        c = N.f64(coords).reshape(-1, 3)
        self._coords = c
        self._check(N.lib().ifea_spd_set_coordinates(self._h, c.shape[0], N.ptr(c, C.c_double)), "set_coordinates")

    def Analyze(self, n: int, colPtr, rowIdx, perm=None):
        cp, ri = N.i32(colPtr), N.i32(rowIdx)
        if perm is None:
            rc = N.lib().ifea_spd_analyze(self._h, int(n), N.ptr(cp, C.c_int), N.ptr(ri, C.c_int))
        else:
            pm = N.i32(perm)
            rc = N.lib().ifea_spd_analyze_perm(self._h, int(n), N.ptr(cp, C.c_int), N.ptr(ri, C.c_int), N.ptr(pm, C.c_int))
        self._check(rc, "Analyze")
        self.n = int(n)

    def Factorize(self, values):
        v = N.f64(values)
        self._check(N.lib().ifea_spd_factorize(self._h, N.ptr(v, C.c_double)), "Factorize")

    def Solve(self, b, x=None):
        b = N.f64(b)
        out = np.empty(self.n) if x is None else x
        if out.dtype != np.float64 or not out.flags.c_contiguous:
            raise ValueError("x must be a contiguous float64 array")
        self._check(N.lib().ifea_spd_solve(self._h, N.ptr(b, C.c_double), N.ptr(out, C.c_double)), "Solve")
        return out

    @property
    def FactorNonZeros(self) -> int:
        return int(N.lib().ifea_spd_factor_nnz(self._h))

    def symbolic(self):
        """(perm, parent, colcount) of the last Analyze, pivot order."""
        p = np.zeros(self.n, np.int32)
        par = np.zeros(self.n, np.int32)
        cc = np.zeros(self.n, np.int32)
        self._check(N.lib().ifea_spd_symbolic(self._h, N.ptr(p, C.c_int), N.ptr(par, C.c_int), N.ptr(cc, C.c_int)),
                    "symbolic")
        return p, par, cc

    def stats(self):
        out = np.zeros(8)
        self._check(N.lib().ifea_spd_stats(self._h, N.ptr(out, C.c_double)), "stats")
        return dict(flops=out[0], supernodes=int(out[1]), levels=int(out[2]), l_words=int(out[3]),
                    u_words=int(out[4]), refine_iterations=int(out[5]), fp64_fallbacks=int(out[6]),
                    factor_ms=out[7])

    def Dispose(self):
        if getattr(self, "_h", None):
            N.lib().ifea_spd_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.Dispose()

    def __del__(self):
        try:
            self.Dispose()
        except Exception:
            pass


def self_test() -> str | None:
    """The backends' 2x2 self-test (CudssSolver.cs:83-93): None on success, else the reason."""
    try:
        with HipSpdSolver() as s:
            s.Analyze(2, [0, 2, 4], [0, 1, 0, 1])
            s.Factorize([4.0, 1.0, 1.0, 3.0])
            x = s.Solve([1.0, 2.0])
            ok = abs(4 * x[0] + x[1] - 1) < 1e-9 and abs(x[0] + 3 * x[1] - 2) < 1e-9
            return None if ok else "self-test gave a wrong answer"
    except Exception as ex:  # noqa: BLE001 - the probe reports any failure as a reason, like TryCreate
        return str(ex)
