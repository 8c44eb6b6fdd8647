"""Failure vocabulary of the solve step (src/ImpressionFEA.Core/Solver/SolverExceptions.cs:4-43) and the mapping of
C-ABI return codes onto it (CudssSolver.cs:141-150, MklPardisoSolver.cs:92-97)."""
from __future__ import annotations


class SimulationError(Exception):
    """SimulationException: base class for all simulation failures."""


class InputValidationError(SimulationError):
    def __init__(self, errors):
        self.errors = list(errors)
        super().__init__("Invalid simulation input: " + "; ".join(self.errors))


class SolverDivergenceError(SimulationError):
    """SolverDivergenceException."""


class MeshDistortionError(SimulationError):
    """MeshDistortionException."""


class MatrixNotPositiveDefiniteError(SimulationError):
    """MatrixNotPositiveDefiniteException(row, pivot)."""

    def __init__(self, row: int = -1, pivot: float = float("nan")):
        self.row = row
        super().__init__(f"Stiffness matrix not positive definite at equation {row} (pivot {pivot:.3E}). "
                         "The billet is probably unrestrained.")


class EngineUnavailableError(SimulationError):
    """EngineUnavailableException: the HIP library or device is missing."""


def raise_for(code: int, what: str, message: str):
    """Raise the reference's exception for a non-zero C-ABI return code."""
    from . import _native as N
    if code == 0:
        return
    if code == N.IFEA_ERR_NOT_SPD:
        raise MatrixNotPositiveDefiniteError(-1, float("nan"))
    if code == N.IFEA_ERR_ORDER:
        raise RuntimeError(message or "Analyze must be called first.")  # InvalidOperationException
    if code == N.IFEA_ERR_NO_DEVICE:
        raise EngineUnavailableError(message or "no HIP device")
    if code == N.IFEA_ERR_ARG and message.startswith("Tetrahedron"):
        raise MeshDistortionError(message)
    if code in (N.IFEA_ERR_ARG, N.IFEA_ERR_PERM, N.IFEA_ERR_SYMBOLIC):
        raise ValueError(f"{what} failed (status {code}): {message}")
    raise SolverDivergenceError(f"{what} failed (status {code}). {message}")
