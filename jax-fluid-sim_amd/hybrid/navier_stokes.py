"""hybrid/navier_stokes.py of the reference: the BR2 operator class lives in
spatial.py next to the Euler operator it extends."""

from .spatial import HybridNavierStokesOperator, polynomial_br2_penalty  # noqa: F401
