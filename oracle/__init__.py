"""CPU FP64 oracle for the hybrid DG hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package may import, call or
execute anything under this directory.  Only ``tests/``, ``__graft_entry__.smoke``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

This is a numpy restatement of the reference solver (superboss2300/jax-fluid-sim,
Python/JAX).  The reference cannot run in this image: jax, jaxlib and meshio are
not installed and there is no network.  Each function cites the reference
``file:line`` it follows.

Pinning status of each oracle function (what checks it, besides retyping):

* elements: Kronecker delta, partition of unity, exact polynomial integration
  against closed-form moments, M symmetric/SPD, M M^-1 = I
  (tests/test_p3_accuracy.py:15-79 restated and extended to P1/P2/all types).
* geometry: volumes and face areas against closed forms, unit normals pointing
  owner -> neighbour, owner/neighbour face points coincide physically.
* solver fluxes: consistency F(U,U)=F(U)·n, HLLC isolated-contact exactness,
  mirror symmetry F(UL,UR,n) = -F(UR,UL,-n), Rusanov closed form.
* Euler residual: free stream gives R == 0 bitwise, periodic conservation,
  order of accuracy for a smooth field, brute-force single element quadrature.
* BR2: affine velocity/temperature field gives an exact gradient and zero
  interior lifts; Couette steady state.
* SSP-RK3: stability polynomial 1 + z + z^2/2 + z^3/6 on a linear problem.
* chemistry: the reference's own known-answer and invariant tests
  (tests/test_air11_chemistry.py, tests/test_reacting_coupler.py), detailed
  balance, finite-difference Jacobian.
* limiter / subcell: P1 field -> sensor 0, closed-form single-node ratio,
  reconstruct∘project = I, conservation of recompute to round-off, TVD and
  WENO traces in closed form where the subcell gradient is exact, the TVD
  trace bound.

Functions without such a check are marked "parity unpinned" in their docstring.
"""
