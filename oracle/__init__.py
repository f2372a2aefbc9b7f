"""CPU oracle for the PyExner per-step hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``pyexner_amd`` imports this package; only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and only as the checker (never as the code measured or
shipped).

It is a numpy restatement, line by line, of the reference JAX functions on the
path (file:line citations are to the PyExner source tree ``src/PyExner/``):

* ``roe_exner.py``   - solvers/kernels/roe_exner.py:14-717
* ``roe.py``         - solvers/kernels/roe.py:13-315
* ``boundaries.py``  - domain/boundary_registry.py:19-187 and the handler files
  under domain/boundaries/
* ``steps.py``       - solvers/roe_exner_solver.py:21-122, solvers/roe_solver.py,
  integrators/forwardeuler.py:32-132, integrators/SSPRK2.py:30-51
* ``closed_forms.py`` - analytic solutions used as pins (Stoker dam break,
  Berthon constant-flux Exner case, MPM closed form)

Every function is dtype-generic: it computes in the dtype of its array inputs
(float64 for the parity oracle, float32 for the CPU baseline).  Python float
constants are weakly typed (NumPy >= 2 / NEP 50), exactly as JAX treats them.

Integer powers follow jax.lax.integer_pow (``x**2 == x*x``,
``x**3 == x*(x*x)``); non-integer powers use ``np.power``.

Pins (what checks each function; see tests/test_oracle_*.py):

* roe_exner.roe_edge         - Roe property (flux jump), C-property (lake at rest),
  rotation consistency, wall symmetry
* roe_exner.get_lambda       - numpy.roots of the coded cubic, ordering, linear
  fallback, clip branch closed form
* roe_exner.compute_dt_2D    - brute force over tiny grids (python loops)
* roe_exner.compute_G        - MPM closed form; zero below critical Shields
* roe_exner.compute_n        - Strickler closed form
* roe_exner.exner_*          - Berthon analytic case, sediment conservation,
  frozen bed when G=0, upwind selection
* roe_exner.momentum_corrections, boundaries.* - brute force on tiny grids
* roe.*                      - Stoker dam break closed form, C-property,
  y-symmetry of column-uniform data
* steps.ssprk2_step          - closed form 1/2(u + FE(FE(u))), second order
  convergence on a smooth problem

Parity unpinned: the reference could not be imported here (jax is absent), so
no function is pinned against reference-produced vectors; last-ulp behaviour of
XLA's transcendental functions is not reproduced.
"""
