"""CPU oracle for the JaxPM particle-mesh step.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy restatement of the
reference algorithm (JaxPM ``jaxpm/painting.py``, ``painting_utils.py``,
``kernels.py``, ``pm.py``, ``ode.py``, ``growth.py``, ``distributed.py`` and
``utils.py``) used as the checker for the HIP implementation in ``jaxpm_amd``.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  Nothing in ``jaxpm_amd`` imports it, and the
product path never falls back to it.

Every function computes in float64 unless a ``dtype`` argument asks for the
float32 mirror, which replays the reference's float32 operation order where
rounding decides an index (relative painting, ``painting_utils.enmesh``).

Pinning (what each function is checked against, in ``tests/test_oracle_*.py``):

* painting: brute-force per-particle loops on 4^3/8^3 meshes, mass
  conservation, lattice identities, ``np.roll`` shift invariance, the
  paint/read transpose identity, absolute == relative painting.
* kernels: closed forms at hand-picked wave numbers; single plane waves
  through ``pm_forces`` give the analytic multiplier.
* pm: plane-wave forces, Zel'dovich displacement of a single mode, the
  2LPT source of two orthogonal modes (analytic).
* cosmology: Einstein-de Sitter limits (D1 = a, f1 = 1, D2 = a^2 after the
  reference's renormalisation, f2 = 2) and the Heath integral for LCDM D1.
* ode: free drift, time reversibility of the KDK step.
* utils.power_spectrum: brute-force binning on a tiny mesh and the spectrum
  of a single plane wave.

Parity unpinned: jax_cosmo's exact growth tables and Eisenstein-Hu
``linear_matter_power`` values, and ``jax.random.normal`` bit streams (no JAX
is installed anywhere this project runs); see the cosmology module header.
"""
from . import cosmology, kernels, ode, painting, pm, utils  # noqa: F401
