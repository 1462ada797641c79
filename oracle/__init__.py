"""CPU oracle for the dualres HMC sampler path.

TEST INFRASTRUCTURE ONLY.  This package is a clean-room numpy/scipy
restatement of the reference's algorithms (asw221/dualres, C++/Eigen/FFTW),
used as the checker for the HIP engine in ``dualres_amd``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path never calls into this package.

Modules and the reference code they restate (paths relative to the
reference checkout):

* ``spectrum``  - embedding rule, circulant torus base, RBF / rational
  quadratic kernels, eigenvalues
  (``include/dualres/MultiResParameters.h:191-284``,
  ``include/dualres/impl/circulant_base.inl:29-89``,
  ``include/dualres/impl/kernels.inl:7-74``).
* ``mrp``       - ``MultiResParameters<T>``: complex state on the torus,
  c2c FFTs, gradient, leapfrog, energies, update, Gibbs draws
  (``include/dualres/MultiResParameters.h``).
* ``images``    - index / mask helpers
  (``include/dualres/impl/nifti_manipulation.inl``).
* ``kriging``   - sparse kriging matrix build
  (``include/dualres/kriging_matrix.h``).
* ``mce``       - minimum-contrast summary data and the weighted LSQ
  objective (``include/dualres/estimate_kernel_parameters.h``,
  ``include/dualres/perturbation_matrix.h``).
* ``hmc``       - dual-averaging step size (``include/dualres/HMCParameters.h``).

Parity status of each function (how it is pinned) is listed in the module
docstrings; functions that nothing but the restatement itself pins say
"parity unpinned".
"""
