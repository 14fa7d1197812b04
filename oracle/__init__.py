"""FP64 CPU restatement of dft-zig's plane-wave H·psi + block eigensolver path.

TEST INFRASTRUCTURE ONLY.  This package is the correctness oracle: it restates
the reference algorithm in numpy, function by function, each citing the
reference file:line it follows (paths relative to the dft-zig source tree).
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the timed CPU
baseline.  The product (``libdzhip.so`` and its Python binding ``dzhip``)
never imports, links or calls anything in here.

How each function is pinned (so that a plausible mistake fails a check):

* FFT sign/normalisation (``fftconv``): brute-force DFT on small sizes and the
  two closed-form tests of the reference (delta-psi and direct convolution,
  ``src/tests.zig:872-1106``).
* Index maps (``pw``): exhaustive comparison with an independent
  frequency-arithmetic restatement, gather(scatter(x)) == x, and the grid
  sizes the reference's own sizing rules give (32^3 for Si-8 at 20 Ry).
* Projector tables (``projectors``): spherical Bessel closed forms
  (``src/tests.zig:1247+``), the Y_lm addition theorem
  (``pseudopotential/nonlocal.zig:151``), Gaussian-Hankel closed forms for the
  radial integral, angular-factor values (``src/tests.zig:1218-1245``).
* Local and nonlocal apply (``hamiltonian``): the dense-matrix brute force
  built independently from V(G-G') and the Legendre-polynomial form of the
  nonlocal matrix (``hamiltonian/hamiltonian.zig:174-374``), Hermiticity.
* Eigensolver (``eigensolver``): exact spectra of the reference's diagonal
  test operators (``linalg/iterative.zig:80-203``) and ``numpy.linalg.eigh``
  of the dense H.
* Density (``density``): integral of rho equals the electron count and a
  brute-force |sum_G c_G e^{iG.r}|^2 on a tiny grid.

Parity that only running the reference could pin (Davidson trajectory,
iteration counts, FFTW's rounding) is "parity unpinned"; converged
eigenvalues and applied vectors do not depend on it.
"""
