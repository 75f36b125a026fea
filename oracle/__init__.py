"""FP64 CPU oracle for the real-time TDDFT split-operator step.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package (``ggbtddft_amd``)
imports, links or executes anything under ``oracle/``.  The only legitimate
users are ``tests/``, ``__graft_entry__.smoke()`` (as the checker of one tiny
GPU invocation) and the ``cpu_baseline`` leg of ``bench.py``.

The oracle is a plain numpy restatement of the equations the GPU path
implements.  Where the equations follow the reference (graham-g-brown/ggbTDDFT)
each function cites the reference file:line it restates; where the equations
are this project's own choice (its scheme: 2x2 laser rotations,
theta-space potentials, Hartree multipoles by radial scans, LDA Slater+PW92,
imaginary-time ground state) the docstring says so and names the closed forms
or published results that pin it.

Modules
-------
mesh        Gauss-Lobatto DVR radial grid and kinetic matrix
angular     normalised associated Legendre functions, Gauss-Legendre theta grid
xc          LDA exchange (Slater) and correlation (Perdew-Wang 1992)
hartree     multipole Hartree potential: direct O(N^2) sum and O(N) scans
atoms       all-electron orbital sets grouped by |m|
field       sin^2 one- and two-colour pulses
groundstate imaginary-time FP64 ground state with Gram-Schmidt
propagate   the real-time step, observables, full runs
spectrum    HHG spectra

Parity pinning (what each function is checked against) is listed in the
module docstrings and exercised in ``tests/test_oracle_*.py``.
"""
