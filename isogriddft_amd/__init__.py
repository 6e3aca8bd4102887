"""MI355X-native kernels for the IsoGridDFT real-space Kohn-Sham inner loop.

Host side: Python on PyTorch-ROCm (device memory, streams) over the C ABI of
``libisogrid_hip.so`` (include/isogrid_hip.h).  Submodules mirror the reference's
call surfaces: ``kinetic`` (ops/kinetic.py), ``hamiltonian`` (ks/local_hamiltonian.py,
ks/static_hamiltonian.py), ``poisson`` (poisson/open_boundary.py, poisson/hartree.py),
``eigensolver`` (ks/eigensolver.py), ``pseudo`` and ``xc``.
"""

__version__ = "0.1.0"
