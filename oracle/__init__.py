"""CPU FP64 oracle for the IsoGridDFT Kohn-Sham inner loop.

TEST INFRASTRUCTURE ONLY.  This package is a plain NumPy/SciPy restatement of the
reference algorithms (jiangweiqi001/IsoGridDFT, files cited per function as
``path:line`` relative to the reference's ``src/isogrid``).  It is imported only by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``, and there only as the checker or the timed CPU baseline.  The product
package ``isogriddft_amd`` never imports it.

Parity pinning of each function (what fixes it independently of this restatement)
is listed in the module docstrings.  Functions with no independent check say
"parity unpinned" in their docstring.
"""

from . import grid, kinetic, pseudo, xc, poisson, eigen, energy  # noqa: F401
