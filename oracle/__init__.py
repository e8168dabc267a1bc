"""CPU oracle for the GSPM LLG time step -- TEST INFRASTRUCTURE ONLY.

This package is a plain FP64 NumPy/SciPy restatement of CuPyMag's
Gauss-Seidel-projection LLG step, written from the reference source read as
text.  It exists to check the HIP implementation in ``cupymag_amd``; it is
never the thing measured or shipped.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``cupymag_amd``) must not import anything from
here and fails loudly when its HIP library is missing.

Modules and the reference files they restate:

* ``oracle.params``  -- ``cupymag/core/parameters.py:60-224`` (reduced units),
  ``cupymag/core/constants.py:17-30``, ``cupymag/utils/rot_111_matrices.py``,
  ``cupymag/utils/sigma_matrices.py``.
* ``oracle.mesh``    -- ``cupymag/mesh/gridHex.py``, ``cupymag/mesh/gridTet.py``,
  ``cupymag/utils/defect_shapes.py``.
* ``oracle.fem``     -- ``cupymag/mesh/ShapeHex.py``, ``cupymag/mesh/ShapeTet.py``,
  ``cupymag/physics/assemble_{demag,Gauss_Seidel,elasticity}.py``,
  ``cupymag/utils/final_assembly.py``, ``cupymag/utils/volume_average.py``.
* ``oracle.gspm``    -- ``cupymag/solvers/linear_solvers.py`` with the CuPy CG
  stopping rule, ``cupymag/utils/compute_derivatives.py``,
  ``cupymag/core/Micromagnetics.py:109-335``.

Pinning status of each function (which independent check fixes it) is listed
in the docstring of each module.  Functions whose only check is "the same
formula retyped" are marked ``parity unpinned`` there.
"""
