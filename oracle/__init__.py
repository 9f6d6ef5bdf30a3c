"""CPU oracle for the belief-propagation / Vidal-gauge / simple-update path.

TEST INFRASTRUCTURE ONLY.  This package is a plain NumPy (complex128) and
SciPy-LAPACK restatement of the reference algorithm in LuchnikovI/dequantizer.
It is the checker against which the HIP implementation in ``dequantizer_amd``
is compared, and the CPU baseline timed beside it.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product package never does.

Modules and the reference files they restate:

* ``tensor_ops``   -- dequantizer/src/tensor_ops.py (all local arithmetic)
* ``graph``        -- dequantizer/src/graph/{tensor_graph,node,edge,element,
                      tensor_initializers,message_initializer}.py
* ``mappings``     -- dequantizer/src/mappings/*.py (+ the two-site density
                      matrix, which has no reference counterpart)
* ``emulator``     -- dequantizer/src/quantum_emulator/{bp,exact}_quantum_emulator.py
* ``annealing``    -- examples/quantum_annealing/quantum_annealing/src/
                      {quantum_annealer,scheduler,energy_evaluation,energy_function}.py

How each function is pinned (checked against something other than a retyped
formula) is stated in its docstring and exercised in ``tests/test_oracle_*.py``.
Random-number-valued inputs (JAX threefry keys in the reference) cannot be
reproduced without JAX; the oracle uses NumPy generators instead and those
values are "parity unpinned" by construction -- parity is always checked on
identical inputs fed to both sides.
"""
