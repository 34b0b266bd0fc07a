"""CPU restatements of the reference fluidLB D2Q9 time step.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, by
``__graft_entry__.smoke()`` and by the CPU-baseline leg of ``bench.py``,
only ever as the checker (or the timed CPU baseline); the engine in
``pylabolt_amd`` never imports anything from here.

* ``d2q9_oracle``: NumPy, pass by pass in the reference's operation order.
* ``d2q9_ref.cpp`` + ``ref_lib``: the same passes in C++/OpenMP (bitwise
  equal to the NumPy form), for long runs and CPU timing.

Pinned by: the reference's own unit-test fixtures (lattice constants,
boundary node indices, circle rasterisation, halo/periodic edge and corner
values), closed forms (equilibrium moments, omega = 1 and omega = 0
collisions, rest-state fixed point, mass conservation of a closed cavity,
MRT with S = omega * I equal to BGK, a pure-shift stream), and the
Ghia et al. (1982) Re=100 cavity centreline table.  See tests/test_oracle*.
"""
