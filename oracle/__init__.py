"""CPU oracle for one EMC iteration -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
import this package, and only as the checker. The product package ``dragonfly_amd`` never
imports it.

* :mod:`oracle.core`  -- ctypes binding of ``liboracle.so`` (our C restatement, emc_oracle.c).
* :mod:`oracle.ref`   -- ctypes binding of ``_ref/libdfref.so``, the reference's own
  ``dragonfly/src/quaternion.c`` and ``dragonfly/src/model.c`` compiled unchanged.
* :mod:`oracle.brute` -- an independent dense numpy formulation of the iteration for tiny
  inputs, used to pin ``orc_maximize``.
"""
