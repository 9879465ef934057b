"""Test infrastructure only: the oracle of the osinco3d time-step path.

Two CPU implementations live here, both used only by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``:

* ``liboracle.so`` - a plain-C restatement of the reference algorithms
  (o3d_oracle.c), pinned by the tests against the items below;
* ``_ref/libo3dref.so`` and ``_ref/o3d_ref_run`` - the reference's own Fortran
  modules compiled from its sources (Makefile); absent when the reference tree
  was not available at build time.

Nothing in ``osinco3d_amd`` imports this package.
"""
from .oracle import Oracle, Reference, load_oracle, load_reference, ref_run_path  # noqa: F401
