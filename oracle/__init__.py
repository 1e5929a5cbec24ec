"""CPU oracle for the GisAPR candidate-scoring path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``gisapr_amd``)
imports, links or calls this package; only ``tests/``, the ``smoke()`` entry
of ``__graft_entry__`` and the ``cpu_baseline`` leg of ``bench.py`` use it, and
there only as the checker (or as the CPU baseline that is timed beside the
GPU path), never as the thing measured or shipped.

The oracle restates, in NumPy, the reference's algorithm for:

* ``density_from_coords``       (pdb2mrc_gpu_ver_fp32_v3.py:41-98)
* rigid transform + 8.3 text round trip (pdb_text.py:123-150; the zip
  release's pdb_text.transformed_density_atoms)
* the integrated projector / whitening / packed template bank
  (v606:837-1155)
* packed CTF (v606:549-692), psi sampling and rotation grids (v606:1163-1256),
  particle preparation (v606:1340-1374)
* the correlation operator and per-particle records (v606:1418-1516)
* the cross-particle score  sum(exp(rawCC/1024))  (new_method..._v4.py:133-134)

where ``v606`` is
test1_with_isspa_weight_varingKK_search_translation_also_v606_torch_optimized_standalone.py.

Every function can run in float32 (mirroring the reference's arithmetic type)
or float64 (the exact value of the same algorithm).  How each function is
pinned (closed forms, brute force on tiny inputs, invariants) is recorded in
the tests under ``tests/test_oracle_*.py``; the reference ships no golden
vectors or fixtures for this path.
"""
