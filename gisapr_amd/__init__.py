"""gisapr_amd -- MI355X (gfx950) engine for GisAPR's candidate-scoring path.

PDB -> density splat, Fourier central-slice projection with CTF / FSC /
whitening, and the v606 FFT cross-correlation orientation/translation search
with its per-particle score reduction, as hand-written HIP kernels behind a
C ABI (``include/gisapr_hip.h``, ``libgisapr_hip.so``), with Python host code
that keeps the reference's call signatures:

* ``gisapr_amd.density``    -- density_from_coords / generate_volume / generate_mrc
* ``gisapr_amd.search``     -- run_search and the v606 module namespace
* ``gisapr_amd.p3d``        -- the project3d_and_whiten_cuda functions and CLI
* ``gisapr_amd.projection`` -- the grid_sample projector pieces of the search namespace
* ``gisapr_amd.engine``     -- device-resident datasets and candidate scoring
* ``gisapr_amd.batch``      -- batched candidate steps (HIP graphs, autotuned tiles)
* ``gisapr_amd.serve``, ``gisapr_amd.worker``, ``gisapr_amd.runtime`` -- serving and the worker protocol
* ``gisapr_amd.restraint``  -- the geometric restraint on the HIP splat
"""

__version__ = "0.1.0"
