"""gisapr_amd -- MI355X (gfx950) engine for GisAPR's candidate-scoring path.

PDB -> density splat, Fourier central-slice projection with CTF / FSC /
whitening, and the v606 FFT cross-correlation orientation/translation search
with its per-particle score reduction, as hand-written HIP kernels behind a
C ABI (``include/gisapr_hip.h``, ``libgisapr_hip.so``), with Python host code
that keeps the reference's call signatures:

* ``gisapr_amd.density``   -- density_from_coords / generate_volume / generate_mrc
* ``gisapr_amd.search``    -- run_search and the v606 module namespace
* ``gisapr_amd.project3d`` -- the project3d_and_whiten_cuda functions
* ``gisapr_amd.engine``    -- device-resident datasets and candidate scoring
"""

__version__ = "0.1.0"
