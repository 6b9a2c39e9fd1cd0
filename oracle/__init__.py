"""CPU oracle for the fused finite-volume time step (test infrastructure only).

This package is the checker, never the product: only ``tests/``, the smoke check in
``__graft_entry__.py`` and the ``cpu_baseline`` leg of ``bench.py`` may import it. The
product path (``geoswe_amd``) never imports it and fails loudly when its HIP library is
missing.

Modules
-------
``swe_oracle``  NumPy restatement (float32 or float64) of GeoSWE's production SRM-HLLC
                step: residual, fused update (rain, Manning friction, wet/dry), CFL
                wave-speed reduction and the dt rule, ghost-cell boundary fill, stage
                boundary and rain forcing lookup.
``flat_mesh``   Restatement of the compressed active-cell ("flat") mesh: ring-2 cross
                dilation, row-major flat order, neighbour ids and the run table that the
                HIP flat kernel reads.
``generators``  Seeded synthetic inputs (Philox value-noise fBm beds, Voronoi Manning
                classes, correlated rain) shared bit-for-bit with the device generators.

Parity status of each function is stated in its docstring ("pinned by ...").
"""
