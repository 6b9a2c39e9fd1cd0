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
``generators``  Seeded synthetic inputs (Philox value-noise fBm beds, Voronoi Manning
                classes, correlated rain) shared bit-for-bit with the device generators.

The flat cell order is not restated here: the product's own builder
(``geoswe_amd.flatmesh``) is pinned directly by the reference's golden ``CompressedMesh2D``
order and neighbour table (tests/test_oracle_golden.py).

Parity status of each function is stated in its docstring ("pinned by ...").
"""
