"""Seeded synthetic workloads for the RF hot paths (matched-filter scenes and
tracer ray batches).  Used by the benchmark, the smoke check and the tests;
not part of the drop-in ``geraf`` package."""
