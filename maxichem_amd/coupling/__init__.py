"""The GCM side of a chemistry call (the reference's maxichem/coupling/composition.py and chemistry.py): tracer
conversions on the GPU by fused HIP kernels (kernels/coupling.hip), the start gas, molar masses, and the stop on
unfinished cells. ``import maxichem_amd`` does not import it, as in the reference."""
