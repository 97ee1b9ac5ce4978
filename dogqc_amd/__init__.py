"""dogqc_amd: an MI355X-native (gfx950) execution backend for DogQC plans.

The plan API (``relationalAlgebra``, ``scalarAlgebra``, ``types``, ``util``,
``dbio``) keeps the reference's names and semantics; ``HipCompiler`` in
``hipTranslator`` replaces the CUDA translator and emits one HIP module per
plan, run through the C ABI of ``include/dogqc_amd.h`` (``runtime``).
"""
