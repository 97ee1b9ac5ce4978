"""Launch-shape knobs of generated pipeline kernels.

``KernelCall.defaultGridSize`` / ``defaultBlockSize`` keep the class-level
knobs of the reference (dogqc/kernel.py:10-11) that drivers set before
translation.  The block size is a codegen constant (it sizes LDS buffers and
``__launch_bounds__``) and must be a multiple of the 64-lane wavefront.  A grid
size of 0 means "derive from the row count and the device's CU count at
launch"; per-pipeline grids can also be overridden at run time through
``dqc_plan_tuning.grid``.  ``defaultRowsPerThread`` (no reference
counterpart) sets how many rows a thread carries through a pipeline at once.
"""


class KernelCall(object):
    defaultGridSize = 0
    defaultBlockSize = 256
    # rows each thread processes per loop trip in pipelines without inner
    # loops (independent rows give the memory system more requests in flight)
    defaultRowsPerThread = 4
