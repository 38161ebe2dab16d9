"""FP64 host oracle for the ASTRA projection hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports, links or
executes anything under ``oracle/``; only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

The oracle restates, in float64 numpy, the sampling schemes of the reference
kernels (each function cites the reference file:line it follows):

* ``geometry``  - angle -> projection-vector generation and normalisation
* ``proj2d``    - 2D parallel / fan Joseph FP and voxel-driven BP
* ``proj3d``    - 3D cone / par3d ray-driven FP and voxel-driven BP
* ``fdk``       - FDK pre-weighting, Ram-Lak filter, row filtering, FDK/FBP
* ``algorithms``- SIRT / CGLS (2D and 3D) over the oracle projectors
* ``cpu_linear.c`` - C restatement of the CPU ``linear`` projector and CPU SIRT

Texture sampling is the *ideal* linear/bilinear interpolation that the CUDA
texture unit approximates with 8-bit weights: texel i has its centre at i+0.5,
and coordinates outside the array read 0 (border addressing).
"""
