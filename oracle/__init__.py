"""Test infrastructure only: CPU oracles for the Gaussian rasterisation path.

Nothing in the product (``pointrix_amd``) imports this package.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use
it, and only as the checker or the timed CPU baseline.

* ``oracle.torch64`` -- FP64 PyTorch-CPU transcription of the six msplat ops
  the reference renderer calls (pointrix/model/renderer/msplat.py:104-151) and
  of the ``MsplatRender`` host glue.  Differentiable through autograd; this is
  the tolerance target (images 1e-5 abs, gradients 1e-4 rel) and the CPU
  baseline.
* ``oracle.ref32`` -- ctypes wrapper over ``ref32.c``, an FP32 C restatement of
  the forward path with the same operation order as the HIP kernels.  This is
  the bit-exact target for the integer outputs (radius, tiles touched, sorted
  ids, tile ranges, per-pixel contributor counts).
"""
