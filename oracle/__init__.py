"""CPU oracle for the GeRaF RF hot paths.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py`` as the checker.  The product
(``geraf-sens_amd/``) never imports it.

* ``oracle.rf`` - ctypes bindings of ``rf_oracle.c`` (FP64 restatements of the
  reference CUDA kernels and an FP32 reference-formula form of the matched
  filter), numpy in / numpy out.
* ``oracle.lensless`` - FP64 restatement of the lensless compositing glue of
  ``rf_rendering.py`` and the per-ray signal, with torch-FP64 autograd for the
  reference-composition gradients.
* ``oracle.cpu_baseline`` - the reference formulas restated in torch-CPU FP32
  (per-term cos/sin), timed as the CPU baseline.
"""
