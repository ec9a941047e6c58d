"""CPU oracle for the tubular-segmentation hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of bench.py, always as the checker (or the
timed CPU baseline), never as part of the product path. The product package
``miccai2016_amd`` never imports anything from here.

* ``oracle.ref32`` - ctypes bindings to ``liboracle.so`` (miccai_oracle.c), the
  FP32 restatement of the reference kernels with the reference's loop order.
* ``oracle.fp64`` - float64 numpy restatements used to state each stage's
  FP32 error, and the host-side steps the reference does in numpy/scipy.
"""
