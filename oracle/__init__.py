"""Oracle: CPU restatement of the reference's Lloyd hot path.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import or run anything in this directory,
and only as the checker or the CPU baseline -- never as the product path.

Contents
  lloyd_oracle.py  NumPy restatement of every reference entry point
                   (src/*.hpp, python_api/_kmeans_lloyd.cpp), each function
                   citing the reference lines it follows, plus an FP64 Lloyd
                   fit used as the accuracy truth.
  restate.c        C restatement (OpenMP) of the assignment arithmetic and of
                   the driver loop, bit-faithful to the reference's
                   per-sample window order (fused multiply-add accumulation,
                   strict-< argmin); used for label parity and as the CPU
                   baseline.
  Makefile         builds restate.c into oracle/build/liboracle.so.
  c_oracle.py      ctypes loader for liboracle.so.

The reference itself (DPC++/SYCL + dpctl) cannot be compiled or imported in
this environment, so there is no oracle/_ref build.
"""
