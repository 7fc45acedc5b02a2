"""CPU oracle of the atmosphere hot path.  TEST INFRASTRUCTURE ONLY: used by
tests/, the smoke check and bench.py's CPU-baseline leg as the checker; the
shipped package never imports it.

  ref_cpu.c / ref_cpu.py  float32 op-for-op restatement of the reference's
                          Taichi kernels (and the same code in float64);
  model_fp64.py           float64 NumPy restatement of the reference's
                          validation model (atmosphere/model.py).
"""
