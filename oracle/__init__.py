"""CPU oracle for the galario model-to-visibility path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library (``galario_amd``)
imports, links or calls anything in this package; only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
it, and only as the checker / the timed CPU comparison.

* ``galario_oracle`` -- numpy restatement of every stage of galario's CPU path
  (``src/galario.cpp`` of galario v1.2.2), plus an independent direct-DFT
  evaluator and brute-force checks.
* ``galario_port.c`` -- the same loops restated in C with OpenMP, built by
  ``oracle/Makefile`` into ``oracle/_build/libgalario_port.so``; ``port.py``
  drives it (FFT stage by ``scipy.fft``) as the timed CPU baseline.
"""
