"""CPU oracle of the receive tail -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker (or the timed
CPU baseline), never as part of the product: ``commkit_amd`` never imports
it and has no CPU compute path.

- :mod:`oracle.cpu`   ctypes wrappers of ``libck_oracle.so`` (plain-C float32
  restatement of the reference's Numba kernels, built by ``oracle/Makefile``);
- :mod:`oracle.llr64` FP64 LLRs from the definition (brute force over points);
- :mod:`oracle.ldpc64` FP64 textbook layered and flooding belief propagation.
"""
