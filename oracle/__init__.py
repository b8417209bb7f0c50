"""CPU oracle for the CommonNN fit path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this package.  ``cnn_oracle`` is the numpy/Python restatement;
``c_oracle`` wraps the C restatement in ``cnn_oracle.c``.
"""
