"""CPU oracle for the de Pina phase loop - test infrastructure only.

Only the test suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may
import this package; the product path never does.
"""
