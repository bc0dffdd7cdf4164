"""The oracle: a CPU FP64 restatement of the reference's CDA hot path.

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this package, and only as the
checker (or the timed CPU baseline). Nothing in lumina_amd/ imports it.

The C++ restatement is oracle/cda_oracle.cpp (built by oracle/Makefile into
oracle/liblumina_oracle.so); oracle/pyoracle.py binds it with ctypes.
"""
