"""Test infrastructure: the oracle (a CPU restatement of the reference path
in C, ``liboracle.so``) and the compiled reference harness
(``_ref/libcadise_ref.so``). Only tests/, bench.py's CPU-baseline leg and
__graft_entry__.smoke() import this package, as the checker."""
