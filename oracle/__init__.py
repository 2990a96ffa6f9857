"""Oracle package — TEST INFRASTRUCTURE ONLY.

``oracle.oracle``: the C restatement of the reference's numeric phase
(glu_oracle.c) plus small dense references; ``oracle.ref``: the
reference's own host code compiled from its sources into oracle/_ref/.
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
this package, and only as the checker.
"""
