"""ORACLE — test infrastructure only (CPU restatement of the reference
algorithm).  Imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py; never by the commit_amd package."""
