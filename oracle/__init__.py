"""ORACLE — test infrastructure, not product code.

CPU restatements of the reference's extensive scan used as the checker by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The
product package (beguidedoff_amd) never imports this.
"""
