"""CPU oracle for ParaGraph's GPU index-build stage.

TEST INFRASTRUCTURE ONLY.  Importable from tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg, where it is the checker (or the timed CPU
baseline), never the thing measured or shipped.  The product package
`paragraph_amd` never imports this package.
"""
