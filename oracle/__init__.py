"""CPU oracle of the reference's TGV-DCT deshade path: test infrastructure only
(tests/, __graft_entry__.smoke() as the checker, bench.py's cpu_baseline)."""
