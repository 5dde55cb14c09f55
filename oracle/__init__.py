"""CPU oracle for the outer-product SpGEMM path.

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py use it as the correctness reference or the timed CPU baseline.  The product
(spgemm_amd/) never imports, links or calls anything here.
"""
