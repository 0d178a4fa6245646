"""CPU oracle: an FP64 restatement of the reference's solve/simulate algorithms.

TEST INFRASTRUCTURE ONLY. Only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` may import anything from this package, and
only as the checker or the timed CPU baseline. The product (`pylcm_amd`) never
imports it and has no CPU fallback.

Each function cites the reference file:line it restates. Parity status of every
function (which golden vectors, closed forms or brute-force checks pin it) is
listed in the module docstrings and in the tests that pin them.
"""
