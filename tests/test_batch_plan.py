"""Host-only replay of the batch engine's access extents (no GPU): the
256-case, 300-step workload lays out every case's ring buffers, chain
records, matrices and load history so that every store and load of every
step 0..300 stays inside its allocation (uvlm_batch_check_plan, the check
uvlm_batch_init also runs)."""
from uvlm_amd import _native as N
from uvlm_amd import workloads as W
from uvlm_amd.gpu import batch_check_plan


def test_c4_plan_in_bounds():
    rc, msg = batch_check_plan(W.batch_c4(steps=300), 300)
    assert rc == N.OK, msg


def test_plan_rejects_odd_columns():
    cases = W.batch_c4(cases=3, steps=10)
    cases[1] = dict(cases[1], cols=5)
    rc, msg = batch_check_plan(cases, 10)
    assert rc == N.ERR_ARG and "case 1" in msg
