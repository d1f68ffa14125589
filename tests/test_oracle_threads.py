"""The oracle's threaded element loops (ora_set_threads, the all-cores CPU
baseline) give the serial results bit for bit: a layered 3D substep with
ground contacts and the stop rule, run with 1 and 4 threads."""
import copy

import numpy as np

import oracle as O
from helpers import beam_case


def test_threaded_oracle_is_bit_identical(oracle_lib):
    m, it, contact = beam_case(12, 4, 4, pct=99.99, z0=0.002)
    it.useRigidContacts = False
    runs = []
    for threads in (1, 4):
        O.lib().ora_set_threads(threads)
        try:
            osim = O.OracleSim(m, copy.deepcopy(it), 1 / 60, contact=contact)
            out = [osim.step() for _ in range(6)]
        finally:
            O.lib().ora_set_threads(1)
        runs.append((osim.p.copy(), osim.v.copy(), osim.lambdas(), out))
    (p1, v1, l1, o1), (p4, v4, l4, o4) = runs
    assert np.array_equal(p1, p4) and np.array_equal(v1, v4) and np.array_equal(l1, l4)
    for a, b in zip(o1, o4):
        assert a["iters"] == b["iters"]
        assert np.array_equal(a["residuals"], b["residuals"])
    assert sum(len(a["iters"]) for a in o1) > 6  # layered substeps ran
