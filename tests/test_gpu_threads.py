"""Host-pointer entry points called from several threads at once.

The host-pointer calls share one default context and its staging memory; each
call holds the context for its whole duration (staging, compute, copy-back),
so concurrent callers are serialised and every result equals the same call
made alone, bit for bit.
"""
from __future__ import annotations

import threading

import numpy as np
import pytest

import thesiscode_amd as tc

pytestmark = pytest.mark.gpu


def test_concurrent_host_calls_equal_serial_calls():
    jobs = []
    for k in range(12):
        A = tc.random_walk(900 + k, 6000 + 1500 * k)
        m = (32, 64, 100)[k % 3]
        fn = (lambda A=A, m=m: tc.selfjoin_f64(A, m)[:2],
              lambda A=A, m=m: tc.selfjoin_f32r(A, m)[:2],
              lambda A=A, m=m: tc.prescrimp_f64(A, m, 0.25)[:2])[k % 3]
        jobs.append(fn)
    serial = [fn() for fn in jobs]
    results = [None] * len(jobs)
    errors = []

    def worker(w):
        try:
            for rep in range(3):
                for k in range(w, len(jobs), 4):
                    p, i = jobs[k]()
                    if results[k] is None:
                        results[k] = (p, i)
                    assert np.array_equal(p, serial[k][0]) and np.array_equal(i, serial[k][1]), (w, k, rep)
        except Exception as e:
            errors.append(e)

    ths = [threading.Thread(target=worker, args=(w,)) for w in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join(300)
    assert not any(t.is_alive() for t in ths)
    assert not errors, errors
    assert all(r is not None for r in results)
