"""The batched LU's static pivots on the CPU (gp_analyse_pivots, host only).

One pivot sequence serves every scenario of a batch, and factorisations
reuse it.  With line trips the library asks the analysis for pivots that are
large in the intact matrix and in the matrix of every distinct trip set.
N-1 screening -- one line out per scenario, every line of the IEEE 39-bus
case -- is checked by emulating the static-pivot factorisation of each
single-trip Jacobian (dt = 0, the event re-solve's matrix, from the oracle
on the network built without that line) in the analysed order: every pivot
stays within a factor 1e-4 of the largest candidate of its column.  Pivots
chosen on the intact matrix alone, or on the all-lines-out matrix alone,
hit zero pivots on some of these trips (the checks below show both), which
is what the per-trip-set rule is for.
"""
import copy
import os

import numpy as np
import pytest

from conftest import get_case

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "psse")


def _worst_pivot_ratio(q, prow, A):
    """Dense LU of A in the fixed order (rows prow, columns q): the smallest
    |pivot| / max |candidate| over the steps (0 at a zero pivot)."""
    B = A[np.ix_(prow, q)].copy()
    worst = 1.0
    for k in range(B.shape[0]):
        mx = np.max(np.abs(B[k:, k]))
        if B[k, k] == 0.0 or mx == 0.0:
            return 0.0
        worst = min(worst, abs(B[k, k]) / mx)
        B[k + 1:, k] /= B[k, k]
        B[k + 1:, k + 1:] -= np.outer(B[k + 1:, k], B[k, k + 1:])
    return worst


@pytest.fixture(scope="module")
def ieee39_trips():
    from gradpower_hip.desc import build_desc
    from gradpower_hip.network import build_network
    from oracle.oracle import Oracle
    case = get_case("case_psse", raw=os.path.join(G, "IEEE39.raw"), dyr=os.path.join(G, "IEEE39_gov.dyr"), M=8)
    J0 = Oracle(case.desc).jacobian_csc(case.desc.z0, 0.0).tocsc()
    rows = J0.indices
    cols = np.repeat(np.arange(J0.shape[1]), np.diff(J0.indptr))
    out = []
    for k in range(len(case.ps.branches)):
        ps = copy.deepcopy(case.ps)
        del ps.branches[k]
        build_network(ps)
        Jd = Oracle(build_desc(ps, case.dp)).jacobian_csc(case.desc.z0, 0.0).toarray()
        vals = Jd[rows, cols]
        dense = np.zeros_like(Jd)
        dense[rows, cols] = vals
        assert np.array_equal(dense, Jd)   # the trip only removes entries of the intact pattern
        out.append((vals, Jd))
    return J0, out


def test_ieee39_every_single_trip_keeps_its_pivots(ieee39_trips):
    from gradpower_hip.lib import analyse_pivots
    J0, trips = ieee39_trips
    assert len(trips) == 46
    q, prow = analyse_pivots(J0.indptr, J0.indices, J0.data, [v for v, _ in trips])
    assert sorted(q) == list(range(J0.shape[0])) and sorted(prow) == list(range(J0.shape[0]))
    worst = [_worst_pivot_ratio(q, prow, Jd) for _, Jd in trips]
    assert min(worst) > 1e-4, np.argsort(worst)[:5]
    assert _worst_pivot_ratio(q, prow, J0.toarray()) > 1e-4


def test_ieee39_pivots_without_the_trip_sets_fail_some_trips(ieee39_trips):
    from gradpower_hip.lib import analyse_pivots
    J0, trips = ieee39_trips
    q, prow = analyse_pivots(J0.indptr, J0.indices, J0.data)
    assert sum(_worst_pivot_ratio(q, prow, Jd) < 1e-8 for _, Jd in trips) >= 5
    allout = J0.data + sum(v - J0.data for v, _ in trips)
    q, prow = analyse_pivots(J0.indptr, J0.indices, J0.data, [allout])
    assert sum(_worst_pivot_ratio(q, prow, Jd) < 1e-8 for _, Jd in trips) >= 1
