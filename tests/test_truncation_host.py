"""Truncation rules of split_svd on hand-computed spectra (no device).

TruncationScheme (Backend.jl:466-483) composes MatrixAlgebraKit's truncrank,
trunctol and truncerror with ``&``.  MatrixAlgebraKit 0.6.9 is not in the
reference tree, so each rule is pinned against its documented definition on
spectra where a plausible mistake (a strict inequality, rtol relative to the
largest value instead of the 2-norm, an off-by-one in the discarded tail)
gives a different rank.
"""
import numpy as np
import pytest

from graft_amd import backend as B


S = np.array([4.0, 3.0, 1.0, 0.5, 0.1, 0.01, 1e-3])
NRM = float(np.linalg.norm(S))          # 5.1225...


def rank(**kw):
    return B.truncation_rank(S, B.TruncationScheme(**kw))


def test_no_truncation_keeps_everything():
    assert rank() == len(S)


def test_maxdim():
    assert rank(maxdim=3) == 3
    assert rank(maxdim=100) == len(S)


def test_atol_keeps_values_equal_to_the_threshold():
    assert rank(atol=0.5) == 4          # 0.5 >= 0.5 is kept
    assert rank(atol=0.50001) == 3
    assert rank(atol=1e-12) == len(S)


def test_rtol_is_relative_to_the_two_norm():
    # 0.7 * ||s|| = 3.586: keeps only 4.0 (0.7 * s_max = 2.8 would keep 3.0 as well)
    assert rank(rtol=0.7) == 1
    assert rank(rtol=3.0 / NRM * (1 - 1e-12)) == 2
    # atol and rtol: the larger threshold wins
    assert rank(atol=0.2, rtol=0.1) == 3     # max(0.2, 0.512) keeps 4, 3, 1
    assert rank(atol=0.6, rtol=0.01) == 3    # max(0.6, 0.051) keeps 4, 3, 1


def test_discarded_weight_keeps_the_fewest_values():
    # the tail 2-norms from index i on
    tail = np.sqrt(np.cumsum((S ** 2)[::-1]))[::-1]
    for i in range(1, len(S)):
        w = tail[i] / NRM
        assert rank(discarded_weight=w) == i               # exactly at the bound: allowed
        assert rank(discarded_weight=w * (1 - 1e-9)) == i + 1 if i + 1 <= len(S) else True


def test_rules_intersect():
    assert rank(maxdim=2, atol=0.05) == 2
    assert rank(maxdim=6, atol=0.05) == 5
    # tail from index 3 has norm 0.5101 > 0.5, from index 4 0.1005: keep 4
    assert rank(maxdim=6, discarded_weight=0.5 / NRM) == 4
    assert rank(maxdim=3, discarded_weight=0.5 / NRM) == 3


def test_at_least_one_value_is_kept():
    assert B.truncation_rank(np.zeros(4), B.TruncationScheme(atol=1e-3)) == 1
    assert rank(atol=100.0) == 1


@pytest.mark.parametrize("n", [1, 5])
def test_every_rank_reachable_by_maxdim(n):
    s = np.sort(np.random.default_rng(n).random(n))[::-1]
    for k in range(1, n + 1):
        assert B.truncation_rank(s, B.TruncationScheme(maxdim=k)) == k
