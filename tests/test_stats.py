"""DDDA statistics (host code of libqmcb_amd.so, no device).

Pinned two ways:
* golden vectors written by the reference's own QMCProperty/QMCStatistic
  (tests/golden/make_golden_stats.py): average, error bar, correlation
  length, serial variance and every level's block variance, for single and
  merged (operator+) accumulations, weighted and unweighted, including the
  "< 100 samples" and "no plateau" answers;
* the mathematics of binary blocking: level i of a series holds the weighted
  means of consecutive blocks of 2^i samples with weight = mean weight of the
  block, the error bar of an AR(1) series approaches
  sigma * sqrt((1 + rho) / ((1 - rho) n)), and merging at a split that is a
  multiple of every used block size equals one sequential accumulation.
"""
import os

import numpy as np
import pytest

from qmcbeaver_amd.stats import Property, series_energy

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ddda_cases.npz")


def close(a, b, rtol=1e-13):
    a, b = float(a), float(b)
    if np.isnan(b):
        return not np.isfinite(a)
    return abs(a - b) <= rtol * max(abs(b), 1e-300)


def build(x, w, split):
    p = Property().add_series(x[:split], w[:split])
    q = Property().add_series(x[split:], w[split:])
    return p.merge(q)


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def test_ddda_matches_reference_golden(gold):
    names = [str(n) for n in gold["names"]]
    assert len(names) >= 7
    for name in names:
        src = str(gold[name + "_src"])
        x, w = gold[src + "_x"], gold[src + "_w"]
        split = int(gold[name + "_split"])
        n, avg, std, clen, svar, sstd = gold[name + "_ref"]
        p = build(x, w, split)
        s = p.summary()
        assert s.n_samples == int(n), name
        assert close(s.average, avg), (name, s.average, avg)
        assert close(s.std_dev, std), (name, s.std_dev, std)
        assert close(s.correlation_length, clen), (name, s.correlation_length, clen)
        assert close(s.serial_variance, svar), (name, s.serial_variance, svar)
        assert close(s.std_dev_std, sstd), (name, s.std_dev_std, sstd)
        bv = gold[name + "_block_var"]
        for i in range(30):
            assert close(p.level(i)[3], bv[i]), (name, i, p.level(i)[3], bv[i])


def test_golden_covers_depth_cases(gold):
    depths = set()
    for name in gold["names"]:
        src = str(gold[str(name) + "_src"])
        p = build(gold[src + "_x"], gold[src + "_w"], int(gold[str(name) + "_split"]))
        depths.add(min(p.decorr_depth, 1))
    assert depths == {-2, -1, 1}


@pytest.mark.parametrize("weighted", [False, True])
def test_levels_are_block_means(weighted):
    rng = np.random.default_rng(3)
    n = 1 << 12
    x = rng.normal(size=n)
    w = rng.uniform(0.5, 2.0, size=n) if weighted else np.ones(n)
    p = Property().add_series(x, w)
    for i in range(0, 11):
        b = 1 << i
        bw = w.reshape(-1, b)
        bx = (x.reshape(-1, b) * bw).sum(1) / bw.sum(1)
        wt = bw.mean(1)
        avg = (bx * wt).sum() / wt.sum()
        var = (bx * bx * wt).sum() / wt.sum() - avg * avg
        ni, swi, ai, bvi = p.level(i)
        assert ni == n // b
        assert abs(swi - wt.sum()) < 1e-9 * wt.sum()
        assert abs(ai - avg) < 1e-12
        assert abs(bvi - var / (ni - 1)) < 1e-9 * var / (ni - 1), i


def test_ar1_error_bar_matches_closed_form():
    rng = np.random.default_rng(11)
    n, rho = 1 << 17, 0.6
    e = rng.normal(size=n)
    x = np.empty(n)
    x[0] = e[0] / np.sqrt(1 - rho * rho)
    for i in range(1, n):
        x[i] = rho * x[i - 1] + e[i]
    s = Property().add_series(x).summary()
    sigma_x = 1.0 / np.sqrt(1 - rho * rho)
    exact = sigma_x * np.sqrt((1 + rho) / (1 - rho) / n)
    assert s.decorr_depth > 0
    assert abs(s.std_dev / exact - 1.0) < 0.2, (s.std_dev, exact)
    naive = np.sqrt(s.serial_variance / n)
    assert s.std_dev > 1.5 * naive  # correlation is visible


def test_merge_at_aligned_split_equals_sequential():
    rng = np.random.default_rng(5)
    x = rng.normal(size=5000)
    w = rng.uniform(0.2, 4.0, size=5000)
    seq = Property().add_series(x, w)
    mer = build(x, w, 4096)
    for i in range(13):
        a, b = seq.level(i), mer.level(i)
        assert a[0] == b[0]
        for u, v in zip(a[1:], b[1:]):
            assert close(u, v, 1e-14) or (not np.isfinite(u) and not np.isfinite(v)), (i, a, b)
    assert seq.std_dev == mer.std_dev


def test_merge_is_binary_addition_of_counts():
    # every level's sample count of a merge is floor-based binary addition
    rng = np.random.default_rng(6)
    for n1, n2 in [(1, 1), (3, 5), (7, 9), (100, 27), (513, 1023)]:
        x = rng.normal(size=n1 + n2)
        m = build(x, np.ones_like(x), n1)
        total = n1 + n2
        for i in range(12):
            assert m.level(i)[0] == total >> i or i == 0 and m.level(0)[0] == total, (n1, n2, i)


def test_short_series_reports_infinite_error():
    s = Property().add_series(np.arange(50.0)).summary()
    assert s.decorr_depth == -2 and s.std_dev == 1e300


def test_series_energy_weights_rows():
    rows = np.array([[-2.0, 2.0, 1.0, 2], [-3.0, 3.0, 1.5, 3], [0.0, 0.0, 0.0, 0]])
    p = series_energy(rows)
    assert p.n_samples == 2
    assert abs(p.average - (-1.0 * 2 + -1.0 * 3) / 5) < 1e-15
