"""Orientation sampling (quat_gen, dragonfly/src/quaternion.c:811-842).

Pins: the committed sha256 of the reference's own output (tests/golden/quat_sha256.json,
written from the compiled reference), the closed-form count 10(5n^3+n), unit norm, weight
sum 1 and the half-sphere rule. Both the oracle restatement and the product library's
host-side generator must match the golden bytes exactly.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from dragonfly_amd import _lib
from dragonfly_amd.quaternion import Quaternion

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, 'golden', 'quat_sha256.json')) as _f:
    GOLDEN = json.load(_f)
FAST = [n for n in GOLDEN if int(n) <= 10]


def _sha(q):
    return hashlib.sha256(np.ascontiguousarray(q).tobytes()).hexdigest()


@pytest.mark.parametrize('n', FAST)
def test_oracle_matches_golden(oracle_lib, n):
    q = oracle_lib.quat_gen(int(n))
    assert q.shape[0] == GOLDEN[n]['num_rot']
    assert _sha(q) == GOLDEN[n]['sha256']


@pytest.mark.parametrize('n', list(GOLDEN))
def test_product_matches_golden(n):
    q = _lib.quat_gen(int(n))
    assert q.shape[0] == GOLDEN[n]['num_rot']
    assert _sha(q) == GOLDEN[n]['sha256']
    for i, row in GOLDEN[n]['rows'].items():
        assert q[int(i)].tolist() == row


@pytest.mark.parametrize('n', [1, 2, 3, 4, 5, 8])
def test_closed_forms(n):
    q = Quaternion(n).quats
    assert q.shape == (10 * (5 * n ** 3 + n), 5)
    assert np.allclose(np.linalg.norm(q[:, :4], axis=1), 1., atol=1e-15)
    assert abs(q[:, 4].sum() - 1.) < 1e-16 * q.shape[0] * 10
    assert (q[:, 4] > 0).all()
    # one of each +-q pair: no two samples are antipodal or equal
    if n > 5:
        return
    d = np.abs(q[:, :4] @ q[:, :4].T)
    np.fill_diagonal(d, 0.)
    assert d.max() < 1. - 1e-9


def test_vertices_are_600_cell():
    q = Quaternion(1).quats
    assert q.shape[0] == 60
    # 120 vertices of the 600-cell = 60 rotations; each has 12 nearest neighbours at tau/2
    d = q[:, :4] @ q[:, :4].T
    tau = (np.sqrt(5.) + 1.) / 2.
    nn = (np.abs(np.abs(d) - tau / 2.) < 1e-12).sum(1)
    assert (nn == 12).all()
    assert np.allclose(q[:, 4], 1. / 60)


def test_matches_compiled_reference_when_present():
    from oracle import ref
    if not ref.available():
        pytest.skip('oracle/_ref not built (reference sources absent)')
    for n in (1, 3, 6):
        assert np.array_equal(ref.quat_gen(n).view('u8'), _lib.quat_gen(n).view('u8'))


def test_divide_uses_integer_division():
    q = Quaternion(2)
    assert q.divide(0, 3) == 140 and q.divide(2, 3) == 140
    assert isinstance(q.divide(0, 1), int)
    q2 = Quaternion(1)
    assert [q2.divide(r, 7) for r in range(7)] == [9, 9, 9, 9, 8, 8, 8]


with open(os.path.join(HERE, 'golden', 'voronoi_sha256.json')) as _f:
    VORONOI = json.load(_f)


@pytest.mark.parametrize('key', list(VORONOI))
def test_voronoi_subset_matches_golden(oracle_lib, key):
    """Nearest coarse orientation of every fine one (quaternion.c:844-862): the product's and
    the oracle's mapping equal the compiled reference's bytes, and every fine orientation's
    cell maximises |q_f . q_c| (brute force over all coarse orientations)."""
    c, f = (int(x) for x in key.split('_'))
    qc, qf = _lib.quat_gen(c), _lib.quat_gen(f)
    m = _lib.voronoi_subset(qc, qf)
    assert m.size == VORONOI[key]['num_fine']
    assert hashlib.sha256(m.astype('<i4').tobytes()).hexdigest() == VORONOI[key]['sha256']
    assert np.array_equal(oracle_lib.voronoi_subset(qc, qf), m)
    rows = np.random.default_rng(3).choice(len(qf), min(len(qf), 1500), replace=False)
    dots = np.abs(sum(qf[rows, k][:, None] * qc[None, :, k] for k in range(4)))
    assert np.allclose(dots[np.arange(len(rows)), m[rows]], dots.max(1), rtol=0, atol=1e-15)
    # every coarse orientation is its own cell's member when it is also a fine one
    assert np.bincount(m, minlength=len(qc)).min() >= 1
