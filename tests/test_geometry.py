"""Geometry of the oracle (slice_gen3d, slice_merge3d, make_rot_quat, symmetrize_friedel,
dragonfly/src/model.c) pinned to golden vectors of the compiled reference and to closed
forms: exact trilinear reproduction of an affine model, rotation orthonormality, merge
weights summing to one, Friedel invariance."""
import os

import numpy as np
import pytest

from oracle import brute

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, 'golden', 'geometry.npz'))
TAGS = ('odd', 'even')


@pytest.mark.parametrize('tag', TAGS)
def test_slice_gen_golden(oracle_lib, tag):
    for q, v in zip(G[tag + '_quats'], G[tag + '_views']):
        got = oracle_lib.slice_gen3d(q, G[tag + '_model'], G[tag + '_qvals'], G[tag + '_corr'])
        assert np.array_equal(got.view('u8'), v.view('u8'))


@pytest.mark.parametrize('tag', TAGS)
def test_slice_merge_golden(oracle_lib, tag):
    S = G[tag + '_model'].shape[0]
    m2 = np.zeros((S,) * 3)
    w2 = np.zeros((S,) * 3)
    fac = np.linspace(0.5, 2., len(G[tag + '_corr']))
    for q, v in zip(G[tag + '_quats'], G[tag + '_views']):
        oracle_lib.slice_merge3d(q, v * fac, m2, w2, G[tag + '_qvals'], G[tag + '_corr'], G[tag + '_mask'])
    assert np.array_equal(m2.view('u8'), G[tag + '_merge_model'].view('u8'))
    assert np.array_equal(w2.view('u8'), G[tag + '_merge_weight'].view('u8'))


@pytest.mark.parametrize('tag', TAGS)
def test_rot_and_symmetrize_golden(oracle_lib, tag):
    for q, rot in zip(G[tag + '_quats'], G[tag + '_rots']):
        assert np.array_equal(oracle_lib.make_rot_quat(q), rot)
    a = G[tag + '_sym_in'].copy()
    oracle_lib.symmetrize_friedel(a, G[tag + '_sym_w'])
    assert np.array_equal(a.view('u8'), G[tag + '_sym_out'].view('u8'))


def test_rotation_closed_forms(oracle_lib):
    assert np.array_equal(oracle_lib.make_rot_quat([1., 0., 0., 0.]), np.eye(3))
    rng = np.random.default_rng(3)
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        R = oracle_lib.make_rot_quat(q)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) - 1.) < 1e-13
        # the dense formulation used by the brute force is the same convention
        assert np.allclose(R, brute.rotation(q), atol=1e-15)
    # 180 degrees about z flips x and y
    assert np.allclose(oracle_lib.make_rot_quat([0., 0., 0., 1.]), np.diag([-1., -1., 1.]))


def test_trilinear_reproduces_affine_model(oracle_lib):
    S = 17
    c = S // 2
    x = np.arange(S)
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    a = np.array([0.3, -0.7, 1.1])
    model = 5. + a[0] * X + a[1] * Y + a[2] * Z
    rng = np.random.default_rng(1)
    qv = rng.normal(size=(300, 3))
    qv *= (rng.uniform(0, 6, 300) / np.linalg.norm(qv, axis=1))[:, None]
    corr = rng.uniform(0.5, 1.5, 300)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    view = oracle_lib.slice_gen3d(q, model, qv, corr)
    p = qv @ oracle_lib.make_rot_quat(q).T + c
    assert np.allclose(view, (5. + p @ a) * corr, rtol=1e-13)


def test_out_of_grid_is_dbl_min(oracle_lib):
    S = 9
    model = np.ones((S,) * 3)
    # center 4: cell = trunc(q + 4), in grid for 0 <= cell <= S-2 = 7
    qv = np.array([[0., 0., 0.], [100., 0., 0.], [0., -4.6, 0.], [0., 0., 3.5], [0., 0., 4.5],
                   [0., -5.2, 0.]])
    view = oracle_lib.slice_gen3d([1., 0., 0., 0.], model, qv, np.full(6, 2.))
    tiny = np.finfo(float).tiny
    # p = -0.6 truncates to cell 0 and is interpolated (weights 1.6 and -0.6 of a flat model)
    assert view[0] == 2. and view[2] == 2. and view[3] == 2.
    # p = 104, 8.5 (cell 8 > 7) and -1.2 (cell -1) are outside: DBL_MIN, not multiplied by corr
    assert view[1] == tiny and view[4] == tiny and view[5] == tiny


def test_merge_weights_sum_to_one(oracle_lib):
    S = 13
    rng = np.random.default_rng(2)
    qv = rng.uniform(-4, 4, (50, 3))
    corr = np.ones(50)
    mask = np.zeros(50, 'u1')
    m2 = np.zeros((S,) * 3)
    w2 = np.zeros((S,) * 3)
    view = rng.random(50)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    oracle_lib.slice_merge3d(q, view, m2, w2, qv, corr, mask)
    assert abs(w2.sum() - 50.) < 1e-12
    assert abs(m2.sum() - view.sum()) < 1e-12
    mask[:10] = 2
    m3 = np.zeros((S,) * 3)
    w3 = np.zeros((S,) * 3)
    oracle_lib.slice_merge3d(q, view, m3, w3, qv, corr, mask)
    assert abs(w3.sum() - 40.) < 1e-12


def test_friedel_invariance(oracle_lib):
    rng = np.random.default_rng(4)
    for S in (11, 12):
        a = rng.random((S, S, S))
        w = rng.random((S, S, S)) + 0.1
        oracle_lib.symmetrize_friedel(a, w)
        lo = 1 if S % 2 == 0 else 0
        sub = a[lo:, lo:, lo:]
        assert np.allclose(sub, sub[::-1, ::-1, ::-1], rtol=1e-13)
