"""Host logic of bench.py that needs no GPU: the one-GPU refusal, the rule that a measured
traffic figure is used only for the kernel sources it was measured on, and the units of the
roofline of the judged proposal iteration (bytes per launch over milliseconds against the
8 TB/s HBM peak)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert 'one GPU' in r.stderr
    assert r.stdout == ''


def _traffic_table():
    with open(os.path.join(ROOT, 'profiles', 'pmc_traffic.json')) as fh:
        return json.load(fh)


def test_traffic_only_for_the_sources_it_was_measured_on(monkeypatch):
    tab = _traffic_table()
    ent = tab['cfg5']
    kernel = 'merge_brick_model_kernel'
    monkeypatch.setattr(bench, 'kernel_source_hash', lambda: ent['_src_sha16'])
    bytes_, note = bench.load_traffic('cfg5', kernel)
    assert bytes_ == pytest.approx(float(ent[kernel]['bytes_per_launch']))
    monkeypatch.setattr(bench, 'kernel_source_hash', lambda: '0' * 16)
    bytes_, note = bench.load_traffic('cfg5', kernel)
    assert bytes_ is None and ent['_src_sha16'] in note
    assert bench.load_traffic('cfg5', 'no_such_kernel')[0] is None
    assert bench.load_traffic('no_such_workload', kernel)[0] is None


def test_source_hash_covers_every_kernel_source(tmp_path, monkeypatch):
    """The hash changes with any .hip, .h or .cpp file under csrc (and only with those)."""
    csrc = tmp_path / 'dragonfly_amd' / 'csrc'
    csrc.mkdir(parents=True)
    (csrc / 'a.hip').write_text('x')
    (csrc / 'b.h').write_text('y')
    (csrc / 'notes.txt').write_text('z')
    monkeypatch.setattr(bench, 'ROOT', str(tmp_path))
    h0 = bench.kernel_source_hash()
    (csrc / 'notes.txt').write_text('changed')
    assert bench.kernel_source_hash() == h0
    (csrc / 'b.h').write_text('y2')
    h1 = bench.kernel_source_hash()
    assert h1 != h0
    (csrc / 'c.cpp').write_text('')
    assert bench.kernel_source_hash() != h1


def test_roofline_of_the_judged_iteration_units():
    """A hand-sized case: 64 slots x 64 pixels (one 64-pixel patch), a 10^3 volume (3^3
    bricks of 4^3 cells), 10^9 nonzero view entries, no fall-back frames. The merge moves
    8 x 64 B of masks + 4 x 10^9 B of views + 16 x 1000 B of volume = 4,000,016,512 B in
    1 ms, i.e. 4000 GB/s, half the peak; the expand 512 x 27 + 4 x 4096 = 30,208 B in
    0.5 us, i.e. 60.4 GB/s."""
    res = dict(stage_ms={'tomogram': 0., 'reduce': 0.25, 'merge': 1.0},
               pass1_stage_ms={'expand': 0.5e-3}, view_nonzero=1e9, fallback_rate=0.)
    out = bench.roofline_learned(res, slots=64, num_pix=64, size=10, num_photons=500, num_frames=100,
                                 workload=None)
    assert out['stage'] == 'merge' and out['kernel'] == 'merge_brick_model_kernel'
    assert out['compulsory_bytes'] == 4000016512
    assert out['achieved'] == pytest.approx(4000.0, abs=0.05)
    assert out['frac'] == pytest.approx(0.5, abs=1e-4)
    assert out['traffic'] is None
    st = out['stages']
    assert st['expand']['compulsory_bytes'] == 512 * 27 + 4 * 64 * 64
    assert st['expand']['frac'] == pytest.approx(60.416 / 8000., abs=1e-4)
    # no fall-back frames: the pair tomogram reads no partial views
    assert st['reduce']['compulsory_bytes'] == 12 * 500 + 8 * 64 + 4 * 10 ** 9
    assert st['tomogram']['frac'] is None


def test_cpu_baseline_fit_on_a_tiny_problem():
    """bench.cpu_baseline end to end on the tiny geometry (3240 orientations, 60 frames, the
    oracle on 2 threads at three orientation subsets of 32 frames): each run's split adds up
    to its wall time, the fitted coefficients are non-negative, and the reported rate is the
    workload's pairs over the fitted time T = alpha N + n (beta N + gamma N^2) + c."""
    from types import SimpleNamespace
    import numpy as np
    import problems as P
    p = P.build(P.tiny_config(num_data=60, num_div=4))
    fp = p['fparams']
    R = len(p['quats'])
    itr = SimpleNamespace(quat=SimpleNamespace(num_rot=R, quats=p['quats']), tot_num_data=60,
                          blacklist=np.zeros(60, 'u1'), beta_start=fp['beta'], fcounts=fp['fcounts'],
                          sum_fact=fp['sum_fact'], mean_count=[fp['mean_count']],
                          model=SimpleNamespace(model1=p['model1']), params=SimpleNamespace(need_scaling=0))
    value, sample, fit = bench.cpu_baseline(p, itr, threads=2, target_rot=(400, 800, 1600), frames=32)
    pts = fit['points']
    assert [q['orientations'] for q in pts] == [len(p['quats'][::k]) for k in (8, 4, 2)]
    for q in pts:
        assert q['frames'] == 32
        parts = q['per_orientation_s'] + q['per_frame_s'] + q['other_s']
        assert parts == pytest.approx(q['seconds'], abs=5e-3)
    a, b, g, c = fit['alpha_s'], fit['beta_s'], fit['gamma_s'], fit['c_s']
    assert a > 0. and b >= 0. and g >= 0.
    # the reported coefficients carry 4 significant digits, c and the time are rounded
    t = a * R + 60 * (b * R + g * R * R) + c
    assert fit['extrapolated_s'] == pytest.approx(t, rel=1e-3, abs=0.051)
    assert value == pytest.approx(60 * R / t, rel=1e-2)
    assert '%d frames' % 32 in sample and '%d orientations' % R in sample
