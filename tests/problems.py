"""Small seeded EMC problems shared by the CPU and GPU tests, plus runners that feed the
same inputs to the oracle (CPU) and to the engine (GPU)."""
import numpy as np

from dragonfly_amd import synth
from dragonfly_amd.iterate import frame_counts, sum_log_factorials

SPARSE_KEYS = ('ones', 'multi', 'ones_accum', 'multi_accum', 'place_ones', 'place_multi',
               'count_multi')


def tiny_config(size=21, nx=22, ny=20, atoms=6, num_data=60, mean_count=40., num_div=2,
                phi=None, gap=0, stoprad=2.):
    return dict(size=size, atoms=atoms, atom_radius=max(2., size / 6.), sigma=1.2, nx=nx, ny=ny,
                gap=gap, stoprad=stoprad, num_data=num_data, mean_count=mean_count, phi=phi,
                num_div=num_div, need_scaling=0, beta=1.0, iters=2)


def build(cfg, quats=None, start_seed=4, num_ave=500, device='cpu'):
    """Detector, frames, start model, per-frame arrays and orientations for a config dict."""
    size = cfg['size']
    det = synth.flat_detector(cfg['nx'], cfg['ny'], qmax=(size - 3) / 2. - 0.05,
                              stoprad=cfg['stoprad'], gap=cfg['gap'])
    assert synth.model_size(det['qvals']) == size
    intens = synth.make_intensity(size, cfg['atoms'], cfg['atom_radius'], cfg['sigma'], seed=0)
    frames = synth.simulate_frames(intens, det, cfg['num_data'], cfg['mean_count'], phi=cfg['phi'],
                                   num_ave=num_ave, batch=256, device=device)
    if quats is None:
        from dragonfly_amd import _lib
        quats = _lib.quat_gen(cfg['num_div'])
    mask = det['mask']
    fc = frame_counts(frames['ones'], frames['multi'], frames['place_ones'], frames['place_multi'],
                      frames['count_multi'], mask)
    sf = sum_log_factorials(frames['multi'], frames['place_multi'], frames['count_multi'], mask)
    D = len(frames['ones'])
    mean_count = fc.astype('f8').sum() / D
    model1 = synth.start_model(size, mean_count, mask, seed=start_seed)
    fparams = dict(scale=np.ones(D), beta=np.full(D, float(cfg.get('beta_value', 1.0))),
                   blacklist=np.zeros(D, 'u1'), fcounts=fc, sum_fact=sf, mean_count=mean_count,
                   num_blacklist=0)
    return dict(det=det, frames={k: frames[k] for k in SPARSE_KEYS}, quats=quats, model1=model1,
                fparams=fparams, intens=intens, true_scale=frames['true_scale'])


def run_oracle(p, flags, model1=None, dense_prob=True, num_threads=1):
    from oracle import core
    fp = dict(p['fparams'])
    fp['scale'] = np.array(fp['scale'], dtype='f8')
    out = core.maximize(p['det'], p['quats'], p['model1'] if model1 is None else model1,
                        p['frames'], fp, flags, num_threads=num_threads, dense_prob=dense_prob)
    out['scale'] = fp['scale']
    return out


def engine_for(p):
    from dragonfly_amd._lib import Engine
    eng = Engine(0)
    eng.set_detector(p['det']['qvals'], p['det']['corr'], p['det']['mask'])
    eng.set_quats(p['quats'])
    eng.set_model(p['model1'])
    f = p['frames']
    eng.set_frames(*[f[k] for k in SPARSE_KEYS])
    fp = p['fparams']
    eng.set_frame_params(fp['scale'], fp['beta'], fp['blacklist'], fp['fcounts'], fp['sum_fact'],
                         fp['mean_count'], fp['num_blacklist'])
    return eng


def run_engine(eng, flags):
    st = eng.maximize(flags.get('iteration', 1), flags.get('need_scaling', 0),
                      flags.get('update_scale', 1), flags.get('known_scale', 0))
    out = {k: eng.get(k) for k in ('u', 'max_exp', 'p_norm', 'rmax', 'likelihood', 'info',
                                   'psum_d', 'psum_r', 'model2', 'inter_weight', 'scale', 'prob')}
    out.update(st)
    return out


def rel_err(a, b):
    a = np.asarray(a, dtype='f8')
    b = np.asarray(b, dtype='f8')
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def densify(p):
    """The same problem with its frames as dense integer frames (ftype DENSE_INT) and the
    dense-frame precompute: fcounts over every pixel (iterate.c:38-46)."""
    from oracle import brute
    K = brute.dense_counts(p['frames'], len(p['det']['corr'])).astype('i4')
    q = dict(p)
    q['frames'] = {'int_frames': K}
    fp = dict(p['fparams'])
    fp['fcounts'] = K.sum(1).astype('i4')
    fp['mean_count'] = fp['fcounts'].astype('f8').sum() / K.shape[0]
    q['fparams'] = fp
    return q


def engine_for_dense(p):
    from dragonfly_amd._lib import Engine
    eng = Engine(0)
    eng.set_detector(p['det']['qvals'], p['det']['corr'], p['det']['mask'])
    eng.set_quats(p['quats'])
    eng.set_model(p['model1'])
    eng.set_frames_dense(p['frames']['int_frames'])
    fp = p['fparams']
    eng.set_frame_params(fp['scale'], fp['beta'], fp['blacklist'], fp['fcounts'], fp['sum_fact'],
                         fp['mean_count'], fp['num_blacklist'])
    return eng


def write_config(p, folder, **keys):
    """Write problem p as the files a reference run reads -- detector (.npz with the HDF5
    dataset names), photons (.emc), start model (raw float64) -- and a config.ini whose
    [emc] section holds these plus `keys`. Returns the config path."""
    import os
    from dragonfly_amd import CDetector, EMCWriter
    os.makedirs(folder, exist_ok=True)
    det = p['det']
    CDetector.from_arrays(det['qvals'], det['corr'], det['mask']).write(os.path.join(folder, 'det.npz'))
    f = p['frames']
    with EMCWriter(os.path.join(folder, 'photons.emc'), len(det['corr'])) as w:
        for d in range(len(f['ones'])):
            o, m = f['ones_accum'][d], f['multi_accum'][d]
            w.write_sparse_frame(f['place_ones'][o:o + f['ones'][d]], f['place_multi'][m:m + f['multi'][d]],
                                 f['count_multi'][m:m + f['multi'][d]])
    p['model1'].astype('f8').tofile(os.path.join(folder, 'start.bin'))
    emc = dict(in_detector_file='det.npz', in_photons_file='photons.emc', start_model_file='start.bin',
               num_div=str(p.get('num_div', 2)), output_folder='out/', log_file='EMC.log')
    emc.update({k: str(v) for k, v in keys.items()})
    path = os.path.join(folder, 'config.ini')
    with open(path, 'w') as fp:
        fp.write('[emc]\n' + ''.join('%s = %s\n' % kv for kv in emc.items()))
    return path


def from_synth(prob, quats=None, start_seed=4, beta='auto'):
    """A problem dict (as build() returns) for a configuration made by synth.make_problem:
    the configuration's detector, frames and orientations, a seeded start model and the
    per-frame precompute; beta 'auto' is calc_beta's per-frame rule with scale factors
    (iterate.c:66-83) or one value for every frame."""
    import math
    from dragonfly_amd import _lib
    cfg, det, f = prob['cfg'], prob['det'], prob['frames']
    mask = det['mask']
    fc = frame_counts(f['ones'], f['multi'], f['place_ones'], f['place_multi'], f['count_multi'], mask)
    sf = sum_log_factorials(f['multi'], f['place_multi'], f['count_multi'], mask)
    D = len(f['ones'])
    mean_count = fc.astype('f8').sum() / D
    if beta == 'auto':
        beta_v = np.array([math.exp(-6.5 * math.pow(c * 1e-5, 0.15)) for c in fc.tolist()])
    else:
        beta_v = np.full(D, float(beta))
    fparams = dict(scale=np.ones(D), beta=beta_v, blacklist=np.zeros(D, 'u1'), fcounts=fc, sum_fact=sf,
                   mean_count=mean_count, num_blacklist=0)
    return dict(det=det, frames={k: f[k] for k in SPARSE_KEYS},
                quats=_lib.quat_gen(cfg['num_div']) if quats is None else quats,
                model1=synth.start_model(cfg['size'], mean_count, mask, seed=start_seed), fparams=fparams,
                intens=prob['intens'], true_scale=f['true_scale'])


def oracle_threads():
    """Host threads for oracle runs that must stay fast (full orientation sets): the result
    equals the one-thread rule whenever every frame keeps every orientation it evaluates."""
    import os
    return max(1, min(16, os.cpu_count() or 1))
