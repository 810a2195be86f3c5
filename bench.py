"""Benchmark: plain FP64 EMC iterations on one MI355X.

Workload (default): 251^3 model from 400 Gaussian atoms in a radius-30 sphere; 260x257
detector as two panels with a 10-px gap; 100,000 sparse Poisson frames, mean 200 photons
with per-frame scale U(0.5, 1.5); num_div 10 (50,100 orientations); scale optimisation on.
A step is one full EMC iteration (maximize + model update + scale normalisation) through
EMCRecon.run_iteration with inputs resident in HBM; output files are not written.

Prints one JSON line: frame x orientation pairs/s, the roofline of the dominant kernel
from HIP events on the engine's stream, and a CPU baseline (the oracle's OpenMP restatement
of the reference engine) timed on a bounded orientation sample.
"""
import argparse
import concurrent.futures
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'frame×orientation pairs/s per EMC iter; speedup vs plain FP64 GPU at ≤1e-3 rel err'
HBM_PEAK_GBS = 8000.


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_recon(prob, cfg, device=0, dense=None, precision=0, start=None, background=None):
    import dragonfly_amd as df
    from dragonfly_amd.iterate import Iterate
    det = prob['det']
    f = prob['frames']
    cdet = df.CDetector.from_arrays(det['qvals'], det['corr'], det['mask'], det['detd'], det['ewald_rad'])
    if background is not None:
        cdet.background = background
    if dense is not None:
        ds = df.CDataset.from_dense(cdet, dense)
    else:
        ds = df.CDataset.from_arrays(cdet, f['ones'], f['multi'], f['place_ones'], f['place_multi'],
                                     f['count_multi'])
    itr = Iterate()
    itr.params.need_scaling = cfg['need_scaling']
    itr.params.alpha = cfg.get('alpha', 0.)
    itr.params.beta_factor = 1.
    itr.params.verbosity = 0
    itr.dets = [cdet]
    itr.set_data(ds)
    itr.set_quat(df.Quaternion(cfg['num_div']))
    from dragonfly_amd import synth
    model = df.Model(cfg['size'])
    model.allocate('')
    model.model1 = synth.start_model(cfg['size'], itr.mean_count[0], det['mask']) if start is None \
        else start.copy()
    itr.set_model(model)
    itr.parse_scale('')
    itr.calc_beta(None if cfg['beta'] == 'auto' else float(cfg['beta']))
    recon = df.EMCRecon(device=device, precision=precision)
    recon.write_output = False
    recon.set_iterate(itr)
    return recon, itr


def cpu_model_name():
    try:
        with open('/proc/cpuinfo') as fh:
            for line in fh:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_baseline(prob, itr, threads, background=None, target_rot=(12800, 25600, 51200), frames=512):
    """Iteration 1 of the oracle engine (OpenMP restatement of the reference engine) on
    `frames` frames and three orientation subsets (every k-th orientation, about target_rot
    orientations), extrapolated to the workload with the reference engine's cost structure:

        T(N, n) = alpha N + n (beta N + gamma N^2) + c

    for N orientations and n frames. alpha N is the per-orientation work (slice generation,
    log, merge), beta N the per-frame photon sums and gamma N^2 the lookup of every (frame,
    orientation) by a linear search of the frame's per-thread keep list of ~N / threads
    entries (max_update.c:86-93); c is the rest (allocation, combining the thread volumes).
    Each run splits its own wall time into these parts with the oracle's per-thread timers
    (thread-seconds in per-orientation work over thread-seconds in the orientation loops), so
    no difference of two noisy run times is taken. Returns pairs/s at the workload's size,
    the sample text and the fit."""
    from oracle import core
    from dragonfly_amd import synth
    f = prob['frames']
    R = itr.quat.num_rot
    D = itr.tot_num_data - int((itr.blacklist != 0).sum())
    n = min(frames, itr.tot_num_data)
    sub = synth.take_frames(f, n)
    pts = []
    for target in target_rot:
        k = max(1, int(round(R / target)))
        quats = itr.quat.quats[::k].copy()
        fp = dict(scale=np.ones(n), beta=itr.beta_start[:n].copy(), blacklist=np.zeros(n, 'u1'),
                  fcounts=itr.fcounts[:n], sum_fact=itr.sum_fact[:n], mean_count=itr.mean_count[0])
        bg = None if background is None else dict(background=background['background'], bgscale=np.ones(n))
        out = core.maximize(prob['det'], quats, itr.model.model1, sub, fp,
                            dict(iteration=1, need_scaling=itr.params.need_scaling, update_scale=1),
                            num_threads=threads, background=bg)
        sec = out['seconds']
        loops = sec[0] + sec[1] + sec[3]
        share = sec[5] / sec[6] if sec[6] > 0 else 0.
        pts.append(dict(orientations=len(quats), frames=n, seconds=round(sec[4], 3),
                        per_orientation_s=round(share * loops, 3),
                        per_frame_s=round((1. - share) * loops + sec[2], 3),
                        other_s=round(sec[4] - loops - sec[2], 3)))
        log('cpu baseline: %d frames x %d orientations, %d threads: %.2f s (%s)' % (n, len(quats), threads, sec[4], pts[-1]))
    # per-orientation work: alpha = mean of its share per orientation; per-frame work per
    # orientation b / N = beta + gamma N, least squares over the points with beta, gamma >= 0
    Ns = np.array([p_['orientations'] for p_ in pts], 'f8')
    alpha = float(np.mean([p_['per_orientation_s'] for p_ in pts] / Ns))
    y = np.array([p_['per_frame_s'] / n for p_ in pts]) / Ns
    A = np.stack([np.ones_like(Ns), Ns], axis=1)
    beta, gamma = np.linalg.lstsq(A, y, rcond=None)[0]
    if gamma < 0.:
        beta, gamma = float(y.mean()), 0.
    elif beta < 0.:
        beta, gamma = 0., float((Ns * y).sum() / (Ns * Ns).sum())
    beta, gamma = float(beta), float(gamma)
    c = float(np.mean([p_['other_s'] for p_ in pts]))
    t_full = alpha * R + D * (beta * R + gamma * R * R) + c
    sample = ('iteration 1 of the oracle engine (OpenMP restatement of the reference, %d threads) on %d frames at '
              '%s of the %d orientations (%s s); fit T = alpha N + n (beta N + gamma N^2) + c from each run\'s '
              'per-orientation / per-frame split (least squares, beta and gamma >= 0): alpha = %.3g s, beta = %.3g s, '
              'gamma = %.3g s, c = %.2f s; extrapolated to %d frames x %d orientations: %.0f s'
              % (threads, n, ', '.join('%d' % N for N in Ns), R, ', '.join('%.1f' % p_['seconds'] for p_ in pts),
                 alpha, beta, gamma, c, D, R, t_full))
    fit = dict(points=pts, alpha_s=float('%.4g' % alpha), beta_s=float('%.4g' % beta),
               gamma_s=float('%.4g' % gamma), c_s=round(c, 3), extrapolated_s=round(t_full, 1))
    return D * R / t_full, sample, fit


def photon_stats(prob, itr):
    """Mean photon pixels per frame on likelihood pixels (mask < 1) and on tomogram pixels
    (mask < 2), and the multi-photon part of each."""
    f = prob['frames']
    mask = prob['det']['mask']
    D = itr.tot_num_data
    out = {}
    for key, lim in (('lik', 1), ('tomo', 2)):
        go = mask[f['place_ones']] < lim
        gm = mask[f['place_multi']] < lim
        out[key] = ((go.sum() + gm.sum()) / D, gm.sum() / D)
    return out


def kernel_source_hash():
    """sha256 (16 hex digits) of the HIP sources the library is built from: a measured
    traffic figure is used only while the kernels it was measured on are unchanged."""
    import hashlib
    h = hashlib.sha256()
    csrc = os.path.join(ROOT, 'dragonfly_amd', 'csrc')
    for name in sorted(os.listdir(csrc)):
        if name.endswith(('.hip', '.h', '.cpp')):
            with open(os.path.join(csrc, name), 'rb') as fh:
                h.update(name.encode() + b'\0' + fh.read())
    return h.hexdigest()[:16]


def load_traffic(workload, kernel):
    """(HBM-side bytes per launch of `kernel` on this workload, profile note) from the
    committed rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE, calibrated on a plain copy),
    or (None, reason) when there is none or it was measured on other kernel sources."""
    path = os.path.join(ROOT, 'profiles', 'pmc_traffic.json')
    try:
        with open(path) as fh:
            tab = json.load(fh)
    except (OSError, ValueError):
        return None, 'no profiles/pmc_traffic.json'
    ent = tab.get(workload, {}).get(kernel)
    if ent is None:
        return None, 'kernel not in profiles/pmc_traffic.json for this workload'
    src = tab.get(workload, {}).get('_src_sha16')
    if src != kernel_source_hash():
        return None, 'profiles/pmc_traffic.json measured on other kernel sources (%s)' % src
    return float(ent['bytes_per_launch']), tab[workload].get('_profile', 'profiles/pmc_traffic.json')


def roofline(prob, itr, stage_tot, steps, nchunks, pairs, workload, value_bytes=8):
    """HBM roofline of the dominant photon kernel (likelihood or tomogram gather, whichever
    takes longer), timed with HIP events on the engine's stream.

    `achieved` counts compulsory bytes: every operand of one launch moved once -- for the
    likelihood the log W chunk (8 B per orientation x pixel), the frames' photon lists
    (int32 pixel + f64 count) and R written (8 B per pair); for the tomogram the posterior
    chunk (8 B per pair), the pixels' photon lists (int32 frame + f64 count) and the view
    written (8 B per orientation x pixel). `gather_bytes` is what the kernel requests through
    the caches instead (one 8-byte log W or P value per photon and orientation, 12 n + 4 m
    bytes per pair); `traffic` is the measured HBM-side bytes per launch of the same kernel on
    the same workload (rocprofv3 FETCH_SIZE + WRITE_SIZE), when a profile of the current
    kernel sources exists. value_bytes 4: the gathered operand (log W, resp. P) is FP32."""
    ps = photon_stats(prob, itr)
    D = itr.tot_num_data - int((itr.blacklist != 0).sum())
    npix = len(prob['det']['mask'])
    C = itr.quat.num_rot / nchunks
    kern = {'likelihood': ('likelihood_kernel', ps['lik']),
            'tomogram': ('tomogram_part_kernel', ps['tomo'])}
    stage = max(kern, key=lambda k: stage_tot.get(k, 0.))
    name, (n, m) = kern[stage]
    launch_ms = stage_tot[stage] / steps / nchunks
    nnz = n * itr.tot_num_data
    vb = float(value_bytes)
    if stage == 'likelihood':
        compulsory = vb * C * npix + 12. * nnz + 8. * D * C
    else:
        compulsory = vb * D * C + 12. * nnz + 8. * C * npix
    gather = pairs / nchunks * ((vb + 4.) * n + 4. * m)
    achieved = compulsory / (launch_ms * 1e-3) / 1e9
    traffic, note = load_traffic(workload, name) if workload else (None, 'not the default workload size')
    out = dict(bound='hbm', achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit='GB/s',
               frac=round(achieved / HBM_PEAK_GBS, 4),
               traffic=None if traffic is None else round(traffic),
               kernel=name, launch_ms=round(launch_ms, 3), compulsory_bytes=round(compulsory),
               gather_bytes=round(gather), gather_GBs=round(gather / (launch_ms * 1e-3) / 1e9, 1),
               traffic_source=note)
    if traffic is not None:
        out['traffic_over_compulsory'] = round(traffic / compulsory, 2)
        out['hbm_frac_measured'] = round(traffic / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    return out


def roofline_learned(res, slots, num_pix, size, num_photons, num_frames, workload):
    """HBM roofline of the judged proposal iteration's longest kernel among the expand (first
    pass), the full frames' tomogram, the pair tomogram and the merge, each one launch per
    iteration (one orientation chunk), timed with HIP events on the engine's stream, on
    compulsory bytes (every operand moved once; FP32 gathers):
    expand: the FP32 model copy read (halo bricks, 512 B per 4^3 cells) and log W written
      (4 B per orientation x pixel);
    tomogram (full frames): their FP32 posterior read (4 B per frame x orientation), their
      photon lists (int32 frame + f64 count) and the partial views written (4 B per
      orientation x pixel);
    reduce (pair tomogram): the partial views read, the photon lists of every frame, the
      nonzero masks (8 B per orientation x 64-pixel patch) and the nonzero view entries
      written (4 B each; the engine counts them);
    merge: the nonzero masks and nonzero view entries read and the merged volume read and
      written (16 B per voxel)."""
    st, st1 = res['stage_ms'], res['pass1_stage_ms']
    nnz = float(res.get('view_nonzero') or 0.)
    npatch = (num_pix + 63) // 64
    nb = (size - 2) // 4 + 1
    full = res['fallback_rate'] * num_frames
    photons_full = num_photons * res['fallback_rate']
    dense = 4. * slots * num_pix
    masks = 8. * slots * npatch
    comp = {'expand': 512. * nb ** 3 + dense,
            'tomogram': 4. * full * slots + 12. * photons_full + dense,
            'reduce': (dense if full > 0 else 0.) + 12. * num_photons + masks + 4. * nnz,
            'merge': masks + 4. * nnz + 16. * float(size) ** 3}
    ms = {'expand': st1.get('expand', 0.), 'tomogram': st.get('tomogram', 0.), 'reduce': st.get('reduce', 0.),
          'merge': st.get('merge', 0.)}
    kernels = {'expand': 'expand_kernel', 'tomogram': 'tomogram_part_w128_kernel', 'reduce': 'pair_tomogram_kernel',
               'merge': 'merge_brick_model_kernel'}
    stage = max(ms, key=ms.get)
    per = {k: dict(kernel=kernels[k], launch_ms=round(ms[k], 3), compulsory_bytes=round(comp[k]),
                   frac=round(comp[k] / (ms[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if ms[k] > 0 else None)
           for k in ms}
    achieved = comp[stage] / (ms[stage] * 1e-3) / 1e9
    traffic, note = load_traffic(workload, kernels[stage]) if workload else (None, 'not the default workload size')
    out = dict(bound='hbm', achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit='GB/s',
               frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None if traffic is None else round(traffic),
               kernel=kernels[stage], stage=stage, launch_ms=round(ms[stage], 3), compulsory_bytes=round(comp[stage]),
               view_nonzero=nnz, traffic_source=note, stages=per)
    if traffic is not None:
        out['traffic_over_compulsory'] = round(traffic / comp[stage], 2)
        out['hbm_frac_measured'] = round(traffic / (ms[stage] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    return out


def roofline_fixed_stages(stage_ms, num_chunks, slots, num_pix, size, value_bytes):
    """HBM roofline of the longer of the two per-orientation stages of an iteration, expand
    and merge, per launch (one launch per orientation chunk), on compulsory bytes: expand
    reads the model once (8 B per voxel) and writes log W (value_bytes per orientation x
    pixel); the merge reads the view (8 B per orientation x pixel) and reads and writes the
    merged volume once (16 B per voxel)."""
    vol = float(size) ** 3
    C = slots / num_chunks
    comp = {'expand': 8. * vol + value_bytes * C * num_pix, 'merge': 8. * C * num_pix + 16. * vol}
    stage = max(comp, key=lambda k: stage_ms.get(k, 0.))
    launch_ms = stage_ms[stage] / num_chunks
    achieved = comp[stage] / (launch_ms * 1e-3) / 1e9
    return dict(bound='hbm', achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit='GB/s',
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                traffic_source='no counter profile of this workload', stage=stage,
                kernel='merge_brick_model_kernel' if stage == 'merge' else 'expand_kernel',
                launch_ms=round(launch_ms, 3), compulsory_bytes=round(comp[stage]))


def run_steps(recon, itr, warmup, steps, keep_models=False):
    """warmup + steps iterations; returns elapsed seconds of the timed steps, the summed
    stage times, the chunk count, and per iteration (avg log-likelihood, model1 copy)."""
    import torch
    trace = []
    for it in range(1, warmup + 1):
        itr.params.iteration = it
        recon.run_iteration()
        trace.append((itr.likelihood, itr.model.model1.copy() if keep_models else None))
        log('warmup iteration %d: %s' % (it, json.dumps({k: round(v, 1) for k, v in recon.last_stats['stage_ms'].items()})))
    torch.cuda.synchronize()
    stage_tot, nchunks, elapsed = {}, 1, 0.
    for k in range(steps):
        itr.params.iteration = warmup + 1 + k
        t0 = time.perf_counter()
        recon.run_iteration()
        torch.cuda.synchronize()
        elapsed += time.perf_counter() - t0
        st = recon.last_stats['stage_ms']
        nchunks = recon.last_stats['num_chunks']
        for name, v in st.items():
            stage_tot[name] = stage_tot.get(name, 0.) + v
        log('iteration %d: %s' % (itr.params.iteration, json.dumps({k2: round(v, 1) for k2, v in st.items()})))
        trace.append((itr.likelihood, itr.model.model1.copy() if keep_models else None))
    return elapsed, stage_tot, nchunks, trace


def bench_dense(args):
    """A reduced-precision path timed, then the plain FP64 path on the same frames from the
    same start model, with the model and log-likelihood relative errors of every iteration:
    dense integer frames with FP16 products on matrix cores (precision 1 or 2), or any frames
    with FP32 log W and posterior gathers (precision 3)."""
    import torch
    from dragonfly_amd import synth
    log('generating %s workload ...' % args.config)
    prob = synth.make_problem(args.config, device='cuda', num_data=args.num_data,
                              num_div=args.num_div, batch=512)
    cfg = prob['cfg']
    mask = prob['det']['mask']
    K = synth.dense_frames(prob['frames'], len(mask)) if cfg.get('dense') else None
    recon, itr = build_recon(prob, cfg, dense=K, precision=args.precision)
    start = itr.model.model1.copy()
    D, R = itr.tot_num_data, itr.quat.num_rot
    pairs = (D - int((itr.blacklist != 0).sum())) * R
    el_mp, st_mp, nch_mp, tr_mp = run_steps(recon, itr, args.warmup, args.steps, keep_models=True)
    recon.free()
    del recon
    torch.cuda.empty_cache()
    out_fp = None
    if not args.no_fp64:
        recon, itr2 = build_recon(prob, cfg, dense=K, precision=0, start=start)
        el_fp, st_fp, nch_fp, tr_fp = run_steps(recon, itr2, args.warmup, args.steps, keep_models=True)
        recon.free()
        model_err = [float(np.abs(a[1] - b[1]).max() / np.abs(b[1]).max()) for a, b in zip(tr_mp, tr_fp)]
        like_err = [abs(a[0] - b[0]) / abs(b[0]) for a, b in zip(tr_mp, tr_fp)]
        out_fp = dict(fp64_value=round(pairs * args.steps / el_fp, 1),
                      fp64_ms_per_step=round(el_fp / args.steps * 1e3, 2),
                      speedup_vs_fp64=round(el_fp / el_mp, 2),
                      model_rel_err=[float('%.3g' % x) for x in model_err],
                      loglik_rel_err=[float('%.3g' % x) for x in like_err],
                      fp64_stage_ms_mean={k: round(v / args.steps, 2) for k, v in st_fp.items()})
    value = pairs * args.steps / el_mp
    terms = 2 if args.precision == 2 else 1
    P = len(mask)
    launch_ms = st_mp['likelihood'] / args.steps / nch_mp
    if args.precision in (1, 2):
        flops = 2. * D * (R / nch_mp) * P * terms
        achieved = flops / (launch_ms * 1e-3) / 1e12
        # measured HBM-side bytes of the one-term product at the default sizes
        default_size = args.num_data is None and args.num_div is None and terms == 1
        traffic, note = load_traffic(args.config, 'mp_gemm_kernel_likelihood') if default_size else \
            (None, 'not the default workload size')
        roof = dict(bound='mfma', achieved=round(achieved, 1), peak=2500., unit='TFLOP/s',
                    frac=round(achieved / 2500., 4), traffic=None if traffic is None else round(traffic),
                    kernel='mp_gemm_kernel (likelihood)', launch_ms=round(launch_ms, 3),
                    algorithmic_flops=flops, traffic_source=note)
        if traffic is not None:
            roof['hbm_frac_measured'] = round(traffic / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        path = 'mixed precision, %d-term FP16 log W' % terms
        dtype = 'f16 products / f32 accumulate / f64 rest'
    else:
        roof = roofline(prob, itr, st_mp, args.steps, nch_mp, pairs, None, value_bytes=4)
        path = 'FP32 log W and posterior gathers, FP64 sums'
        dtype = 'f32 gathers / f64 rest'
    
    out = {
        'metric': METRIC, 'value': round(value, 1), 'unit': 'frame×orientation pairs/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(el_mp / args.steps * 1e3, 2), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': dtype,
        'data': 'synthetic',
        'config': {'workload': '%s: %d %s frames, %d orientations (num_div %d), %d^3 model, %d pixels'
                               % (args.config, D, 'dense' if K is not None else 'sparse', R, cfg['num_div'],
                                  cfg['size'], P),
                   'path': path, 'frames': D, 'num_rot': R, 'num_pix': P},
        'stage_ms_mean': {k: round(v / args.steps, 2) for k, v in st_mp.items()},
        'roofline': roof, 'cpu_baseline': None,
    }
    if out_fp:
        out.update(out_fp)
    print(json.dumps(out), flush=True)


def learned_case(prob, cfg, beta_factor, args, timed_steps, warmup):
    """One beta setting of the learned-proposal workload (fine num_div 16, coarse num_div 4).

    (1) Plain FP64 iterations 1..k over every fine orientation from the true intensity at
    photon scale (`--converge-iters` k, so that iteration k + 1 is a late iteration of a
    converging run); iteration k's posterior summed over the Voronoi cells of the coarse
    sampling gives the training targets (the method's own output of an earlier iteration), its
    top cells the no-network proposals, and its updated model and scales the input of iteration
    k + 1. (2) Plain FP64 iteration k + 1, timed (`fp64_steps` times from the same input): the
    denominator and the reference model. (3) The network (polar photon histogram -> coarse-cell
    logits) is trained on 90% of the frames. (4) Iteration k + 1 from the same input with
    proposals, timed: a full search over the 3240 coarse orientations (expectation step only),
    each frame's proposal (method 'learned': the network's top-k cells joined with iteration k's
    top-k cells and the coarse search's top cells), pass 1 over the proposals, the likelihood
    judge (mass outside each proposal estimated from the coarse search against the fine maximum
    inside it; frames above 1 - judge_threshold fall back to the full fine search) and the
    M-step (resumed from pass 1 when no frame falls back, else a second pass), with FP32
    gathers, and the model update. (5) The same with the other proposal sources beside it:
    'previous_posterior' (iteration k's cells, no network), 'network_only' (no posterior of an
    earlier fine iteration) and 'coarse_only' (the coarse search's own top-k cells: neither)."""
    import torch
    from dragonfly_amd import _lib, proposal
    det, mask = prob['det'], prob['det']['mask']
    true_model = np.ascontiguousarray(prob['intens'] * prob['frames']['rescale'])
    k_top, k_coarse = cfg['top_k'], args.coarse_k
    tau = cfg['mass_threshold'] if args.judge_threshold is None else args.judge_threshold

    def recon_for(num_div, precision=0):
        recon, itr = build_recon(prob, dict(cfg, num_div=num_div), start=true_model, precision=precision)
        itr.params.beta_factor = beta_factor
        itr.params.iteration = 1
        return recon, itr

    # (1) iterations 1..k, plain: targets, previous proposals and the input of iteration k + 1
    recon, itr = recon_for(cfg['num_div'])
    D, R = itr.tot_num_data, itr.quat.num_rot
    qcoarse = _lib.quat_gen(cfg['coarse_div'])
    num_coarse = len(qcoarse)
    qmap = _lib.voronoi_subset(qcoarse, itr.quat.quats)
    kk = max(1, args.converge_iters)
    for it in range(1, kk + 1):
        itr.params.iteration = it
        t0 = time.perf_counter()
        recon.run_iteration()
        log('beta x %g: iteration %d (plain FP64): %.3f s, rms change %.3e' % (beta_factor, it, time.perf_counter() - t0,
                                                                                itr.rms_change))
    post = recon.engine.cell_posterior(num_coarse, qmap)
    prev_top, _ = recon.engine.cell_topk(num_coarse, qmap, k_top)
    model_k1 = itr.model.model1.copy()
    scale_k1 = np.array(itr.scale, dtype='f8')

    def start_k1(rec, it):
        it.model.model1 = model_k1
        it.scale[:] = scale_k1
        rec.engine.set_scale(scale_k1)
        it.params.iteration = kk + 1

    # (2) iteration k + 1, plain, timed
    plain_times = []
    for _ in range(args.fp64_steps):
        start_k1(recon, itr)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        recon.run_iteration()
        torch.cuda.synchronize()
        plain_times.append(time.perf_counter() - t0)
    plain_s = float(np.median(plain_times))
    log('beta x %g: iteration k + 1, plain FP64: %s s' % (beta_factor, ['%.3f' % t for t in plain_times]))
    st_plain = recon.last_stats['stage_ms']
    model_plain = itr.model.model1.copy()
    rmax_plain = recon.rmax.copy()
    beta_k1 = itr.beta.copy()
    recon.free()
    # (3) training
    net = proposal.OrientationProposal(num_coarse, nbins=4096, width=1024, depth=3, seed=0)
    off, pix, cnt = proposal.photon_lists(prob['frames'], mask)
    X = net.features(off, pix, cnt, proposal.polar_bins(det['qvals'], mask))
    Y = torch.tensor(post, dtype=torch.float32, device='cuda')
    del post
    ntr = int(0.9 * D)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = net.fit(X, Y, ntr, epochs=args.epochs, batch=1024, lr=1e-3, wd=args.wd)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log('beta x %g: trained %d epochs in %.1f s, loss %.4f' % (beta_factor, args.epochs, train_s, hist[-1]))
    del Y

    # (4), (5) iteration k + 1 with judged proposals
    judge_itr = itr
    judge_itr.model.model1 = model_k1
    judge = proposal.CoarseJudge(judge_itr, cfg['coarse_div'], precision=args.judge_precision)
    recon, itr = recon_for(cfg['num_div'], precision=args.precision)
    if args.fused:
        recon.engine.set_option('fused', 1)
    if args.sig_full_rows is not None:
        recon.engine.set_option('sig_full_rows', args.sig_full_rows)
    if args.unit_items is not None:
        recon.engine.set_option('unit_items', args.unit_items)
    start_k1(recon, itr)

    def sources(source):
        """The proposal arrays of a method (the coarse search's top cells are added to each)."""
        if source == 'learned':
            return [net.propose(X, D, k_top), prev_top]
        if source == 'network_only':
            return [net.propose(X, D, k_top)]
        return [prev_top]  # previous_posterior (coarse_only takes the coarse search's own top cells)

    pool = concurrent.futures.ThreadPoolExecutor(max_workers=1)

    def one_step(source):
        """Judge, proposals, both passes and the model update; returns (seconds, info). The
        coarse search runs on its own engine and stream from a second host thread (taking the
        fine engine's model on the device) while this one computes the proposals that do not
        need it (network, previous posterior)."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        coarse = pool.submit(judge.estep, recon.engine, scale_k1, beta_k1, kk + 1, cfg['need_scaling'])
        tops = sources(source) if source != 'coarse_only' else []
        t1 = time.perf_counter()
        coarse.result()
        t2 = time.perf_counter()
        if source == 'coarse_only':
            tops = [judge.top_cells(k_top)]
        elif k_coarse > 0:
            tops.append(judge.top_cells(k_coarse))
        t3 = time.perf_counter()
        info = proposal.judged_iteration(recon, itr, judge, tops, cfg['coarse_div'], tau, widen=args.widen)
        torch.cuda.synchronize()
        info['phase_ms'] = dict(proposals=round((t1 - t0) * 1e3, 1), coarse_search=round((t2 - t0) * 1e3, 1),
                                coarse_top=round((t3 - t2) * 1e3, 1), **info['phase_ms'])
        return time.perf_counter() - t0, info

    results = {}
    others = [m for m in ('previous_posterior', 'network_only', 'coarse_only') if m in args.methods.split(',')]
    methods = [('learned', timed_steps, warmup)] + [(m, args.baseline_steps, 1) for m in others]
    for source, nsteps, nwarm in methods:
        for _ in range(max(nwarm, 1)):
            start_k1(recon, itr)
            one_step(source)
        times = []
        for _ in range(nsteps):
            start_k1(recon, itr)
            dt, info = one_step(source)
            times.append(dt)
        ml = itr.model.model1.copy()
        step_s = float(np.median(times))
        acc = info['accept']
        held = np.arange(D) >= ntr
        res = dict(step_ms=round(step_s * 1e3, 1), steps_ms=[round(t * 1e3, 1) for t in times],
                   num_chunks=recon.last_stats['num_chunks'], passes=info['passes'],
                   cells_per_frame=round(info['cells_per_frame'], 2), speedup_vs_fp64=round(plain_s / step_s, 2),
                   fallback_rate=round(float(1. - acc.mean() - info['widened'].mean()), 4),
                   widened_rate=round(float(info['widened'].mean()), 4),
                   heldout_fallback_rate=round(float(1. - (acc | info['widened'])[held].mean()), 4),
                   judge_outside_accepted_mean=float('%.3g' % info['outside'][acc].mean()) if acc.any() else None,
                   model_rel_err=float('%.3g' % (np.abs(ml - model_plain).max() / np.abs(model_plain).max())),
                   model_rms_rel_err=float('%.3g' % np.sqrt(np.mean((ml - model_plain) ** 2) / np.mean(model_plain ** 2))),
                   rmax_agree=round(float(np.mean(recon.rmax == rmax_plain)), 5),
                   view_nonzero=recon.last_stats.get('view_nonzero', 0.),
                   stage_ms={k: round(v, 1) for k, v in recon.last_stats['stage_ms'].items()},
                   pass1_stage_ms={k: round(v, 1) for k, v in info['estep_stage_ms'].items()},
                   phase_ms=info['phase_ms'])
        results[source] = res
        log('beta x %g, %s: %s' % (beta_factor, source, json.dumps(res)))
    pool.shutdown()
    judge_ms = judge.stats['stage_ms'] if judge.stats else {}
    judge.close()
    recon.free()
    f = prob['frames']
    num_photons = int((mask[f['place_ones']] < 2).sum() + (mask[f['place_multi']] < 2).sum())
    return dict(D=D, R=R, num_coarse=num_coarse, plain_s=plain_s, plain_times=plain_times, st_plain=st_plain,
                num_photons=num_photons,
                train_s=train_s, epochs=args.epochs, loss=hist[-1], results=results,
                judge_stage_ms={k: round(v, 1) for k, v in judge_ms.items()}, itr=itr, k=kk, tau=tau)


def bench_learned(args):
    """Learned orientation proposal at fine num_div 16 against the plain FP64 full search,
    at beta_start x 4 (the reference schedule's beta after 40 iterations at sqrt(2) per 10,
    the late iterations in which proposals are meant to be used; the timed headline) and at
    the configuration's own beta_start (reported beside it)."""
    from dragonfly_amd import synth
    log('generating %s workload ...' % args.config)
    t0 = time.perf_counter()
    prob = synth.make_problem(args.config, device='cuda', num_data=args.num_data, num_div=args.num_div,
                              batch=1024)
    log('generated in %.1f s' % (time.perf_counter() - t0))
    cfg = dict(prob['cfg'])
    if args.top_k:
        cfg['top_k'] = args.top_k
    late = float(cfg.get('beta_factor', 1.)) if args.beta_factor is None else args.beta_factor
    default_size = args.num_data is None and args.num_div is None
    head = learned_case(prob, cfg, late, args, args.steps, args.warmup)
    other = None
    if not args.single_beta and late != 1.:
        other = learned_case(prob, cfg, 1., args, max(1, min(args.steps, 3)), 1)
    D, R = head['D'], head['R']
    net = head['results']['learned']
    step_s = net['step_ms'] * 1e-3
    cpu = None
    if not args.no_cpu:
        itr = head['itr']
        itr.model.model1 = np.ascontiguousarray(prob['intens'] * prob['frames']['rescale'])
        threads = min(16, os.cpu_count() or 1)
        v_cpu, sample, fit = cpu_baseline(prob, itr, threads)
        cpu = dict(value=round(v_cpu, 1), unit='frame×orientation pairs/s', cores=threads, kind='port',
                   sample=sample + ' (full search at fine num_div %d, the reference engine has no proposals)'
                   % cfg['num_div'], fit=fit, nproc=os.cpu_count(), cpu_model=cpu_model_name())

    def summary(case):
        return dict(beta_factor=case.get('beta_factor'), fp64_ms_per_step=round(case['plain_s'] * 1e3, 1),
                    fp64_steps_ms=[round(t * 1e3, 1) for t in case['plain_times']],
                    train_s=round(case['train_s'], 2), train_loss=round(case['loss'], 4),
                    judge_stage_ms=case['judge_stage_ms'], methods=case['results'])
    head['beta_factor'] = late
    out = {
        'metric': METRIC, 'value': round(D * R / step_s, 1), 'unit': 'frame×orientation pairs/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': net['step_ms'],
        'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': net['speedup_vs_fp64'],
        'vs_baseline_def': 'speedup over the plain FP64 GPU full search on the same iteration (the denominator '
                           'the metric names); model_rel_err is the max-abs error of the updated model relative '
                           'to its maximum',
        'model_rel_err': net['model_rel_err'],
        'dtype': ('f32 gathers, f64 sums' if args.precision == 3 else 'f64') + ' EMC / bf16 proposal network',
        'data': 'synthetic',
        'config': {'workload': '%s: %d frames (mean 200 photons, scale U(0.5, 1.5)), 251^3 model, 260x257 '
                               'detector, fine num_div %d (%d orientations), coarse num_div %d (%d); iteration '
                               '%d from %d plain iterations started at the true intensity; proposals = top-%d '
                               'cells of the network + top-%d of the previous iteration\'s posterior + top-%d of a '
                               'coarse full search; likelihood judge (fall back to the full search where the '
                               'estimated mass outside exceeds %.4g); '
                               'beta_start x %g (timed); beta_start x 1 beside it'
                               % (args.config, D, cfg['num_div'], R, cfg['coarse_div'], head['num_coarse'],
                                  head['k'] + 1, head['k'], cfg['top_k'], cfg['top_k'], args.coarse_k,
                                  1. - head['tau'], late),
                   'path': 'learned proposal (+ previous posterior) + likelihood judge + sparse refinement',
                   'precision': args.precision,
                   'frames': D, 'num_rot': R},
        'speedup_vs_fp64': net['speedup_vs_fp64'], 'fallback_rate': net['fallback_rate'],
        'late_beta': summary(head),
        'config_beta': summary(dict(other, beta_factor=1.)) if other else None,
        'roofline': (roofline_learned(net, R, len(prob['det']['mask']), cfg['size'], head['num_photons'], D,
                                      'cfg5' if default_size else None)
                     if args.precision == 3 and net['num_chunks'] == 1 else
                     roofline_fixed_stages(net['stage_ms'], net['num_chunks'], R, len(prob['det']['mask']),
                                           cfg['size'], 4 if args.precision == 3 else 8)),
        'cpu_baseline': cpu,
    }
    print(json.dumps(out), flush=True)


def bench_batched(args):
    """Uneven batching: 1M frames with lognormal photon counts (median 150, 20-20000) and
    per-frame orientation subsets of 50-5000 of the num_div 12 orientations, whole Voronoi
    cells of num_div 3 around a prior 3 degrees from the true orientation, from a model at the
    truth (photon scale), beta_start. Every frame's subset is evaluated sparsely: its (frame,
    64-orientation tile) pairs in one likelihood launch per orientation chunk, pairs walked
    tile-major, per-frame reductions over the pair rows. A step is one iteration; the metric
    counts the evaluated frame x orientation pairs (sum of the subset sizes).

    Beside it, on the first `--compare-frames` frames (a slice whose dense posterior fits the
    GPU): the same iteration over every pair with the excluded ones masked (the plain FP64
    search restricted to the subsets), for the speedup and the relative model difference."""
    import torch
    from dragonfly_amd import _lib, synth
    log('generating %s workload ...' % args.config)
    t0 = time.perf_counter()
    prob = synth.make_problem(args.config, device='cuda', num_data=args.num_data, num_div=args.num_div,
                              batch=2048)
    cfg = prob['cfg']
    log('generated in %.1f s' % (time.perf_counter() - t0))
    true_model = np.ascontiguousarray(prob['intens'] * prob['frames']['rescale'])
    qcoarse = _lib.quat_gen(cfg['prior_div'])
    nfine = _lib.num_rot(cfg['num_div'])
    rel_off, rel_quat = synth.prior_subsets(prob['frames']['true_quats'], qcoarse, nfine, cfg['subset_range'],
                                            cfg['prior_error_deg'])
    prec = 0 if args.precision is None else args.precision

    def setup(p, n):
        recon, itr = build_recon(p, cfg, start=true_model, precision=prec)
        itr.set_refine_lists(cfg['prior_div'], rel_off[:n + 1], rel_quat[:rel_off[n]])
        recon.engine.set_refine(itr.num_coarse, itr.quat_mapping, itr.rel_off, itr.rel_quat)
        cell_size = np.bincount(itr.quat_mapping, minlength=itr.num_coarse)
        pairs = int(cell_size[itr.rel_quat].sum())
        return recon, itr, pairs

    recon, itr, pairs = setup(prob, len(prob['frames']['ones']))
    D = itr.tot_num_data
    counts = itr.fcounts.astype('f8')
    log('%d frames, photons per frame median %.0f, mean %.0f, max %d; %.3g evaluated pairs (subset sizes '
        '%d-%d)' % (D, np.median(counts), counts.mean(), counts.max(), pairs,
                    int(np.diff(rel_off).min() * nfine / len(qcoarse)), int(np.diff(rel_off).max() * nfine / len(qcoarse))))
    for it in range(1, args.warmup + 1):
        itr.params.iteration = it
        recon.run_iteration()
        log('warmup iteration %d: %s' % (it, json.dumps({k: round(v, 1) for k, v in recon.last_stats['stage_ms'].items()})))
    times, stages = [], []
    for it in range(args.warmup + 1, args.warmup + args.steps + 1):
        itr.params.iteration = it
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        recon.run_iteration()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        stages.append(recon.last_stats['stage_ms'])
        log('iteration %d: %.3f s %s' % (it, times[-1], json.dumps({k: round(v, 1) for k, v in stages[-1].items()})))
    recon.free()
    step_s = float(np.mean(times))

    # sparse vs masked on a slice of frames
    n = min(D, args.compare_frames)
    sub = dict(prob, frames=synth.take_frames(prob['frames'], n))
    res = {}
    for mode in ('sparse', 'masked'):
        r2, i2, p2 = setup(sub, n)
        r2.engine.set_option('subset', 1 if mode == 'sparse' else 0)
        i2.params.iteration = 1
        r2.run_iteration()
        i2.model.model1 = true_model
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r2.run_iteration()
        torch.cuda.synchronize()
        res[mode] = dict(s=time.perf_counter() - t0, model=i2.model.model1.copy(), pairs=p2)
        log('%d-frame slice, %s: %.3f s' % (n, mode, res[mode]['s']))
        r2.free()
    mdiff = float(np.abs(res['sparse']['model'] - res['masked']['model']).max() / np.abs(res['masked']['model']).max())
    out = {
        'metric': METRIC, 'value': round(pairs / step_s, 1), 'unit': 'frame×orientation pairs/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': round(step_s * 1e3, 1),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32 gathers, f64 sums' if prec == 3 else 'f64', 'data': 'synthetic',
        'config': {'workload': '%s: %d frames, lognormal photons (median %.0f, 20-20000), num_div %d (%d '
                               'orientations), subsets of %d-%d orientations (num_div %d cells around a prior)'
                               % (args.config, D, np.median(counts), cfg['num_div'], nfine, cfg['subset_range'][0],
                                  cfg['subset_range'][1], cfg['prior_div']),
                   'path': 'sparse refinement subsets', 'frames': D, 'evaluated_pairs': pairs,
                   'mean_photons': round(float(counts.mean()), 1), 'max_photons': int(counts.max())},
        'stage_ms_mean': {k: round(float(np.mean([s[k] for s in stages])), 1) for k in stages[0]},
        'slice_compare': {'frames': n, 'pairs': res['sparse']['pairs'],
                          'sparse_ms': round(res['sparse']['s'] * 1e3, 1), 'masked_ms': round(res['masked']['s'] * 1e3, 1),
                          'speedup': round(res['masked']['s'] / res['sparse']['s'], 2), 'model_rel_diff': float('%.3g' % mdiff)},
        'roofline': None, 'cpu_baseline': None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--config', default='cfg5',
                    help='workload: cfg5 (default: learned proposals at num_div 16, the metric\'s headline), '
                         'cfg1, cfg2 (plain FP64), cfg3 (dense frames, mixed precision), cfg4 (uneven batching)')
    ap.add_argument('--num-data', type=int, default=None)
    ap.add_argument('--num-div', type=int, default=None)

    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--precision', type=int, default=None,
                    help='dense-frame configurations: 1 (default) or 2 FP16 terms of log W; 3 (any '
                    'configuration): FP32 log W and posterior gathers, compared with FP64 (the '
                    'default of the learned-proposal iteration; 0 there: FP64)')
    ap.add_argument('--epochs', type=int, default=80, help='learned-proposal workload: training epochs')
    ap.add_argument('--wd', type=float, default=0., help='learned-proposal workload: L2 weight decay')
    ap.add_argument('--top-k', type=int, default=0, help='learned-proposal workload: cells proposed per frame')
    ap.add_argument('--judge-threshold', type=float, default=None,
                    help='learned-proposal workload: estimated posterior mass a proposal must hold '
                         '(default: the configuration\'s mass_threshold, 0.99)')
    ap.add_argument('--coarse-k', type=int, default=4,
                    help='learned-proposal workload: cells of the coarse search joined to every proposal')
    ap.add_argument('--methods', default='previous_posterior,network_only,coarse_only',
                    help='learned-proposal workload: proposal sources timed beside the headline')
    ap.add_argument('--sig-full-rows', type=int, default=None,
                    help='learned-proposal workload: mean significant full-frame rows per tile up to which the '
                         'pair tomogram reads them directly (engine option sig_full_rows)')
    ap.add_argument('--widen', type=int, default=0,
                    help='learned-proposal workload: at most this many coarse cells the judge adds to a rejected '
                         'frame before it falls back to the full search (0: no widening)')
    ap.add_argument('--unit-items', type=int, default=None,
                    help='learned-proposal workload: brick-merge list entries per work unit (engine option '
                         'unit_items)')
    ap.add_argument('--fused', action='store_true',
                    help='learned-proposal workload: expand and pair likelihood in one persistent launch')
    ap.add_argument('--converge-iters', type=int, default=3,
                    help='learned-proposal workload: plain iterations before the timed one (k)')
    ap.add_argument('--judge-precision', type=int, default=3,
                    help='learned-proposal workload: arithmetic of the coarse judging search (0 FP64, 3 FP32 gathers)')
    ap.add_argument('--beta-factor', type=float, default=None,
                    help='learned-proposal workload: beta_start factor of the timed case (default: the config\'s)')
    ap.add_argument('--single-beta', action='store_true',
                    help='learned-proposal workload: skip the case at the configuration\'s own beta')
    ap.add_argument('--fp64-steps', type=int, default=2,
                    help='learned-proposal workload: timed plain FP64 iterations (the denominator)')
    ap.add_argument('--baseline-steps', type=int, default=2,
                    help='learned-proposal workload: timed steps of the no-network proposals')
    ap.add_argument('--compare-frames', type=int, default=50000,
                    help='uneven-batching workload: frames of the sparse vs masked comparison')
    ap.add_argument('--autotune', action='store_true',
                    help='plain workloads: pick kernel variants on the live workload before timing')
    ap.add_argument('--no-fp64', action='store_true',
                    help='dense-frame configurations: skip the FP64 comparison run')
    ap.add_argument('--background', type=float, default=0.,
                    help='detector background of this many photons per pixel (background-aware path, '
                         'scale optimisation on)')
    ap.add_argument('--copy-calibration-mib', type=int, default=0,
                    help='run one device copy of this many MiB first (counter calibration)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: there is no multi-GPU path (the work would shard '
              'by frames, each GPU taking every orientation). Run with --gpus 1.', file=sys.stderr)
        sys.exit(2)

    import torch
    from dragonfly_amd import synth
    if args.copy_calibration_mib:
        n = args.copy_calibration_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device='cuda')
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log('copy calibration: %d MiB read + %d MiB written' % (args.copy_calibration_mib, args.copy_calibration_mib))
        del a, b
    if synth.CONFIGS.get(args.config, {}).get('batched'):
        bench_batched(args)
        return
    if synth.CONFIGS.get(args.config, {}).get('learned'):
        # the learned-proposal iteration gathers in FP32 unless told otherwise (0: FP64)
        args.precision = 3 if args.precision is None else args.precision
        if args.precision not in (0, 3):
            sys.exit('the learned-proposal workload has sparse frames: --precision 0 or 3')
        bench_learned(args)
        return
    if args.precision is None:
        args.precision = 1 if synth.CONFIGS.get(args.config, {}).get('dense') else 0
    if synth.CONFIGS.get(args.config, {}).get('dense') or args.precision == 3:
        bench_dense(args)
        return
    t0 = time.perf_counter()
    log('generating %s workload ...' % args.config)
    prob = synth.make_problem(args.config, device='cuda', num_data=args.num_data,
                              num_div=args.num_div, batch=1024)
    cfg = prob['cfg']
    log('generated in %.1f s' % (time.perf_counter() - t0))
    bgd = None
    if args.background > 0:
        # a radial background profile of mean args.background photons per pixel, unit
        # background scales; the background-aware updates run with the scale factors on
        q = prob['det']['qvals']
        rad = np.sqrt((q ** 2).sum(1))
        prof = 1. + rad / rad.max()
        bgd = args.background * prof / prof.mean()
        cfg = dict(cfg, need_scaling=1)
    # host copy of the start model for the CPU baseline before the GPU changes it
    recon, itr = build_recon(prob, cfg, background=bgd)
    start_model = itr.model.model1.copy()
    tuned = None
    if args.autotune:
        from dragonfly_amd import autotune
        t0 = time.perf_counter()
        tuned, _ = autotune.tune(recon.engine, iteration=1, need_scaling=cfg['need_scaling'], log=log)
        log('autotune: %s in %.1f s' % (tuned, time.perf_counter() - t0))
    mask = prob['det']['mask']
    D = itr.tot_num_data
    nbar = photon_stats(prob, itr)['lik'][0]
    R = itr.quat.num_rot
    pairs = (D - int((itr.blacklist != 0).sum())) * R

    for it in range(1, args.warmup + 1):
        itr.params.iteration = it
        recon.run_iteration()
        log('warmup iteration %d: %s' % (it, json.dumps({k: round(v, 1) for k, v in recon.last_stats['stage_ms'].items()})))
    torch.cuda.synchronize()
    stage_tot, nchunks = {}, 1
    t_start = time.perf_counter()
    for k in range(args.steps):
        itr.params.iteration = args.warmup + 1 + k
        recon.run_iteration()
        st = recon.last_stats['stage_ms']
        nchunks = recon.last_stats['num_chunks']
        for name, v in st.items():
            stage_tot[name] = stage_tot.get(name, 0.) + v
        log('iteration %d: %s' % (itr.params.iteration, json.dumps({k2: round(v, 1) for k2, v in st.items()})))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_per_step = elapsed / args.steps * 1e3
    value = pairs * args.steps / elapsed

    default_size = args.num_data is None and args.num_div is None
    roof = None if bgd is not None else roofline(prob, itr, stage_tot, args.steps, nchunks, pairs,
                                                 args.config if default_size else None)
    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        recon.free()
        itr.model.model1 = start_model
        v_cpu, sample, fit = cpu_baseline(prob, itr, threads, None if bgd is None else dict(background=bgd))
        cpu = dict(value=round(v_cpu, 1), unit='frame×orientation pairs/s', cores=threads,
                   kind='port', sample=sample, fit=fit, nproc=os.cpu_count(), cpu_model=cpu_model_name())
    out = {
        'metric': METRIC, 'value': round(value, 1), 'unit': 'frame×orientation pairs/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': round(ms_per_step, 2),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': 'synthetic',
        'config': {'workload': '%s: %d frames, %d orientations (num_div %d), %d^3 model, %d pixels, '
                               'need_scaling %d' % (args.config, D, R, cfg['num_div'], cfg['size'],
                                                    len(mask), cfg['need_scaling']),
                   'path': 'plain FP64' if bgd is None else
                           'background-aware FP64 (background %.3g photons/pixel)' % args.background,
                   'frames': D, 'num_rot': R, 'num_pix': int(len(mask)),
                   'mean_photon_pixels': round(float(nbar), 2)},
        'stage_ms_mean': {k: round(v / args.steps, 2) for k, v in stage_tot.items()},
        'roofline': roof,
        'cpu_baseline': cpu,
    }
    if tuned is not None:
        out['autotune'] = tuned
    if bgd is not None:
        out['roofline_note'] = ('background-aware path: the tomogram is a per-(pixel, orientation) root '
                                'search whose gradient passes depend on the data; see the kernel trace')
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
