"""EMCRecon: one EMC iteration on the GPU (mirror of dragonfly/recon.pyx:19-445).

run_iteration keeps the reference's sequence (recon.pyx:72-110): update beta, maximize,
update the model (Friedel symmetrise, divide, rms, blend), normalise scale factors, write
outputs and the log line. maximize, update_model and normalize_scale run in the engine
(libemc_hip.so); the per-frame and per-orientation outputs are read back on demand.
"""
import argparse
import os
import time

import numpy as np

from . import _lib
from .iterate import Iterate

PROB_MIN = 1.e-6


class EMCRecon:
    """Main class for EMC reconstruction on one GPU.

    Example:
        >>> recon = EMCRecon()
        >>> itr = Iterate('config.ini')
        >>> recon.set_iterate(itr)
        >>> for itr.params.iteration in range(1, 11):
        ...     recon.run_iteration()
    """

    def __init__(self, num_threads=-1, device=None, precision=0):
        # num_threads is kept for interface compatibility; the GPU engine does not use it.
        # precision selects the arithmetic of the photon contractions for dense frames
        # (0: FP64 as the reference; 1, 2: FP16 products on matrix cores, see
        # Engine.set_precision).
        self.num_threads = num_threads
        self.device = int(os.environ.get('DRAGONFLY_DEVICE', '0')) if device is None else device
        self.precision = precision
        self.engine = None
        self.iter = None
        self.write_output = True
        self.last_stats = None

    def set_iterate(self, itr):
        if itr.data is None:
            print('Set data for iterate first')
            return
        if itr.model is None:
            print('Set model for iterate first')
            return
        if itr.quat is None:
            print('Set quat for iterate first')
            return
        if itr.params is None:
            print('Set params for iterate first')
            return
        par = itr.params
        itr.quat.num_rot_p = itr.quat.divide(0, 1, par.num_modes, par.nonrot_modes)
        par.rank, par.num_proc = 0, 1
        if itr.scale is None:
            itr.parse_scale('')
        if itr.beta_start is None:
            itr.calc_beta()
        self.iter = itr
        if self.engine is None:
            self.engine = _lib.Engine(self.device)
        eng = self.engine
        eng.set_precision(self.precision)
        det = itr.dets[0]
        eng.set_detector(det.qvals, det.corr, det.raw_mask)
        eng.set_modes(par.num_modes)
        eng.set_quats(itr.quat.quats)
        eng.set_model(itr.model.model1)
        if getattr(itr, 'dense', False):
            eng.set_frames_dense(itr.dense_frames())
        else:
            sp = itr.sparse_arrays()
            eng.set_frames(sp['ones'], sp['multi'], sp['ones_accum'], sp['multi_accum'],
                           sp['place_ones'], sp['place_multi'], sp['count_multi'])
        # iter->num_blacklist is never set on the reference's Cython path (iterate.pyx:35
        # callocs it and nothing writes it), so combine_information divides the summed
        # likelihood and information by tot_num_data (max_prob.c:373-374) and update_beta
        # its beta sum likewise (recon.pyx:172); blacklisted frames still contribute nothing.
        itr.num_blacklist = 0
        eng.set_frame_params(itr.scale, itr.beta_start, itr.blacklist, itr.fcounts,
                             itr.sum_fact, itr.mean_count[0], itr.num_blacklist)
        self._engine_beta = None  # the engine now holds beta_start
        if det.with_bg:
            if itr.bgscale is None:
                itr.parse_scale('', bg=True)
            eng.set_background(det.background, itr.bgscale)
        if itr.quat_mapping is not None:
            eng.set_refine(itr.num_coarse, itr.quat_mapping, itr.rel_off, itr.rel_quat)
        itr.model._bind(eng)

    def run_iteration(self, dynamic=False):
        stime = time.time()
        itr = self.iter
        if itr is None:
            print('Set iterate first')
            return
        par = itr.params
        if par.iteration == 1 and self.write_output:
            self.write_log_file_header()
        beta_mean = self.update_beta()
        stats = self.engine.maximize(par.iteration, par.need_scaling, par.update_scale,
                                     par.known_scale)
        self.last_stats = stats
        itr.rescale[0] = stats['rescale']
        itr.mutual_info = stats['mutual_info']
        itr.likelihood = stats['avg_likelihood']
        self.update_model()
        if par.need_scaling == 1:
            mean_scale = self.engine.normalize_scale()
            itr.rms_change *= mean_scale
        if par.need_scaling:
            itr.scale[:] = self.engine.get('scale')
        itr.model._mark_stale()
        if self.write_output:
            self.save_output(dynamic)
            self.update_log_file(itr.likelihood, beta_mean, time.time() - stime)
        if par.verbosity > 0:
            print('Finished iteration', par.iteration, '(%e)' % itr.rms_change)

    def update_model(self):
        """Friedel-symmetrise, divide, rms change, blend (recon.pyx:112-149), on the GPU."""
        self.iter.rms_change = self.engine.update_model(self.iter.params.alpha)

    def update_beta(self):
        """beta_d = min(beta_start_d * jump^((it-1)//period) * factor, 1) (recon.pyx:151-174)."""
        itr = self.iter
        par = itr.params
        factor = par.beta_jump ** ((par.iteration - 1) // par.beta_period)
        factor *= par.beta_factor
        good = itr.blacklist != 1
        beta = itr.beta
        beta[good] = np.minimum(itr.beta_start[good] * factor, 1.)
        # the engine's copy is refreshed only when beta changed (a set-up counts as a change
        # of the iteration's state, which an expectation step already run would then not match)
        last = getattr(self, '_engine_beta', None)
        if last is None or last.shape != beta.shape or not np.array_equal(last, beta):
            self.engine.set_beta(beta)
            self._engine_beta = beta.copy()
        return beta[good].sum() / (itr.tot_num_data - itr.num_blacklist)

    # ---- outputs ---------------------------------------------------------------------
    def write_log_file_header(self):
        itr = self.iter
        par = itr.params
        mask = itr.dets[0].raw_mask
        with open(par.log_fname, 'w') as fp:
            fp.write('Cryptotomography with the EMC algorithm on one MI355X GPU\n\n')
            fp.write('Data parameters:\n')
            if itr.num_blacklist == 0:
                fp.write('\tnum_data = %d\n\tmean_count = %f\n\n' % (itr.tot_num_data, itr.mean_count[0]))
            else:
                fp.write('\tnum_data = %d/%d\n\tmean_count = %f\n\n' % (
                    itr.tot_num_data - itr.num_blacklist, itr.tot_num_data, itr.mean_count[0]))
            fp.write('System size:\n')
            fp.write('\tnum_rot = %d\n\tnum_pix = %d/%d\n\t' % (itr.quat.num_rot, (mask == 0).sum(), mask.size))
            s = itr.model.size
            fp.write('system_volume = %d X %ld X %ld X %ld\n\n' % (itr.model.num_modes, s, s, s))
            fp.write('Reconstruction parameters:\n')
            fp.write('\tnum_threads = %d\n\tnum_proc = %d\n\talpha = %.6f\n\tbeta = %.6f\n\tneed_scaling = %s' % (
                self.num_threads, par.num_proc, par.alpha, itr.beta_start[0],
                'yes' if par.need_scaling == 1 else 'no'))
            fp.write('\n\nIter\ttime\trms_change\tinfo_rate\tlog-likelihood\tnum_rot\tbeta\n')

    def update_log_file(self, likelihood, beta, iter_time):
        itr = self.iter
        with open(itr.params.log_fname, 'a') as fp:
            fp.write('%d\t' % itr.params.iteration)
            fp.write('%4.2f\t' % iter_time)
            fp.write('%1.4e\t%f\t%.6e\t%-7d\t%f\n' % (itr.rms_change, itr.mutual_info, likelihood,
                                                      itr.quat.num_rot, beta))

    def _output_dict(self):
        itr = self.iter
        itr.model.sync()
        m1, w = itr.model.model1, itr.model.inter_weight
        if itr.model.num_modes == 1:
            m1, w = m1[None], w[None]
        out = {'orientations': self.rmax, 'mutual_info': self.info, 'likelihood': self.likelihood,
               'intens': m1, 'inter_weight': w, 'quaternions': itr.quat.quats}
        if itr.model.num_modes > 1:
            out['occupancies'] = self.quat_norm
        if itr.params.need_scaling != 0:
            out['scale'] = itr.scale
        for k, v in vars(itr.params).items():
            out['params/' + k] = np.asarray(v)
        out['params/num_rot'] = np.asarray(itr.quat.num_rot)
        out['params/num_div'] = np.asarray(itr.quat.num_div)
        if itr.params.save_prob:
            num, place, prob = self.sparse_prob()
            out['probabilities/num_rot'] = np.asarray([itr.quat.num_rot])
            out['probabilities/num_prob'] = num
            out['probabilities/place'] = place
            out['probabilities/prob'] = prob
        return out

    def save_output(self, dynamic=False):
        """Per-iteration output file (recon.pyx:230-279). Written as HDF5 with h5py when it is
        importable, else as .npz with the same dataset names."""
        par = self.iter.params
        os.makedirs(par.output_folder, exist_ok=True)
        stem = 'output_dynamic' if dynamic else 'output_%.3d' % par.iteration
        out = self._output_dict()
        try:
            import h5py
        except ImportError:
            h5py = None
        if h5py is not None:
            with h5py.File(os.path.join(par.output_folder, stem + '.h5'), 'w') as f:
                for k, v in out.items():
                    f[k] = v
        else:
            np.savez(os.path.join(par.output_folder, stem + '.npz'), **out)

    def save_initial_model(self):
        par = self.iter.params
        os.makedirs(par.output_folder, exist_ok=True)
        m1 = self.iter.model.model1
        out = {'intens': m1[None] if self.iter.model.num_modes == 1 else m1}
        if par.need_scaling != 0:
            out['scale'] = self.iter.scale
        np.savez(os.path.join(par.output_folder, 'output_000.npz'), **out)

    # ---- per-frame / per-orientation results (recon.pyx:294-415) -----------------------
    @property
    def max_exp(self):
        return self.engine.get('max_exp')

    @property
    def p_norm(self):
        return self.engine.get('p_norm')

    @property
    def u(self):
        return [self.engine.get('u')]

    @property
    def info(self):
        return self.engine.get('info')

    @property
    def likelihood(self):
        return self.engine.get('likelihood')

    @property
    def rmax(self):
        return self.engine.get('rmax')

    @property
    def psum_d(self):
        return self.engine.get('psum_d')

    @property
    def psum_r(self):
        return self.engine.get('psum_r')

    @property
    def quat_norm(self):
        """(num_data, num_modes) occupancy of each frame in each mode (max_update.c:122-124,
        recon.pyx:259-260); num_modes > 1 only."""
        return self.engine.get('quat_norm')

    @property
    def prob_dense(self):
        """(num_data, num_rot) posterior, 0 where an orientation was not kept."""
        return self.engine.get('prob')

    def sparse_prob(self):
        """(num_prob, place, prob) of the saved sparse posterior (P > PROB_MIN,
        max_prob.c:323-325), entries frame after frame, compacted on the GPU."""
        return self.engine.sparse_prob()

    @property
    def place_prob(self):
        """Per frame, orientations with P > PROB_MIN (recon.pyx:413-415)."""
        num, place, _ = self.sparse_prob()
        return np.split(place, np.cumsum(num)[:-1])

    @property
    def prob(self):
        """Per frame, the posterior of those orientations (recon.pyx:405-411)."""
        num, _, prob = self.sparse_prob()
        return np.split(prob, np.cumsum(num)[:-1])

    @property
    def num_prob(self):
        return self.sparse_prob()[0]

    def free(self):
        """Release the engine. The model's host arrays are refreshed first and then detached
        from it, so they stay readable (and assignable) afterwards."""
        if self.engine is not None:
            model = self.iter.model if self.iter is not None else None
            if model is not None and model._engine is self.engine:
                model.sync()
                model._engine = None
            self.engine.close()
            self.engine = None


def main():
    """Command-line entry point (recon.pyx:418-445)."""
    parser = argparse.ArgumentParser(description='EMC reconstruction on one MI355X GPU')
    parser.add_argument('niter', help='Number of iterations', type=int)
    parser.add_argument('-c', '--config_file', help='Path to configuration file. Default: config.ini',
                        default='config.ini')
    parser.add_argument('-r', '--resume', help='Whether to resume from last iteration', action='store_true')
    parser.add_argument('-t', '--threads', help='Kept for compatibility (unused on the GPU)', type=int,
                        default=-1)
    args = parser.parse_args()
    recon = EMCRecon(args.threads)
    itr = Iterate(args.config_file, resume=args.resume)
    itr.params.num_iter = args.niter
    recon.set_iterate(itr)
    st = itr.params.start_iter
    if st == 1:
        recon.save_initial_model()
    for itr.params.iteration in range(st, st + itr.params.num_iter):
        recon.run_iteration()
    if itr.params.verbosity > 0:
        print('Finished all iterations')


if __name__ == '__main__':
    main()
