"""Iterate: the per-iteration state (mirror of dragonfly/iterate.pyx:21-640).

Holds parameters, detector, photon data, orientations, model and the per-frame arrays, and
computes the per-frame precompute of dragonfly/src/iterate.c (frame counts, mean count,
sum of log-factorials, beta).
"""
import configparser
import math
import os.path as op

import numpy as np

from .config import MyConfigParser
from .detector import CDetector
from .emcfile import CDataset
from .model import Model
from .params import EMCParams
from .quaternion import Quaternion

_LNFACT = None


def lnfact(counts):
    """log(n!) as GSL's gsl_sf_lnfact: log of the double-rounded n! for n <= 170, else
    lgamma(n + 1)."""
    global _LNFACT
    if _LNFACT is None:
        _LNFACT = np.array([math.log(float(math.factorial(n))) for n in range(171)])
    counts = np.asarray(counts, dtype='i8')
    out = np.empty(counts.shape)
    small = counts <= 170
    out[small] = _LNFACT[counts[small]]
    out[~small] = [math.lgamma(n + 1.) for n in counts[~small]]
    return out


def frame_counts(ones, multi, place_ones, place_multi, count_multi, mask):
    """fcounts: photons per frame on mask < 1 pixels (calc_frame_counts, iterate.c:24-35)."""
    frame_o = np.repeat(np.arange(len(ones)), ones)
    frame_m = np.repeat(np.arange(len(multi)), multi)
    good_o = mask[place_ones] < 1
    good_m = mask[place_multi] < 1
    f = np.bincount(frame_o[good_o], minlength=len(ones)).astype('i8')
    f += np.bincount(frame_m[good_m], weights=count_multi[good_m], minlength=len(ones)).astype('i8')
    return f.astype('i4')


def sum_log_factorials(multi, place_multi, count_multi, mask):
    """sum_fact: sum of log(K!) over a frame's multi-photon pixels with mask < 1
    (calc_sum_fact, iterate.c:85-117), accumulated in pixel order."""
    out = np.zeros(len(multi))
    frame_m = np.repeat(np.arange(len(multi)), multi)
    good = mask[place_multi] < 1
    lf = lnfact(count_multi)
    np.add.at(out, frame_m[good], lf[good])
    return out


class Iterate:
    """State for each EMC iteration. Construct from a config file, or set the pieces with
    set_params / set_data / set_quat / set_model."""

    def __init__(self, config_fname='', section_name='emc', resume=False):
        self.params = EMCParams()
        self.dets = []
        self.data = None
        self.quat = None
        self.model = None
        self.tot_num_data = 0
        self.num_blacklist = 0
        self.fcounts = self.scale = self.bgscale = None
        self.beta = self.beta_start = self.sum_fact = self.blacklist = None
        self.mean_count = None
        self.rescale = None
        self.likelihood = self.mutual_info = self.rms_change = 0.
        # refinement: coarse cell of every orientation and per-frame coarse sets (CSR)
        self.quat_mapping = self.rel_off = self.rel_quat = None
        self.num_coarse = 0
        if config_fname != '':
            self.from_config(config_fname, section_name, resume=resume)

    # ---- configuration pipeline (iterate.pyx:40-182) ----------------------------------
    def from_config(self, config_fname, section_name='emc', resume=False):
        param = EMCParams().from_config(config_fname, section_name)
        self.set_params(param)
        config_folder = op.dirname(config_fname)
        config = MyConfigParser()
        config.read(config_fname)
        det_fname = config.get(section_name, 'in_detector_file', fallback=None)
        if det_fname is None:
            raise ValueError('Need in_detector_file (in_detector_list is not implemented).')
        if ':::' in det_fname:
            sname, oname = det_fname.split(':::')
            det_fname = config.get(sname, oname)
        det = CDetector(op.join(config_folder, det_fname), norm=True)
        ph_fname = config.get(section_name, 'in_photons_file', fallback=None)
        ph_flist = config.get(section_name, 'in_photons_list', fallback=None)
        if ph_fname is not None and ph_flist is not None:
            raise ValueError('Both in_photons_file and in_photons_list specified. Pick one.')
        if ph_fname is not None:
            if ':::' in ph_fname:
                sname, oname = ph_fname.split(':::')
                ph_fname = config.get(sname, oname)
            frames = CDataset(op.join(config_folder, ph_fname), det)
        elif ph_flist is not None:
            if ':::' in ph_flist:
                sname, oname = ph_flist.split(':::')
                ph_flist = config.get(sname, oname)
            with open(ph_flist) as fptr:
                fnames = [op.join(config_folder, line.strip()) for line in fptr if line.strip()]
            frames = CDataset(fnames[0], det)
            for fn in fnames[1:]:
                frames.append(CDataset(fn, det))
        else:
            raise ValueError('Need either in_photons_file or in_photons_list.')
        self.dets = [det]
        # detector background (iterate.pyx:106-125); one detector here, so a list must
        # name exactly one file, relative to the list's folder
        background_file = config.get_filename(section_name, 'background_file', fallback=None)
        background_list = config.get_filename(section_name, 'background_list', fallback=None)
        if background_file is not None and background_list is not None:
            raise ValueError('Both background_file and background_list specified. Pick one.')
        if background_list is not None:
            with open(background_list) as f:
                bg_lines = [line.strip() for line in f if line.strip()]
            if len(bg_lines) != len(self.dets):
                raise ValueError('Background list has %d entries but there are %d unique detectors'
                                 % (len(bg_lines), len(self.dets)))
            det.parse_background(op.join(op.dirname(background_list), bg_lines[0]))
        elif background_file is not None:
            det.parse_background(background_file)
        self.set_data(frames)
        self.set_quat(Quaternion().from_config(config_fname, section_name))
        try:
            model_file = config.get_filename(section_name, 'start_model_file')
        except configparser.NoOptionError:
            model_file = ''
        try:
            scale_file = config.get_filename(section_name, 'scale_file')
        except configparser.NoOptionError:
            scale_file = ''
        try:
            bgscale_file = config.get_filename(section_name, 'bgscale_file')
        except configparser.NoOptionError:
            bgscale_file = ''
        if resume:
            with open(param.log_fname) as fp:
                last_iter = int(fp.readlines()[-1].split()[0])
            self.params.start_iter = last_iter + 1
            model_file = self.output_file(param.output_folder, last_iter)
            scale_file = model_file
            print('Resuming from iteration', self.params.start_iter)
        model = Model(self.calculate_size(max(d.qmax() for d in self.dets)), self.params.num_modes)
        model_mean = self.mean_count[0] / (self.dets[0].raw_mask == 0).sum() * 2.
        model.allocate(model_file, model_mean, fixed_seed=bool(param.fixed_seed))
        self.set_model(model)
        if self.params.need_scaling == 1:
            self.parse_scale(scale_file)
            self.parse_scale(bgscale_file, bg=True)
        sel_string = config.get(section_name, 'selection', fallback=None)
        try:
            blacklist_file = config.get_filename(section_name, 'blacklist_file')
        except configparser.NoOptionError:
            blacklist_file = ''
        self.parse_blacklist(blacklist_file, sel_string)
        beta_str = config.get(section_name, 'beta', fallback='auto')
        self.calc_beta(None if beta_str == 'auto' else float(beta_str))
        if self.params.refine:
            try:
                probs = config.get_filename(section_name, 'probs_file')
            except configparser.NoOptionError:
                raise ValueError('Need coarse probabilities in probs_file to do refinement')
            self.setup_refine(self.params.coarse_div, self.parse_rel_quat(probs, self.params.coarse_div))
        return self

    # ---- refinement (legacy src/iterate.c:103-130, :364-525) ----------------------------
    def setup_refine(self, coarse_div, rel_quat):
        """Map every (fine) orientation to its nearest coarse one (voronoi_subset) and attach
        the per-frame coarse sets: rel_quat is a list of int arrays, one per frame."""
        if len(rel_quat) != self.tot_num_data:
            raise ValueError('rel_quat has %d frames, data has %d' % (len(rel_quat), self.tot_num_data))
        counts = np.array([len(r) for r in rel_quat], dtype='i8')
        rel_off = np.zeros(len(rel_quat) + 1, dtype='i8')
        np.cumsum(counts, out=rel_off[1:])
        flat = np.concatenate([np.asarray(r, 'i4') for r in rel_quat]) if len(rel_quat) \
            else np.zeros(0, 'i4')
        self.set_refine_lists(coarse_div, rel_off, flat)

    def set_refine_lists(self, coarse_div, rel_off, rel_quat):
        """As setup_refine with the sets already flattened: frame d's coarse orientations are
        rel_quat[rel_off[d]:rel_off[d + 1]]. The orientation -> cell map is kept while the
        orientations and coarse_div do not change."""
        from . import _lib
        rel_off = np.ascontiguousarray(rel_off, dtype='i8')
        rel_quat = np.ascontiguousarray(rel_quat, dtype='i4')
        if rel_off.shape != (self.tot_num_data + 1,) or rel_off[0] != 0 or rel_off[-1] != rel_quat.size:
            raise ValueError('rel_off must hold num_data + 1 ascending offsets into rel_quat')
        key = (int(coarse_div), id(self.quat), self.quat.num_rot)
        if getattr(self, '_qmap_key', None) != key:
            qcoarse = _lib.quat_gen(coarse_div)
            self.quat_mapping = _lib.voronoi_subset(qcoarse, self.quat.quats)
            self.num_coarse = qcoarse.shape[0]
            self._qmap_key = key
        self.rel_off = rel_off
        self.rel_quat = rel_quat  # entries are range-checked by the engine (emc_set_refine)

    def parse_rel_quat(self, fname, coarse_div):
        """Per-frame coarse orientation sets from a previous coarse run: the binary probs
        format (int num_data, int num_rot, 1024-byte header, int num_rel_quat[num_data], then
        the lists), or an output file with 'probabilities/num_prob' + 'probabilities/place'
        (sparse posteriors) or else 'orientations' (one per frame)."""
        from . import _lib
        num_rot_coarse = _lib.num_rot(coarse_div)
        with open(fname, 'rb') as f:
            head = f.read(8)
        if head == b'\x89HDF\r\n\x1a\n':
            try:
                import h5py
            except ImportError:
                raise TypeError('HDF5 probability files need h5py, which is not available here')
            with h5py.File(fname, 'r') as f:
                data = {k: f[k][()] for k in ('probabilities/num_prob', 'probabilities/place',
                                              'orientations') if k in f}
        elif fname.endswith('.npz'):
            with np.load(fname) as f:
                data = {k: f[k] for k in f.files}
        else:
            with open(fname, 'rb') as f:
                num_data, num_rot = (int(x) for x in np.fromfile(f, dtype='i4', count=2))
                if num_data != self.tot_num_data:
                    raise ValueError('num_data of rel_quat (%d) does not match' % num_data)
                if num_rot != num_rot_coarse:
                    raise ValueError('Rotation sampling in probs file does not match num_rot_coarse '
                                     '(%d vs %d)' % (num_rot, num_rot_coarse))
                f.seek(1024)
                counts = np.fromfile(f, dtype='i4', count=num_data)
                flat = np.fromfile(f, dtype='i4', count=int(counts.sum()))
            return np.split(flat, np.cumsum(counts)[:-1])
        if 'probabilities/num_prob' in data:
            counts = np.asarray(data['probabilities/num_prob'], 'i8')
            return np.split(np.asarray(data['probabilities/place'], 'i4'), np.cumsum(counts)[:-1])
        # frames without an orientation (rmax -1) match no coarse cell
        return [np.array([o] if o >= 0 else [], 'i4') for o in np.asarray(data['orientations']).ravel()]

    def set_params(self, param):
        self.params = param

    def set_quat(self, quaternion):
        self.quat = quaternion

    def set_model(self, model):
        self.model = model

    def set_data(self, in_dset):
        """Attach photon data and compute the per-frame precompute (iterate.pyx:241-266)."""
        if self.tot_num_data != 0:
            raise ValueError('Please use update_data() to change dataset')
        total = 0
        kinds = set()
        for ds in in_dset:
            ds.num_offset = total
            total += ds.num_data
            kinds.add(ds.ftype)
        if not kinds <= {'sparse', 'dense_integer'} or len(kinds) != 1:
            raise NotImplementedError('photon data must be all sparse or all dense-integer frames')
        self.tot_num_data = total
        self.data = in_dset
        self.dense = kinds == {'dense_integer'}
        if not self.dets:
            self.dets = [in_dset.det]
        if any(ds.det is not self.dets[0] for ds in in_dset):
            raise NotImplementedError('all datasets must share one detector')
        mask = self.dets[0].raw_mask
        if self.dense:
            fr = self.dense_frames()
            # calc_frame_counts counts every pixel of a dense frame (iterate.c:38-46);
            # calc_sum_fact only mask < 1 pixels (iterate.c:103-108)
            self.fcounts = fr.sum(1, dtype='i8').astype('i4')
            good = mask < 1
            # sum of log(K!) over mask < 1 pixels, from a table of log(k!) per count value
            table = lnfact(np.arange(int(fr.max()) + 1 if fr.size else 1))
            self.sum_fact = np.zeros(total)
            for d0 in range(0, total, 512):
                self.sum_fact[d0:d0 + 512] = table[fr[d0:d0 + 512][:, good]].sum(1)
        else:
            sp = self.sparse_arrays()
            self.fcounts = frame_counts(sp['ones'], sp['multi'], sp['place_ones'], sp['place_multi'],
                                        sp['count_multi'], mask)
            self.sum_fact = sum_log_factorials(sp['multi'], sp['place_multi'], sp['count_multi'], mask)
        self.mean_count = np.array([self.fcounts.astype('f8').sum() / total])
        self.rescale = np.zeros(1)
        self.blacklist = np.zeros(total, dtype='u1')

    def dense_frames(self):
        """All dense datasets stacked, (tot_num_data, num_pix) int32."""
        parts = [p.int_frames for p in self.data]
        if len(parts) == 1:
            return np.ascontiguousarray(parts[0], dtype='i4')
        return np.concatenate(parts).astype('i4')

    def sparse_arrays(self):
        """All datasets concatenated into one frame list (linked list order)."""
        parts = list(self.data)
        cat = {k: np.concatenate([getattr(p, k) for p in parts]).astype('i4')
               for k in ('ones', 'multi', 'place_ones', 'place_multi', 'count_multi')}
        from .emcfile import _accum
        cat['ones_accum'] = _accum(cat['ones'])
        cat['multi_accum'] = _accum(cat['multi'])
        return cat

    def parse_scale(self, fname, bg=False):
        """Scale factors (bg: background scale factors, bgscale) from a file, or ones
        (iterate.pyx:284-315)."""
        if not fname or not op.isfile(fname):
            if self.tot_num_data == 0:
                raise AttributeError('Need tot_num_data to initialize scale factors')
            scale = np.ones(self.tot_num_data)
        elif fname.endswith('.npz'):
            # the 'scale' dataset, for background scales too (iterate.pyx:295-297)
            with np.load(fname) as f:
                scale = np.asarray(f['scale'], dtype='f8')
        else:
            with open(fname, 'rb') as f:
                is_hdf5 = f.read(8) == b'\x89HDF\r\n\x1a\n'
            if is_hdf5:
                try:
                    import h5py
                except ImportError:
                    raise TypeError('%s is an HDF5 file and h5py is not available' % fname)
                with h5py.File(fname, 'r') as f:
                    scale = np.asarray(f['scale'][:], dtype='f8')
            else:
                scale = np.loadtxt(fname, dtype='f8').ravel()
        if self.tot_num_data < scale.shape[0]:
            print('More scale factors than required. Taking only first', self.tot_num_data)
            scale = scale[:self.tot_num_data]
        elif self.tot_num_data > scale.shape[0]:
            print('Insufficient scale factors in file. Setting rest to unity')
            scale = np.append(scale, np.ones(self.tot_num_data - scale.shape[0]))
        if bg:
            self.bgscale = np.ascontiguousarray(scale, dtype='f8').copy()
        else:
            self.scale = np.ascontiguousarray(scale, dtype='f8').copy()

    def parse_blacklist(self, fname, sel_string=None, refresh=False):
        """Blacklist from a 0/1 text file and/or odd_only / even_only (iterate.pyx:317-359)."""
        if self.tot_num_data == 0:
            raise AttributeError('Need to define tot_num_data before generating blacklist')
        if self.blacklist is None or refresh:
            self.blacklist = np.zeros(self.tot_num_data, dtype='u1')
        if fname and op.isfile(fname):
            arr = np.loadtxt(fname, dtype='u1')
            if arr.shape[0] != self.tot_num_data:
                raise ValueError('Mismatched number of frames in blacklist file')
            self.blacklist[:] = arr
        if sel_string is None:
            return
        if sel_string == 'odd_only':
            curr = 0
        elif sel_string == 'even_only':
            curr = 1
        else:
            raise ValueError('Unrecognized sel_string, %s' % sel_string)
        for d in range(self.tot_num_data):
            if self.blacklist[d] == 0:
                self.blacklist[d] = curr
                curr = 1 - curr

    def normalize_scale(self):
        """Mean-normalise scale factors into the model (iterate.pyx:361-384), host arrays."""
        good = self.blacklist == 0
        mean_scale = self.scale[good].mean()
        if self.model is not None and self.model.model1 is not None:
            self.model.model1 *= mean_scale
        self.scale[good] /= mean_scale
        self.rms_change *= mean_scale
        return mean_scale

    def calc_beta(self, setval=None):
        """beta_start per frame (calc_beta, iterate.c:66-83)."""
        if self.data is None:
            raise AttributeError('Add data first before calculating beta for each frame')
        start = -1. if setval is None else float(setval)
        if not self.params.need_scaling and start < 0:
            start = math.exp(-6.5 * math.pow(self.mean_count[0] * 1.e-5, 0.15))
        if start > 0:
            self.beta_start = np.full(self.tot_num_data, start)
        else:
            self.beta_start = np.array([math.exp(-6.5 * math.pow(f * 1.e-5, 0.15))
                                        for f in self.fcounts.tolist()])
        self.beta = self.beta_start.copy()

    @staticmethod
    def output_file(output_folder, iteration):
        """The output file of an iteration: output_NNN.h5 as the reference writes it
        (recon.pyx:230-279), or output_NNN.npz where h5py is absent."""
        stem = op.join(output_folder, 'output_%.3d' % iteration)
        for ext in ('.h5', '.npz'):
            if op.isfile(stem + ext):
                return stem + ext
        raise FileNotFoundError('No output file %s.h5 or %s.npz to resume from' % (stem, stem))

    @staticmethod
    def calculate_size(qmax):
        return int(2 * np.ceil(qmax) + 3)
