"""Model: the 3D intensity volumes (mirror of dragonfly/model.pyx:18-313, 3D)."""
import os

import numpy as np

_HDF5_SIGNATURE = b'\x89HDF\r\n\x1a\n'


def _read_intens(fname):
    """The 'intens' dataset of an output file (.npz written here, or HDF5 as the reference
    writes, recon.pyx:230-279), or None when fname is a raw float64 volume."""
    if fname.endswith('.npz'):
        with np.load(fname) as f:
            return np.asarray(f['intens'], dtype='f8')
    with open(fname, 'rb') as f:
        if f.read(8) != _HDF5_SIGNATURE:
            return None
    try:
        import h5py
    except ImportError:
        raise TypeError('%s is an HDF5 file and h5py is not available' % fname)
    with h5py.File(fname, 'r') as f:
        return np.asarray(f['intens'][()], dtype='f8')


class Model:
    """model1 (current estimate), model2 (merged numerator) and inter_weight, each
    (size, size, size) float64 for one mode and (num_modes, size, size, size) for several
    (model.pyx:58), index x*size*size + y*size + z within a mode as the reference."""

    def __init__(self, size=0, num_modes=1, model_type='3d'):
        if model_type.lower() != '3d':
            raise NotImplementedError('only 3d models are implemented')
        if num_modes < 1:
            raise ValueError('num_modes must be at least 1')
        self.mtype = 'MODEL_3D'
        self.num_modes = num_modes
        self.size = int(size)
        self._engine = None
        self._stale = False
        self._model1 = self._model2 = self._weight = None

    # While bound to an engine the live volumes are on the GPU; reading one after an
    # iteration copies it back, assigning model1 sends it to the engine.
    @property
    def model1(self):
        self.sync()
        return self._model1

    @model1.setter
    def model1(self, arr):
        self._model1 = arr
        if self._engine is not None and arr is not None:
            self.sync()
            self._model1 = arr
            self._engine.set_model(arr)

    @property
    def model2(self):
        self.sync()
        return self._model2

    @model2.setter
    def model2(self, arr):
        self._model2 = arr

    @property
    def inter_weight(self):
        self.sync()
        return self._weight

    @inter_weight.setter
    def inter_weight(self, arr):
        self._weight = arr

    @property
    def center(self):
        return self.size // 2

    @property
    def vol(self):
        return self.size ** 3

    def allocate(self, fname='', model_mean=1., rank=0, fixed_seed=False):
        """Load model1 from a raw float64 file, or draw a random start (model.pyx:43-103):
        uniform [0, model_mean), numpy's legacy generator seeded 0x5EED when fixed_seed."""
        shape = (self.size,) * 3 if self.num_modes == 1 else (self.num_modes,) + (self.size,) * 3
        self.model2 = np.zeros(shape)
        self.inter_weight = np.zeros(shape)
        if fname and os.path.isfile(fname):
            intens = _read_intens(fname)
            if intens is not None:
                # an output file's 'intens' (num_modes, size, size, size): extra modes are
                # dropped, missing ones drawn uniform [0, model_mean) (model.pyx:67-82)
                mshape = (self.num_modes,) + (self.size,) * 3
                intens = intens.reshape((-1,) + intens.shape[-3:])
                if intens.shape[1:] != mshape[1:]:
                    raise ValueError('Input model has wrong size (%d)' % intens.shape[1], intens.shape[1])
                if intens.shape[0] < self.num_modes:
                    extra = np.random.random((self.num_modes - intens.shape[0],) + mshape[1:]) * model_mean
                    intens = np.concatenate([intens, extra])
                self.model1 = np.ascontiguousarray(intens[:self.num_modes], dtype='f8').reshape(shape).copy()
                return self
            if fname.endswith('.npy'):
                arr = np.load(fname)
            else:
                raw = np.fromfile(fname, dtype='f8')
                fsize = int(np.rint(raw.size ** (1 / 3.)))
                if fsize != self.size:
                    raise ValueError('Wrong volume size in binary file (%d)' % fsize, fsize)
                arr = raw
            self.model1 = np.ascontiguousarray(arr, dtype='f8').reshape(shape).copy()
        else:
            rs = np.random.RandomState(0x5EED) if fixed_seed else np.random
            self.model1 = rs.random_sample(shape) * model_mean
        return self

    # device mirror: the engine owns the live copies while a reconstruction runs
    def _bind(self, engine):
        self._engine = engine
        self._stale = False

    def _mark_stale(self):
        self._stale = True

    def sync(self):
        """Refresh the host arrays from the engine after an iteration."""
        if self._engine is not None and self._stale:
            self._stale = False
            self._model1 = self._engine.get('model1')
            self._model2 = self._engine.get('model2')
            self._weight = self._engine.get('inter_weight')
        return self

    def _slice_engine(self, det, vol):
        """A private engine holding `det` and one mode's volume `vol` (never the engine of a
        running reconstruction, whose state these calls would disturb)."""
        from ._lib import Engine
        eng = getattr(self, '_slicer', None)
        if eng is None:
            eng = self._slicer = Engine(0)
        eng.set_detector(det.qvals, det.corr, det.raw_mask)
        eng.set_model(np.ascontiguousarray(vol, dtype='f8'))
        return eng

    def _mode_view(self, arr, mode):
        if not 0 <= mode < self.num_modes:
            raise ValueError('mode %d out of range (num_modes %d)' % (mode, self.num_modes))
        return arr if self.num_modes == 1 else arr[mode]

    def slice_gen(self, quat, det, mode=0, view=None):
        """View of model1 (mode `mode`) at one orientation on `det` (Model.slice_gen,
        model.pyx:105-131 -> slice_gen3d, model.c:11-59), on the GPU; written into `view`
        when given."""
        if self.model1 is None:
            raise AttributeError('Allocate model1 first')
        eng = self._slice_engine(det, self._mode_view(self.model1, mode))
        out = eng.slice_gen(np.ascontiguousarray(quat, dtype='f8')[:4])
        if view is None:
            return out
        view[:] = out
        return view

    def slice_merge(self, quat, view, det, mode=0):
        """Merge one view into model2 / inter_weight of mode `mode` (Model.slice_merge,
        model.pyx:133-150 -> slice_merge3d, model.c:171-244). The view is merged on the GPU
        into zeroed volumes, which are then added to the host arrays."""
        if self.model2 is None:
            raise AttributeError('Allocate model2 first')
        eng = self._slice_engine(det, self._mode_view(self.model2, mode))
        eng.clear_merge()
        eng.slice_merge(np.ascontiguousarray(quat, dtype='f8')[:4], view)
        self._mode_view(self.model2, mode)[...] += eng.get('model2').reshape((self.size,) * 3)
        self._mode_view(self.inter_weight, mode)[...] += eng.get('inter_weight').reshape((self.size,) * 3)
