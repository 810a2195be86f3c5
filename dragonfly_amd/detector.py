"""CDetector: detector geometry (mirror of dragonfly/detector.pyx:15-222)."""
import os

import numpy as np


class CDetector:
    """Pixel q-vectors (voxel units), correction factors and mask (0 good, 1 ignored by the
    likelihood, 2 bad).

    The reference reads an HDF5 file with datasets qx, qy, qz, corr, mask, detd, ewald_rad
    (detector.pyx:31-74). The same datasets are read from HDF5 when h5py is importable, or
    from a ``.npz`` archive with those keys; arrays can also be set directly.
    """

    def __init__(self, fname=None, norm=True, rtype='3d'):
        self.fname = fname or ''
        self._qvals = self._corr = self._mask = None
        self.detd = 0.
        self.ewald_rad = 0.
        self.with_bg = 0
        self._background = None
        if fname is not None:
            self.parse(fname, norm=norm, rtype=rtype)

    def parse(self, fname, norm=True, rtype='3d'):
        if rtype != '3d':
            raise NotImplementedError('only 3d detectors are implemented')
        self.fname = fname
        if fname.endswith('.npz'):
            with np.load(fname) as f:
                data = {k: f[k] for k in f.files}
        else:
            try:
                import h5py
            except ImportError as exc:
                raise TypeError('reading an HDF5 detector needs h5py; use a .npz with the same '
                                'datasets') from exc
            with h5py.File(fname, 'r') as f:
                data = {k: f[k][()] for k in ('qx', 'qy', 'qz', 'corr', 'mask', 'detd', 'ewald_rad')}
        corr = np.asarray(data['corr'], dtype='f8').ravel().copy()
        if norm:
            corr /= corr.mean()
        self._qvals = np.ascontiguousarray(
            np.array([np.ravel(data['qx']), np.ravel(data['qy']), np.ravel(data['qz'])]).T, dtype='f8')
        self._corr = corr
        self._mask = np.asarray(data['mask']).astype('u1').ravel()
        self.detd = float(data['detd'])
        self.ewald_rad = float(data['ewald_rad'])

    def write(self, fname):
        """Save in the .npz layout read by parse()."""
        q = self.qvals
        np.savez(fname, qx=q[:, 0], qy=q[:, 1], qz=q[:, 2], corr=self.corr, mask=self.raw_mask,
                 detd=self.detd, ewald_rad=self.ewald_rad)

    def qmax(self):
        """Largest |q| over all pixels (detector.pyx:106-117)."""
        return float(np.sqrt((self.qvals ** 2).sum(1).max()))

    @property
    def num_pix(self):
        return 0 if self._corr is None else self._corr.shape[0]

    def _check(self, n):
        if self.num_pix > 0 and n != self.num_pix:
            raise ValueError('num_pix mismatch with other data (%d vs %d)' % (n, self.num_pix))

    @property
    def qvals(self):
        return self._qvals

    @qvals.setter
    def qvals(self, arr):
        arr = np.asarray(arr)
        if arr.ndim != 2 or arr.shape[1] != 3:
            raise ValueError('qvals must be  2D array of shape (N, 3)')
        if arr.dtype != np.float64:
            raise TypeError('qvals must be double precision floats (float64)')
        if self._corr is not None:
            self._check(arr.shape[0])
        self._qvals = np.ascontiguousarray(arr).copy()

    @property
    def corr(self):
        return self._corr

    @corr.setter
    def corr(self, arr):
        arr = np.asarray(arr)
        if arr.ndim != 1 or arr.dtype != np.float64:
            raise ValueError('corr must be  1D array of float64 dtype')
        if self._qvals is not None and arr.shape[0] != self._qvals.shape[0]:
            raise ValueError('num_pix mismatch with other data (%d vs %d)' % (arr.shape[0], self._qvals.shape[0]))
        self._corr = arr.copy()

    @property
    def raw_mask(self):
        return self._mask

    @raw_mask.setter
    def raw_mask(self, arr):
        arr = np.asarray(arr)
        if arr.ndim != 1 or arr.dtype != np.uint8:
            raise ValueError('raw_mask must be  1D array of uint8 dtype')
        n = self._corr.shape[0] if self._corr is not None else (
            self._qvals.shape[0] if self._qvals is not None else arr.shape[0])
        if arr.shape[0] != n:
            raise ValueError('num_pix mismatch with other data (%d vs %d)' % (arr.shape[0], n))
        self._mask = arr.copy()

    @property
    def background(self):
        """Per-pixel background (detector.pyx:207-220), None without one."""
        return self._background

    @background.setter
    def background(self, arr):
        if arr is None:
            self._background, self.with_bg = None, 0
            return
        arr = np.ascontiguousarray(arr, dtype='f8').ravel()
        if arr.shape[0] != self.num_pix:
            raise ValueError('background must have num_pix entries')
        self._background, self.with_bg = arr.copy(), 1

    def parse_background(self, fname):
        """The 'background' dataset of a .npz (or HDF5, with h5py) file (detector.pyx:101-104)."""
        if fname.endswith('.npz'):
            with np.load(fname) as f:
                self.background = f['background']
        else:
            import h5py
            with h5py.File(fname, 'r') as f:
                self.background = f['background'][:]

    @classmethod
    def from_arrays(cls, qvals, corr, raw_mask, detd=0., ewald_rad=0., fname='in-memory'):
        det = cls()
        det.qvals = np.asarray(qvals, dtype='f8')
        det.corr = np.asarray(corr, dtype='f8')
        det.raw_mask = np.asarray(raw_mask, dtype='u1')
        det.detd, det.ewald_rad, det.fname = float(detd), float(ewald_rad), fname
        return det

    def __repr__(self):
        return 'CDetector(%s, num_pix=%d)' % (os.path.basename(self.fname), self.num_pix)
