"""Sparse photon data: the binary .emc format (mirror of dragonfly/emcfile.pyx:18-616 and
parse_binarydataset, dragonfly/src/emcfile.c:3-57).

Layout: a 1024-byte header of int32 (num_data, num_pix, ftype, ...), then for sparse files
int32 arrays ones[num_data], multi[num_data], place_ones[sum ones], place_multi[sum multi],
count_multi[sum multi]; for dense-int files int32 frames[num_data][num_pix].
"""
import os

import numpy as np

SPARSE, DENSE_INT, DENSE_DOUBLE = 0, 1, 2
_FTYPES = ['sparse', 'dense_integer', 'dense_double']
HEADER_BYTES = 1024


def _accum(counts):
    out = np.zeros(len(counts), dtype='i8')
    if len(counts) > 1:
        np.cumsum(counts[:-1], out=out[1:])
    return out


class CDataset:
    """One photon file; datasets chain through ``next`` as the reference's linked list."""

    def __init__(self, fname=None, det=None):
        self.fname = fname or ''
        self.det = det
        self.next = None
        self.num_offset = 0
        self._ftype = SPARSE
        self.num_data = 0
        self.num_pix = 0
        self.ones = self.multi = self.place_ones = self.place_multi = self.count_multi = None
        self.int_frames = None
        if fname is not None:
            self.parse(fname)

    def parse(self, fname):
        if self.det is None:
            raise AttributeError('Set detector before parsing')
        self.fname = fname
        with open(fname, 'rb') as f:
            head = f.read(8)
        if head == b'\x89HDF\r\n\x1a\n':
            raise TypeError('HDF5 photon files need h5py, which is not available here')
        with open(fname, 'rb') as f:
            hdr = np.fromfile(f, dtype='i4', count=3)
            self.num_data, self.num_pix, self._ftype = int(hdr[0]), int(hdr[1]), int(hdr[2])
            f.seek(HEADER_BYTES)
            if self._ftype == SPARSE:
                self.ones = np.fromfile(f, dtype='i4', count=self.num_data)
                self.multi = np.fromfile(f, dtype='i4', count=self.num_data)
                self.place_ones = np.fromfile(f, dtype='i4', count=int(self.ones.sum()))
                self.place_multi = np.fromfile(f, dtype='i4', count=int(self.multi.sum()))
                self.count_multi = np.fromfile(f, dtype='i4', count=int(self.multi.sum()))
            elif self._ftype == DENSE_INT:
                self.int_frames = np.fromfile(f, dtype='i4', count=self.num_data * self.num_pix)
                self.int_frames = self.int_frames.reshape(self.num_data, self.num_pix)
            else:
                raise NotImplementedError('dense double frames are not implemented')
        if self.det is not None and self.num_pix != self.det.num_pix:
            print('WARNING! The detector file and photons file %s do not have the same number '
                  'of pixels' % fname)

    @classmethod
    def from_arrays(cls, det, ones, multi, place_ones, place_multi, count_multi, fname='in-memory'):
        ds = cls(det=det)
        ds.fname = fname
        ds.ones = np.ascontiguousarray(ones, dtype='i4')
        ds.multi = np.ascontiguousarray(multi, dtype='i4')
        ds.place_ones = np.ascontiguousarray(place_ones, dtype='i4')
        ds.place_multi = np.ascontiguousarray(place_multi, dtype='i4')
        ds.count_multi = np.ascontiguousarray(count_multi, dtype='i4')
        ds.num_data = ds.ones.shape[0]
        ds.num_pix = det.num_pix if det is not None else 0
        if ds.place_ones.size != ds.ones.sum() or ds.place_multi.size != ds.multi.sum() or \
                ds.count_multi.size != ds.place_multi.size:
            raise ValueError('sparse arrays are inconsistent with ones/multi counts')
        return ds

    @classmethod
    def from_dense(cls, det, frames, fname='in-memory'):
        """Dense integer frames (ftype DENSE_INT), shape (num_data, num_pix)."""
        ds = cls(det=det)
        ds.fname = fname
        ds._ftype = DENSE_INT
        ds.int_frames = np.ascontiguousarray(frames, dtype='i4')
        ds.num_data, ds.num_pix = ds.int_frames.shape
        return ds

    def append(self, next_dset):
        curr = self
        while curr.next is not None:
            curr = curr.next
        curr.next = next_dset

    def __iter__(self):
        curr = self
        while curr is not None:
            yield curr
            curr = curr.next

    @property
    def ftype(self):
        return _FTYPES[self._ftype]

    @property
    def ones_accum(self):
        return _accum(self.ones)

    @property
    def multi_accum(self):
        return _accum(self.multi)

    @property
    def ones_total(self):
        return int(self.ones.sum())

    @property
    def multi_total(self):
        return int(self.multi.sum())

    @property
    def mean_count(self):
        """Mean masked photon count per frame (emcfile.c:171-193)."""
        m = self.det.raw_mask
        if self._ftype == DENSE_INT:
            return self.int_frames[:, m < 1].sum() / self.num_data
        tot = (m[self.place_ones] < 1).sum() + self.count_multi[m[self.place_multi] < 1].sum()
        return tot / self.num_data


class EMCWriter:
    """Writes sparse .emc files (mirror of EMCWriter, emcfile.pyx:418-616, binary branch)."""

    def __init__(self, out_fname, num_pix, hdf5=False):
        if hdf5:
            raise NotImplementedError('HDF5 photon output needs h5py')
        self.out_fname = out_fname
        self.num_pix = int(num_pix)
        self._ones, self._multi, self._po, self._pm, self._cm = [], [], [], [], []

    def __enter__(self):
        return self

    def __exit__(self, etype, val, tb):
        self.finish_write()

    @property
    def num_data(self):
        return len(self._ones)

    def write_frame(self, frame, fraction=1., partition=1):
        frame = np.asarray(frame)
        if frame.ndim != 1 or frame.shape[0] != self.num_pix or not np.issubdtype(frame.dtype, np.integer):
            raise ValueError('write_frame needs 1D array of integers: ' + str(frame.shape))
        if fraction != 1. or partition != 1:
            raise NotImplementedError('photon sub-sampling is not implemented')
        place_ones = np.where(frame == 1)[0]
        place_multi = np.where(frame > 1)[0]
        self.write_sparse_frame(place_ones, place_multi, frame[place_multi])

    def write_sparse_frame(self, place_ones, place_multi, count_multi):
        self._ones.append(len(place_ones))
        self._multi.append(len(place_multi))
        self._po.append(np.asarray(place_ones, dtype='i4'))
        self._pm.append(np.asarray(place_multi, dtype='i4'))
        self._cm.append(np.asarray(count_multi, dtype='i4'))

    def finish_write(self, header_nums=None):
        if self.num_data == 0:
            print('No frames to write')
            return
        header = np.zeros(HEADER_BYTES // 4, dtype='i4')
        header[0] = self.num_data
        header[1] = self.num_pix
        if header_nums is not None:
            vals = np.asarray(header_nums).astype('i4')[:254]
            header[2:2 + len(vals)] = vals
        d = os.path.dirname(self.out_fname)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(self.out_fname, 'wb') as f:
            header.tofile(f)
            np.asarray(self._ones, 'i4').tofile(f)
            np.asarray(self._multi, 'i4').tofile(f)
            for arrs in (self._po, self._pm, self._cm):
                for a in arrs:
                    a.tofile(f)


def write_dense(fname, frames):
    """Write dense integer frames (num_data, num_pix) as a DENSE_INT .emc file."""
    frames = np.ascontiguousarray(frames, dtype='i4')
    header = np.zeros(HEADER_BYTES // 4, dtype='i4')
    header[0], header[1], header[2] = frames.shape[0], frames.shape[1], DENSE_INT
    with open(fname, 'wb') as f:
        header.tofile(f)
        frames.tofile(f)


def write_sparse(fname, num_pix, ones, multi, place_ones, place_multi, count_multi):
    """Write already-concatenated sparse arrays to a binary .emc file."""
    header = np.zeros(HEADER_BYTES // 4, dtype='i4')
    header[0], header[1] = len(ones), num_pix
    with open(fname, 'wb') as f:
        header.tofile(f)
        for a in (ones, multi, place_ones, place_multi, count_multi):
            np.asarray(a, dtype='i4').tofile(f)
