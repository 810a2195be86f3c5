"""Quaternion: orientation samples of SO(3) (mirror of dragonfly/quaternion.pyx:25-175)."""
import sys

import numpy as np

from . import _lib
from .config import MyConfigParser


class Quaternion:
    """600-cell orientation sampling; ``quats`` is (num_rot, 5) = (q0, q1, q2, q3, weight).

    generate_3d calls the engine library's host-side quat_gen, which reproduces the reference
    (dragonfly/src/quaternion.c:811-842) bit for bit; it needs no GPU.
    """

    def __init__(self, num_div=0, num_rot=0, point_group=''):
        self.quats = None
        self.num_div = 0
        self.num_rot_p = 0
        self.reduced = False
        if point_group not in ('', '1'):
            raise NotImplementedError('point groups other than 1 are not implemented')
        if num_div > 0 and num_rot > 0:
            raise ValueError('Cannot specify both num_div and num_rot')
        if num_div > 0:
            self.generate_3d(num_div)
        elif num_rot > 0:
            raise NotImplementedError('2D in-plane rotations are not implemented')

    def from_config(self, config_fname, section_name='emc'):
        config = MyConfigParser()
        config.read(config_fname)
        rtype = config.get(section_name, 'recon_type', fallback='3d').lower()
        if rtype != '3d':
            raise NotImplementedError('only 3d orientation sampling is implemented')
        num_div = config.get(section_name, 'num_div', fallback='0').split()
        # "num_div = fine coarse" selects refinement (params.pyx:121-126): the fine sampling
        # is the orientation set, the coarse one indexes the per-frame subsets
        self.generate_3d(int(num_div[0]))
        self.coarse_div = int(num_div[1]) if len(num_div) > 1 else 0
        point_group = config.get(section_name, 'point_group', fallback='1')
        if point_group != '1':
            raise NotImplementedError('point group %s is not implemented' % point_group)
        return self

    def generate_3d(self, num_div):
        self.num_div = int(num_div)
        self.reduced = False
        self.quats = _lib.quat_gen(self.num_div)

    @property
    def num_rot(self):
        return 0 if self.quats is None else self.quats.shape[0]

    def divide(self, rank, num_proc, num_modes=1, num_nonrot_modes=0):
        """Orientations owned by one rank (quaternion.pyx:156-175), with integer division."""
        tot = num_modes * self.num_rot + num_nonrot_modes
        self.num_rot_p = tot // num_proc
        if rank < tot % num_proc:
            self.num_rot_p += 1
        if num_proc > 1:
            sys.stderr.write('%d: num_rot_p = %d/%d\n' % (rank, self.num_rot_p, tot))
        return self.num_rot_p

    def save(self, fname):
        np.save(fname, self.quats)
