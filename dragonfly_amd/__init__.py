"""MI355X-native EMC engine with the reference's Python interface.

Objects mirror duaneloh/Dragonfly's (dragonfly/__init__.py:1-6): EMCRecon, Iterate, Model,
CDetector, CDataset/EMCWriter, Quaternion, EMCParams. One EMC iteration runs as hand-written
HIP kernels for gfx950 in libemc_hip.so, called through a C ABI (include/emc_hip.h).
"""
from .detector import CDetector
from .emcfile import CDataset, EMCWriter
from .iterate import Iterate
from .model import Model
from .params import EMCParams
from .quaternion import Quaternion
from .recon import EMCRecon

__all__ = ['CDetector', 'CDataset', 'EMCWriter', 'Iterate', 'Model', 'EMCParams', 'Quaternion',
           'EMCRecon']
