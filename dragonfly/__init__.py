"""The reference's import name (dragonfly/__init__.py:1-7) for the MI355X engine.

`import dragonfly`, `from dragonfly.iterate import Iterate` and the like resolve to the
modules of dragonfly_amd, so scripts written against the reference's Python objects run
unchanged; dragonfly.recon is a module of its own so that `python -m dragonfly.recon` runs
the reconstruction command.
Names of the reference that sit outside one EMC iteration (Detector, EMCReader: the viewers'
file readers) are not provided.
"""
import sys

import dragonfly_amd
from dragonfly_amd import detector, emcfile, iterate, model, params, quaternion
from dragonfly_amd import CDataset, CDetector, EMCParams, EMCRecon, EMCWriter, Iterate, Model, Quaternion

for _mod in (detector, emcfile, iterate, model, params, quaternion):
    sys.modules[__name__ + '.' + _mod.__name__.rsplit('.', 1)[1]] = _mod
del _mod

__all__ = ['CDataset', 'CDetector', 'EMCParams', 'EMCRecon', 'EMCWriter', 'Iterate', 'Model',
           'Quaternion']
