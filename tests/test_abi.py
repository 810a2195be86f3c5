"""libemc_hip.so loads and exports every entry point include/emc_hip.h declares; the
host-only entry points work without a GPU; the ctypes structs match the header."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

from dragonfly_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, 'include', f) for f in os.listdir(os.path.join(ROOT, 'include'))
           if f.endswith('.h')]


def _declared():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
        names |= set(re.findall(r'\b(emc_[a-z_0-9]+)\s*\(', text))
    return names


def test_library_exports_declared_symbols():
    lib = _lib.lib()
    declared = _declared()
    assert len(declared) >= 20
    for name in sorted(declared):
        assert hasattr(lib, name), name
    out = subprocess.run(['nm', '-D', '--defined-only', _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r' T (emc_[a-z_0-9]+)', out))
    assert declared <= exported, declared - exported


def test_host_entry_points_without_gpu():
    assert _lib.num_rot(4) == 3240
    with np.testing.assert_raises(ValueError):
        _lib.num_rot(0)
    q = _lib.quat_gen(2)
    assert q.shape == (420, 5)
    assert _lib.lib().emc_quat_gen(0, q.ctypes.data_as(C.POINTER(C.c_double))) < 0


def test_struct_layouts_match_header():
    text = open(os.path.join(ROOT, 'include', 'emc_hip.h')).read()
    opts = re.search(r'typedef struct emc_iter_opts \{(.*?)\}', text, re.S).group(1)
    stats = re.search(r'typedef struct emc_iter_stats \{(.*?)\}', text, re.S).group(1)
    assert [f[0] for f in _lib.IterOpts._fields_] == re.findall(r'int (\w+);', opts)
    fields = re.findall(r'(?:double|int) (\w+)', stats)
    assert [f[0] for f in _lib.IterStats._fields_] == fields
    nstages = int(re.search(r'#define EMC_NUM_STAGES (\d+)', text).group(1))
    assert nstages == len(_lib.STAGES) == _lib.IterStats.stage_ms.size // 8
    enum = re.search(r'enum emc_field \{(.*?)\};', text, re.S).group(1)
    enum = re.sub(r'/\*.*?\*/', '', enum, flags=re.S)
    names = [n.strip().split('=')[0].strip() for n in enum.split(',') if n.strip()]
    names = [n for n in names if n != 'EMC_FIELD_COUNT']
    by_id = {v[0]: k for k, v in _lib.FIELDS.items()}
    for i, n in enumerate(names):
        assert by_id[i].upper() == n.replace('EMC_FIELD_', ''), (i, n)
