"""Write the golden fixtures of the reference's own orientation sampling and geometry.

Runs the reference's dragonfly/src/quaternion.c and model.c, compiled unchanged into
oracle/_ref/libdfref.so by `make -C oracle ref` (needs the reference sources, so this script
runs only where they are mounted). Writes, next to this script:

  quat_sha256.json  sha256 of the raw float64 bytes of quat_gen(n) for n in NUM_DIVS, plus
                    the row count and a few rows in full
  geometry.npz      inputs and outputs of slice_gen3d, slice_merge3d, make_rot_quat and
                    symmetrize_friedel on small seeded cases (odd and even grid sizes)
  voronoi_sha256.json  sha256 of the int32 output of voronoi_subset(quat_gen(c), quat_gen(f))
                    for the (coarse, fine) pairs in VORONOI, plus its first entries

`python make_golden.py voronoi` writes only the last file.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

NUM_DIVS = (1, 2, 3, 4, 5, 6, 7, 10, 12, 16)
VORONOI = ((1, 3), (2, 5), (4, 8))


def write_voronoi(ref):
    out = {}
    for c, f in VORONOI:
        m = ref.voronoi_subset(ref.quat_gen(c), ref.quat_gen(f)).astype('<i4')
        out['%d_%d' % (c, f)] = dict(num_fine=int(m.size), sha256=hashlib.sha256(m.tobytes()).hexdigest(),
                                     head=m[:16].tolist())
    with open(os.path.join(HERE, 'voronoi_sha256.json'), 'w') as fh:
        json.dump(out, fh, indent=1)


def main():
    from oracle import ref
    if not ref.available():
        raise SystemExit('oracle/_ref/libdfref.so missing: run `make -C oracle ref` first')
    write_voronoi(ref)
    if sys.argv[1:] == ['voronoi']:
        return
    quats = {}
    for n in NUM_DIVS:
        q = ref.quat_gen(n)
        quats[str(n)] = dict(num_rot=int(q.shape[0]), sha256=hashlib.sha256(q.tobytes()).hexdigest(),
                             rows={str(i): q[i].tolist() for i in (0, 1, q.shape[0] // 2, q.shape[0] - 1)})
    with open(os.path.join(HERE, 'quat_sha256.json'), 'w') as f:
        json.dump(quats, f, indent=1)

    rng = np.random.default_rng(20261202)
    out = {}
    for tag, S, N in (('odd', 15, 400), ('even', 14, 300)):
        model = rng.random((S, S, S))
        model[rng.random((S, S, S)) < 0.05] = 0.
        qv = rng.uniform(-(S - 1) / 2. - 1.5, (S - 1) / 2. + 1.5, (N, 3))
        corr = rng.uniform(0.5, 1.5, N)
        mask = rng.integers(0, 3, N).astype('u1')
        q = ref.quat_gen(3)[[0, 17, 333, 1000, 1379]]
        views = np.array([ref.slice_gen3d(qq, model, qv, corr, mask) for qq in q])
        rots = np.array([ref.make_rot_quat(qq) for qq in q])
        sym_in = rng.random((S, S, S))
        sym_w = rng.random((S, S, S))
        sym_w[sym_w < 0.1] = 0.
        sym_out = sym_in.copy()
        ref.symmetrize_friedel(sym_out, sym_w)
        out.update({tag + '_model': model, tag + '_qvals': qv, tag + '_corr': corr,
                    tag + '_mask': mask, tag + '_quats': q, tag + '_views': views,
                    tag + '_rots': rots, tag + '_sym_in': sym_in, tag + '_sym_w': sym_w,
                    tag + '_sym_out': sym_out})
        # merge the views scaled by fixed per-pixel factors
        fac = np.linspace(0.5, 2., N)
        m2 = np.zeros((S,) * 3)
        w2 = np.zeros((S,) * 3)
        for qq, v in zip(q, views):
            ref.slice_merge3d(qq, v * fac, m2, w2, qv, corr, mask)
        out[tag + '_merge_model'] = m2
        out[tag + '_merge_weight'] = w2
    np.savez_compressed(os.path.join(HERE, 'geometry.npz'), **out)
    print('wrote', os.path.join(HERE, 'quat_sha256.json'), os.path.join(HERE, 'geometry.npz'))


if __name__ == '__main__':
    main()
