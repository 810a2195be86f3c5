"""Seeded synthetic EMC problems of the benchmark configurations.

Stands in for the reference's simulation pipeline (dragonfly/utils/sim_setup.py ->
make_densities / make_intensities / make_detector / make_data), which needs PDB files and
GSL. Semantics kept from the reference:

* detector q-vectors and polarisation/solid-angle correction as Detector.calc_from_coords
  (dragonfly/detector.pyx:321-356), corr normalised to mean 1 (detector.pyx:55-57); mask 1
  outside the inscribed circle and 2 inside the beam stop (dragonfly/utils/make_detector.py:60-67);
* intensity = |FFT(density)|^2 centred at size//2 (dragonfly/utils/make_intensities.py:68-72);
* uniform random orientations by rejection in the 4-ball (make_data.c:24-40), intensities
  rescaled so the mean photon count over random views is the target (make_data.c:42-113),
  Poisson counts on pixels with mask <= 1 (make_data.c:183-189), sparse ones/multi split.

Seeds (numpy default_rng / torch generator): model 0, orientations 1, Poisson 2, scale 3,
start model 4. Frame slicing and Poisson sampling run in torch on the requested device.
"""
import math

import numpy as np

DBL_MIN = np.finfo(np.float64).tiny

# Benchmark configurations (sizes as stated for each workload).
CONFIGS = {
    'cfg1': dict(size=125, atoms=40, atom_radius=12., sigma=1.5, nx=151, ny=151, gap=0,
                 stoprad=5., num_data=3000, mean_count=100., phi=None, num_div=4,
                 need_scaling=0, beta=1.0, iters=10),
    'cfg2': dict(size=251, atoms=400, atom_radius=30., sigma=1.5, nx=260, ny=257, gap=10,
                 stoprad=5., num_data=100000, mean_count=200., phi=(0.5, 1.5), num_div=10,
                 need_scaling=1, beta='auto', iters=5),
    # dense integer frames (ftype DENSE_INT) of ~5000 photons for the mixed-precision path
    'cfg3': dict(size=251, atoms=400, atom_radius=30., sigma=1.5, nx=260, ny=257, gap=10,
                 stoprad=5., num_data=20000, mean_count=5000., phi=None, num_div=10,
                 need_scaling=0, beta='auto', iters=5, dense=True, alpha=0.5),
    # uneven batching: 1M frames with lognormal photon counts (median 150, clipped to
    # 20-20000), per-frame orientation subsets of 50-5000 of the num_div 12 orientations around
    # a prior (whole Voronoi cells of num_div 3, about 63 orientations each)
    'cfg4': dict(size=251, atoms=400, atom_radius=30., sigma=1.5, nx=260, ny=257, gap=10,
                 stoprad=5., num_data=1000000, mean_count=150.,
                 phi=dict(dist='lognormal', sigma=1.0, range=(20., 20000.)), num_div=12,
                 prior_div=3, subset_range=(50, 5000), prior_error_deg=3., need_scaling=1,
                 beta='auto', iters=5, batched=True),
    # learned orientation proposal: the 100k-frame data set, coarse num_div 4, fine num_div 16;
    # beta_factor 4 is the reference schedule's beta after 40 iterations (jump sqrt(2) every
    # 10), the late iterations in which proposals are used
    'cfg5': dict(size=251, atoms=400, atom_radius=30., sigma=1.5, nx=260, ny=257, gap=10,
                 stoprad=5., num_data=100000, mean_count=200., phi=(0.5, 1.5), num_div=16,
                 coarse_div=4, top_k=16, mass_threshold=0.99, need_scaling=1, beta='auto',
                 beta_factor=4., iters=5, learned=True),
}


def sphere_atoms(num_atoms, radius, rng):
    pts = []
    while len(pts) < num_atoms:
        p = rng.uniform(-radius, radius, 3)
        if p @ p <= radius * radius:
            pts.append(p)
    return np.array(pts)


def gaussian_density(size, positions, sigma):
    """Sum of isotropic Gaussians (width sigma voxels) centred at size//2 + position."""
    dens = np.zeros((size,) * 3)
    c = size // 2
    h = int(math.ceil(5 * sigma))
    ax = np.arange(-h, h + 1)
    for p in positions:
        ctr = c + p
        i0 = np.floor(ctr).astype(int)
        g = [np.exp(-((i0[k] + ax) - ctr[k]) ** 2 / (2 * sigma * sigma)) for k in range(3)]
        sl = [slice(i0[k] - h, i0[k] + h + 1) for k in range(3)]
        dens[sl[0], sl[1], sl[2]] += g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]
    return dens


def intensity_from_density(dens):
    import scipy.fft as sfft
    f = sfft.fftn(dens, workers=-1)
    return np.abs(np.fft.fftshift(f)) ** 2


def make_intensity(size, atoms, atom_radius, sigma, seed=0):
    rng = np.random.default_rng(seed)
    return intensity_from_density(gaussian_density(size, sphere_atoms(atoms, atom_radius, rng), sigma))


def flat_detector(nx, ny, qmax, detd=None, stoprad=5., gap=0, pol='x'):
    """Flat detector of nx x ny pixels whose largest |q| (at a corner) is qmax voxels.

    Returns qvals (N,3), corr (N,) normalised to mean 1, raw_mask (N,) u1, detd, ewald_rad.
    A horizontal gap of `gap` rows through the centre splits it into two panels (mask 2).
    """
    detc_x, detc_y = (nx - 1) / 2., (ny - 1) / 2.
    x, y = np.meshgrid(np.arange(nx) - detc_x, np.arange(ny) - detc_y, indexing='ij')
    cx, cy = x.ravel(), y.ravel()
    if detd is None:
        detd = 2. * max(detc_x, detc_y)
    fac = np.sqrt(cx ** 2 + cy ** 2 + detd ** 2)
    # |q|^2 = 2 ER^2 (1 - detd/fac): pick ER so the corner reaches qmax
    ewald_rad = qmax / np.sqrt(2. * (1. - detd / fac.max()))
    qvals = np.empty((cx.size, 3))
    qvals[:, 0] = cx * ewald_rad / fac
    qvals[:, 1] = cy * ewald_rad / fac
    qvals[:, 2] = ewald_rad * (detd / fac - 1.)
    if pol == 'x':
        corr = detd / fac ** 3 * (1. - cx ** 2 / fac ** 2)
    elif pol == 'y':
        corr = detd / fac ** 3 * (1. - cy ** 2 / fac ** 2)
    else:
        corr = detd / fac ** 3 * (1. - (cx ** 2 + cy ** 2) / 2 / fac ** 2)
    corr = corr / corr.mean()
    rad = np.sqrt(cx ** 2 + cy ** 2)
    mask = np.zeros(cx.size, 'u1')
    mask[rad > min(detc_x, detc_y)] = 1
    mask[rad < stoprad] = 2
    if gap > 0:
        g0 = nx // 2 - gap // 2
        rows = np.repeat(np.arange(nx), ny)
        mask[(rows >= g0) & (rows < g0 + gap)] = 2
    return dict(qvals=qvals, corr=corr, mask=mask, detd=detd, ewald_rad=ewald_rad)


def model_size(qvals):
    """Iterate.calculate_size of the largest |q| over all pixels (iterate.pyx:459-468,
    detector.pyx:106-117)."""
    qmax = np.sqrt((qvals ** 2).sum(1).max())
    return int(2 * np.ceil(qmax) + 3)


def random_quaternions(n, rng):
    """Uniform unit quaternions by rejection in the 4-ball (make_data.c:24-40)."""
    out = np.empty((0, 4))
    while out.shape[0] < n:
        q = rng.random((2 * (n - out.shape[0]) + 16, 4)) - .5
        qq = (q * q).sum(1)
        q = q[qq <= .25]
        out = np.concatenate([out, q / np.sqrt((q * q).sum(1))[:, None]])
    return out[:n]


def rot_matrices(q):
    """make_rot_quat convention (dragonfly/src/model.c:759-786), vectorised."""
    q0, q1, q2, q3 = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    r = np.empty((q.shape[0], 3, 3))
    r[:, 0, 0] = 1. - 2. * (q2 * q2 + q3 * q3)
    r[:, 0, 1] = 2. * (q1 * q2 + q0 * q3)
    r[:, 0, 2] = 2. * (q1 * q3 - q0 * q2)
    r[:, 1, 0] = 2. * (q1 * q2 - q0 * q3)
    r[:, 1, 1] = 1. - 2. * (q1 * q1 + q3 * q3)
    r[:, 1, 2] = 2. * (q0 * q1 + q2 * q3)
    r[:, 2, 0] = 2. * (q0 * q2 + q1 * q3)
    r[:, 2, 1] = 2. * (q2 * q3 - q0 * q1)
    r[:, 2, 2] = 1. - 2. * (q1 * q1 + q2 * q2)
    return r


def _slices(torch, model_flat, size, rots, qv, corr):
    """Trilinear views (B, N) of the model at rotations rots (B,3,3), slice_gen3d semantics."""
    center = float(size // 2)
    p = [rots[:, i, 0:1] * qv[None, :, 0] + rots[:, i, 1:2] * qv[None, :, 1]
         + rots[:, i, 2:3] * qv[None, :, 2] + center for i in range(3)]
    cell = [torch.trunc(pi) for pi in p]
    inb = torch.ones_like(p[0], dtype=torch.bool)
    for ci in cell:
        inb &= (ci >= 0) & (ci <= size - 2)
    f = [p[i] - cell[i] for i in range(3)]
    ix = [torch.where(inb, cell[i], torch.zeros_like(cell[i])).long() for i in range(3)]
    val = torch.zeros_like(p[0])
    for bx in (0, 1):
        wx = f[0] if bx else 1. - f[0]
        xx = (ix[0] + bx) % size
        for by in (0, 1):
            wy = f[1] if by else 1. - f[1]
            yy = (ix[1] + by) % size
            for bz in (0, 1):
                wz = f[2] if bz else 1. - f[2]
                zz = (ix[2] + bz) % size
                val += wx * wy * wz * model_flat[(xx * size + yy) * size + zz]
    val = val * corr[None, :]
    return torch.where(inb, val, torch.full_like(val, DBL_MIN))


def simulate_frames(intens, det, num_data, mean_count, phi=None, device='cpu', batch=1024,
                    seed_quat=1, seed_poisson=2, seed_phi=3, num_ave=5000):
    """Sparse Poisson frames of `intens` at uniform random orientations.

    Returns a dict with the reference's sparse layout (ones, multi, place_ones,
    place_multi, count_multi, ones_accum, multi_accum) plus the true quaternions and scales.
    """
    import torch
    size = intens.shape[0]
    dev = torch.device(device)
    model_flat = torch.as_tensor(np.ascontiguousarray(intens).ravel(), device=dev)
    qv = torch.as_tensor(det['qvals'], device=dev)
    corr = torch.as_tensor(det['corr'], device=dev)
    keep = torch.as_tensor(det['mask'] <= 1, device=dev)
    rng = np.random.default_rng(seed_quat)
    q_ave = random_quaternions(num_ave, rng)
    quats = random_quaternions(num_data, rng)
    # rescale_intens: mean photon count over random views (make_data.c:42-113)
    tot = 0.
    for b0 in range(0, num_ave, batch):
        rots = torch.as_tensor(rot_matrices(q_ave[b0:b0 + batch]), device=dev)
        v = _slices(torch, model_flat, size, rots, qv, corr)
        tot += float((v * keep[None, :]).sum())
    rescale = mean_count / (tot / num_ave)
    if isinstance(phi, dict) and phi.get('dist') == 'lognormal':
        # per-frame photon count ~ lognormal with the given median, clipped to [lo, hi]
        lo, hi = phi['range']
        counts = np.random.default_rng(seed_phi).lognormal(np.log(mean_count), phi['sigma'], num_data)
        scales = np.clip(counts, lo, hi) / mean_count
    elif phi is not None:
        scales = np.random.default_rng(seed_phi).uniform(phi[0], phi[1], num_data)
    else:
        scales = np.ones(num_data)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed_poisson)
    ones, multi, po, pm, cm = [], [], [], [], []
    for b0 in range(0, num_data, batch):
        b1 = min(num_data, b0 + batch)
        rots = torch.as_tensor(rot_matrices(quats[b0:b1]), device=dev)
        lam = _slices(torch, model_flat, size, rots, qv, corr) * rescale
        lam = lam * torch.as_tensor(scales[b0:b1], device=dev)[:, None]
        lam = torch.where(keep[None, :], lam, torch.zeros_like(lam))
        counts = torch.poisson(lam, generator=gen).to(torch.int32)
        nz = counts.nonzero()
        vals = counts[nz[:, 0], nz[:, 1]]
        is1 = vals == 1
        ism = vals > 1
        B = b1 - b0
        ones.append(torch.bincount(nz[is1, 0], minlength=B).cpu().numpy())
        multi.append(torch.bincount(nz[ism, 0], minlength=B).cpu().numpy())
        po.append(nz[is1, 1].to(torch.int32).cpu().numpy())
        pm.append(nz[ism, 1].to(torch.int32).cpu().numpy())
        cm.append(vals[ism].cpu().numpy())
        del lam, counts, nz, vals
    ones = np.concatenate(ones).astype('i4')
    multi = np.concatenate(multi).astype('i4')
    out = dict(ones=ones, multi=multi, place_ones=np.concatenate(po).astype('i4'),
               place_multi=np.concatenate(pm).astype('i4'),
               count_multi=np.concatenate(cm).astype('i4'), true_quats=quats, true_scale=scales,
               rescale=rescale)
    out['ones_accum'] = np.concatenate([[0], np.cumsum(ones)[:-1]]).astype('i8')
    out['multi_accum'] = np.concatenate([[0], np.cumsum(multi)[:-1]]).astype('i8')
    return out


def dense_frames(frames, num_pix):
    """(num_data, num_pix) int32 counts from the sparse lists of simulate_frames."""
    D = len(frames['ones'])
    K = np.zeros((D, num_pix), dtype='i4')
    rows = np.repeat(np.arange(D), frames['ones'])
    K[rows, frames['place_ones']] = 1
    rows = np.repeat(np.arange(D), frames['multi'])
    K[rows, frames['place_multi']] = frames['count_multi']
    return K


def start_model(size, mean_count, mask, seed=4):
    """Random start model with the mean Iterate.from_config uses (iterate.pyx:165-166)."""
    model_mean = mean_count / (np.asarray(mask) == 0).sum() * 2.
    return np.random.default_rng(seed).random((size,) * 3) * model_mean


def make_problem(name='cfg1', device='cpu', num_data=None, num_div=None, batch=1024, **over):
    """Build the detector, true intensities, frames and start model of a configuration.

    `num_data` / `num_div` / other keys override the configuration (for reduced tests).
    """
    cfg = dict(CONFIGS[name]) if isinstance(name, str) else dict(name)
    cfg.update(over)
    if num_data is not None:
        cfg['num_data'] = num_data
    if num_div is not None:
        cfg['num_div'] = num_div
    size = cfg['size']
    det = flat_detector(cfg['nx'], cfg['ny'], qmax=(size - 3) / 2. - 0.05, stoprad=cfg['stoprad'],
                        gap=cfg['gap'])
    if model_size(det['qvals']) != size:
        raise RuntimeError('detector does not give model size %d' % size)
    intens = make_intensity(size, cfg['atoms'], cfg['atom_radius'], cfg['sigma'], seed=0)
    frames = simulate_frames(intens, det, cfg['num_data'], cfg['mean_count'], phi=cfg['phi'],
                             device=device, batch=batch)
    return dict(cfg=cfg, det=det, intens=intens, frames=frames)


def prior_subsets(true_quats, qcoarse, num_fine, subset_range, prior_error_deg, seed=5, device='cuda'):
    """Per-frame refinement sets around a prior orientation: the prior is the true orientation
    turned by prior_error_deg about a random axis; a frame takes the k Voronoi cells of
    qcoarse nearest its prior, k chosen so that the set holds a log-uniform target number of
    fine orientations in subset_range (num_fine / len(qcoarse) per cell on average).
    Returns (rel_off int64[N + 1], rel_quat int32[...])."""
    import torch
    rng = np.random.default_rng(seed)
    N = len(true_quats)
    axis = rng.normal(size=(N, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    half = np.radians(prior_error_deg) / 2.
    turn = np.concatenate([np.full((N, 1), np.cos(half)), axis * np.sin(half)], axis=1)
    q = np.asarray(true_quats, 'f8')[:, :4]
    # Hamilton product q * turn
    w1, x1, y1, z1 = q.T
    w2, x2, y2, z2 = turn.T
    prior = np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                      w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], axis=1)
    per_cell = num_fine / len(qcoarse)
    lo, hi = subset_range
    target = np.exp(rng.uniform(np.log(lo), np.log(hi), N))
    k = np.clip(np.rint(target / per_cell), 1, len(qcoarse)).astype('i8')
    kmax = int(k.max())
    qc = torch.as_tensor(np.asarray(qcoarse, 'f8')[:, :4], device=device)
    idx = np.empty((N, kmax), 'i4')
    for b0 in range(0, N, 65536):
        pr = torch.as_tensor(prior[b0:b0 + 65536], device=device)
        top = torch.topk((pr @ qc.T).abs(), kmax, dim=1).indices
        idx[b0:b0 + len(pr)] = top.to(torch.int32).cpu().numpy()
    rel_off = np.zeros(N + 1, 'i8')
    np.cumsum(k, out=rel_off[1:])
    rel_quat = idx[np.arange(kmax)[None, :] < k[:, None]]
    return rel_off, rel_quat.astype('i4')


def take_frames(frames, n):
    """The first n frames of a sparse frame set (same layout)."""
    out = dict(frames)
    no = int(frames['ones'][:n].sum())
    nm = int(frames['multi'][:n].sum())
    for k in ('ones', 'multi', 'ones_accum', 'multi_accum', 'true_quats', 'true_scale'):
        if k in frames:
            out[k] = frames[k][:n]
    out['place_ones'] = frames['place_ones'][:no]
    out['place_multi'] = frames['place_multi'][:nm]
    out['count_multi'] = frames['count_multi'][:nm]
    return out
