"""Dense brute-force EMC iteration for tiny inputs -- test infrastructure only.

An independent formulation of one reference iteration (dragonfly/src/maximize.c:8-56 and
the functions it calls) written as dense numpy over all (frame, orientation, pixel)
triples, with no keep lists, no linear searches and no C code. It pins the oracle's
orc_maximize: both are derived from the reference's semantics separately, so a dropped term,
a wrong mask threshold or a swapped operand in either shows up as a mismatch.
"""
import numpy as np

DBL_MIN = np.finfo(np.float64).tiny


def rotation(q):
    """make_rot_quat (model.c:759-786)."""
    q0, q1, q2, q3 = q[:4]
    return np.array([
        [1. - 2. * (q2 * q2 + q3 * q3), 2. * (q1 * q2 + q0 * q3), 2. * (q1 * q3 - q0 * q2)],
        [2. * (q1 * q2 - q0 * q3), 1. - 2. * (q1 * q1 + q3 * q3), 2. * (q0 * q1 + q2 * q3)],
        [2. * (q0 * q2 + q1 * q3), 2. * (q2 * q3 - q0 * q1), 1. - 2. * (q1 * q1 + q2 * q2)]])


def _cells(q, qvals, size):
    p = qvals @ rotation(q).T + size // 2
    cell = np.trunc(p).astype(np.int64)
    inb = np.all((cell >= 0) & (cell <= size - 2), axis=1)
    return p, cell, inb


def _corners(cell, f, size):
    for bx in (0, 1):
        for by in (0, 1):
            for bz in (0, 1):
                w = ((f[:, 0] if bx else 1 - f[:, 0]) * (f[:, 1] if by else 1 - f[:, 1])
                     * (f[:, 2] if bz else 1 - f[:, 2]))
                idx = (((cell[:, 0] + bx) % size) * size + (cell[:, 1] + by) % size) * size \
                    + (cell[:, 2] + bz) % size
                yield w, idx


def slice_view(model, q, qvals, corr):
    """slice_gen3d (model.c:11-59): trilinear view times corr, DBL_MIN outside the grid."""
    size = model.shape[0]
    p, cell, inb = _cells(q, qvals, size)
    f = p - cell
    flat = model.ravel()
    val = np.zeros(len(corr))
    for w, idx in _corners(np.where(inb[:, None], cell, 0), f, size):
        val += w * flat[idx]
    return np.where(inb, val * corr, DBL_MIN)


def merge_view(model2, weight, q, view, qvals, corr, mask):
    """slice_merge3d (model.c:171-244) with np.add.at."""
    size = model2.shape[0]
    p, cell, inb = _cells(q, qvals, size)
    sel = inb & (mask <= 1)
    f = (p - cell)[sel]
    w_val = view[sel] / corr[sel]
    for w, idx in _corners(cell[sel], f, size):
        np.add.at(weight.ravel(), idx, w)
        np.add.at(model2.ravel(), idx, w * w_val)


def dense_counts(frames, num_pix):
    if 'int_frames' in frames:
        return np.asarray(frames['int_frames'], dtype=np.int64)
    D = len(frames['ones'])
    K = np.zeros((D, num_pix), dtype=np.int64)
    for d in range(D):
        o0, m0 = frames['ones_accum'][d], frames['multi_accum'][d]
        K[d, frames['place_ones'][o0:o0 + frames['ones'][d]]] = 1
        sl = slice(m0, m0 + frames['multi'][d])
        K[d, frames['place_multi'][sl]] = frames['count_multi'][sl]
    return K


def iterate(det, quats, model1, frames, fparams, flags, refine=None):
    """One iteration; returns the same fields as oracle.core.maximize. refine (optional):
    dict quat_mapping, rel_off, rel_quat -- frame d only takes orientations whose coarse cell
    is in its set (legacy src/max_emc.c:264-274). model1 (M, S, S, S) runs M modes over the
    combined orientations r = rotation * M + mode (max_prob.c:80-84) and adds quat_norm."""
    qvals, corr, mask = det['qvals'], det['corr'], det['mask']
    nm = model1.shape[0] if model1.ndim == 4 else 1
    vols = model1 if model1.ndim == 4 else model1[None]
    quats = np.repeat(quats, nm, axis=0)  # rotation of each combined orientation
    mode_of = np.tile(np.arange(nm), quats.shape[0] // nm)
    Rn, P = quats.shape[0], len(corr)
    size = vols.shape[-1]
    beta, scale = fparams['beta'], fparams['scale']
    black = np.asarray(fparams['blacklist']).astype(bool)
    sum_fact = fparams['sum_fact']
    need_scaling = flags.get('need_scaling', 0)
    update_scale = flags.get('update_scale', 1)
    use_scale = need_scaling and (flags.get('iteration', 1) > 1 or flags.get('known_scale', 0))
    K = dense_counts(frames, P)
    D = K.shape[0]
    dense = 'int_frames' in frames
    W = np.array([slice_view(vols[m], q, qvals, corr) for q, m in zip(quats, mode_of)])
    good = mask < 1
    u = W[:, good].sum(1)
    w_r = quats[:, 4]
    rescale = fparams['mean_count'] / (w_r * u).sum() * nm
    with np.errstate(divide='ignore'):
        logW = np.log(W)
    R = np.full((D, Rn), -np.inf)
    for d in range(D):
        if black[d]:
            continue
        phi = scale[d] if use_scale else rescale
        # sparse frames sum their photons; dense frames every mask < 1 pixel, where a zero
        # count times log W = -inf is NaN (max_prob.c:173-177)
        t = np.nonzero(good)[0] if dense else np.nonzero((K[d] > 0) & good)[0]
        with np.errstate(invalid='ignore'):
            R[d] = np.log(w_r) - u * phi + (K[d, t][None, :] * logW[:, t]).sum(1)
        if refine is not None:
            cells = refine['rel_quat'][refine['rel_off'][d]:refine['rel_off'][d + 1]]
            R[d][~np.isin(refine['quat_mapping'], cells)] = -np.inf
    M = np.full(D, -1.e100)
    rmax = np.full(D, -1, dtype=np.int64)
    Pm = np.zeros((D, Rn))
    like, info, psum_d, p_norm = np.zeros(D), np.zeros(D), np.zeros(D), np.zeros(D)
    for d in range(D):
        if black[d]:
            continue
        valid = R[d] > -1.e100
        if not valid.any():
            continue
        M[d] = R[d][valid].max()
        rmax[d] = int(np.nonzero(R[d] == M[d])[0][0])
        kept = R[d] > M[d] - 14. / beta[d]
        e = np.where(kept, np.exp(beta[d] * (R[d] - M[d])), 0.)
        p_norm[d] = e.sum()
        Pm[d] = e / p_norm[d]
        like[d] = (Pm[d] * np.where(kept, R[d] - sum_fact[d], 0.)).sum()
        big = kept & (Pm[d] >= 1e-6)
        info[d] = (Pm[d][big] * np.log(Pm[d][big] / w_r[big] * nm)).sum()
        psum_d[d] = (Pm[d] * u * rescale).sum()
    phi_p = np.where(black, 0., scale if need_scaling else 1.)
    psum_r = (Pm * phi_p[:, None]).sum(0)
    Psig = np.where(Pm > 1e-6, Pm, 0.)
    Psig[black] = 0.
    K2 = np.where((mask < 2)[None, :], K, 0)
    views = Psig.T @ K2
    model2 = np.zeros((nm,) + (size,) * 3)
    weight = np.zeros((nm,) + (size,) * 3)
    for r in range(Rn):
        if psum_r[r] > 0:
            merge_view(model2[mode_of[r]], weight[mode_of[r]], quats[r], views[r] / psum_r[r], qvals, corr,
                       mask)
    if model1.ndim == 3:
        model2, weight = model2[0], weight[0]
    # combine_information divides by tot_num_data - iter->num_blacklist (max_prob.c:373-374);
    # the Cython path never sets that field (iterate.pyx:35 callocs it), so callers pass 0
    nb = D - int(fparams.get('num_blacklist', 0))
    out = dict(u=u, rescale=rescale, max_exp=M, rmax=rmax, p_norm=p_norm, likelihood=like,
               info=info, psum_d=psum_d, psum_r=psum_r, prob=Pm, model2=model2,
               inter_weight=weight, R=R, avg_likelihood=like[~black].sum() / nb,
               mutual_info=info[~black].sum() / nb)
    if nm > 1:
        out['quat_norm'] = np.stack([np.where((mode_of == m)[None, :] & (Pm >= 1e-6), Pm, 0.).sum(1)
                                     for m in range(nm)], axis=1)
    if need_scaling and update_scale:
        new_scale = scale.copy()
        new_scale[~black] = fparams['fcounts'][~black] / psum_d[~black]
        out['scale'] = new_scale
    return out


def update_model(model1, model2, weight, alpha=0.):
    """EMCRecon.update_model in Cython order (recon.pyx:112-149): Friedel average of the
    weighted sums, then divide by the weight, rms change, alpha blend. The Friedel partner of
    voxel i is 2c - i with c = size // 2 on every axis; for even sizes index 0 has no partner
    and is left alone (model.c:470-490 starts its loops at 1 then), which is the same
    vectorised flip applied to the [1:, 1:, 1:] block."""
    size = model1.shape[0]
    lo = 1 if size % 2 == 0 else 0
    m2 = np.array(model2, dtype='f8', copy=True)
    a = m2[lo:, lo:, lo:]
    w = weight[lo:, lo:, lo:]
    num = 0.5 * (a * w + a[::-1, ::-1, ::-1] * w[::-1, ::-1, ::-1])
    den = 0.5 * (w + w[::-1, ::-1, ::-1])
    m2[lo:, lo:, lo:] = np.where(den > 0, num / np.where(den > 0, den, 1.), 0.)
    m2 = np.where(weight > 0, m2 / np.where(weight > 0, weight, 1.), m2)
    rms = np.sqrt(((m2 - model1) ** 2).mean())
    new1 = alpha * model1 + (1 - alpha) * m2 if alpha > 0 else m2
    return new1, m2, rms
