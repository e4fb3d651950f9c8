"""Photon-history comparison: seqhis chi2 between two samples and per-photon divergence
and end-point error between two runs on the same random streams.

history_chi2 follows the reference's A-B history comparison (ana/qcf.py:59-170 with
chi2 from ana/nbase.py:339-380): count each distinct seqhis in A and in B, and over
the histories whose combined count exceeds a cut (C2CUT, default 30) sum
(a-b)^2/(a+b); chi2/ndf is that sum over the number of histories kept.
"""
from __future__ import annotations

import numpy as np


def seqhis(seq):
    """seqhis as (n, 2) uint64 from an sseq array (n, 2, 2) / (n, 4) u64 (sysrap/sseq.h:54-65)."""
    return np.asarray(seq, dtype=np.uint64).reshape(len(seq), -1)[:, :2]


def seqbnd(seq):
    return np.asarray(seq, dtype=np.uint64).reshape(len(seq), -1)[:, 2:4]


def num_points(his):
    """Number of recorded points (non-zero nibbles) of each seqhis (n, 2) u64, at most 32."""
    his = np.asarray(his, dtype=np.uint64)
    n = np.zeros(len(his), dtype=np.int32)
    for k in range(32):
        w, sh = divmod(k, 16)
        n += ((his[:, w] >> np.uint64(4 * sh)) & np.uint64(15)) != 0
    return n


def first_divergence(ha, hb):
    """Slot of the first differing seqhis nibble per photon, -1 where the histories agree."""
    ha = np.asarray(ha, dtype=np.uint64)
    hb = np.asarray(hb, dtype=np.uint64)
    out = np.full(len(ha), -1, dtype=np.int32)
    for k in range(32):
        w, sh = divmod(k, 16)
        na = (ha[:, w] >> np.uint64(4 * sh)) & np.uint64(15)
        nb = (hb[:, w] >> np.uint64(4 * sh)) & np.uint64(15)
        m = (out < 0) & (na != nb)
        out[m] = k
    return out


def history_chi2(ha, hb, cut=30):
    """chi2 between the seqhis populations of two samples of equal size.

    Returns dict(c2sum, c2n, c2c, c2per, histories): c2n histories pass the count cut,
    c2c do not; c2per = c2sum / c2n (0 when nothing passes)."""
    ha = np.ascontiguousarray(np.asarray(ha, dtype=np.uint64).reshape(len(ha), 2))
    hb = np.ascontiguousarray(np.asarray(hb, dtype=np.uint64).reshape(len(hb), 2))
    ka = ha.view([('lo', np.uint64), ('hi', np.uint64)]).ravel()
    kb = hb.view([('lo', np.uint64), ('hi', np.uint64)]).ravel()
    keys, inv = np.unique(np.concatenate([ka, kb]), return_inverse=True)
    a = np.bincount(inv[:len(ka)], minlength=len(keys)).astype(np.float64)
    b = np.bincount(inv[len(ka):], minlength=len(keys)).astype(np.float64)
    msk = a + b > cut
    c2 = np.zeros_like(a)
    c2[msk] = (a[msk] - b[msk]) ** 2 / (a[msk] + b[msk])
    c2n = int(msk.sum())
    c2sum = float(c2.sum())
    return dict(c2sum=c2sum, c2n=c2n, c2c=int((~msk).sum()), c2per=c2sum / c2n if c2n else 0.0,
                histories=int(len(keys)))


def endpoint_errors(photon_a, seq_a, photon_b, seq_b):
    """Same-stream comparison of run A against run B (B the more precise).

    A photon's history is fully recorded when it has fewer than 32 points. Over those,
    'divergent' is the fraction whose seqhis or seqbnd differ; for the rest the end
    position |dA-dB| (mm) and time |tA-tB| (ns) errors are summarised."""
    pa = np.asarray(photon_a, dtype=np.float32).reshape(-1, 4, 4)
    pb = np.asarray(photon_b, dtype=np.float32).reshape(-1, 4, 4)
    sa = np.asarray(seq_a).reshape(len(pa), -1)
    sb = np.asarray(seq_b).reshape(len(pb), -1)
    full = num_points(sb[:, :2]) < 32
    same = np.all(sa == sb, axis=1) & full
    nfull = int(full.sum())
    d = np.linalg.norm(pa[same, 0, :3].astype(np.float64) - pb[same, 0, :3].astype(np.float64), axis=1)
    dt = np.abs(pa[same, 0, 3].astype(np.float64) - pb[same, 0, 3].astype(np.float64))
    fd = first_divergence(sa[full, :2], sb[full, :2])
    fd = fd[fd >= 0]

    def pct(x, q):
        return float(np.percentile(x, q)) if len(x) else 0.0

    return dict(photons=int(len(pa)), fully_recorded=nfull,
                divergent_fraction=float(1.0 - same.sum() / nfull) if nfull else 0.0,
                first_divergence_median=float(np.median(fd)) if len(fd) else None,
                pos_mm_p50=pct(d, 50), pos_mm_p99=pct(d, 99), pos_mm_max=float(d.max()) if len(d) else 0.0,
                time_ns_p50=pct(dt, 50), time_ns_p99=pct(dt, 99), time_ns_max=float(dt.max()) if len(dt) else 0.0)
