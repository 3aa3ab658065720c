"""CPU emulation of the workgroup counting sort at the top of the shadow and
continuation shade rounds (wf_shade_kernel in voxeltoy_amd/csrc/vt_trace.hip):
per wave a ballot per class, the lane's rank among the wave's lanes of its
class, the per-wave class counts in LDS, then the position = the counts of
all lower classes, plus the counts of its class in lower waves, plus the
rank.  The positions must be a permutation of the workgroup's threads that
orders the paths by class (hits, then sky misses, then no path), keeping
wave and lane order inside a class."""
import numpy as np
import pytest

THREADS, WAVE, CLASSES = 256, 64, 3


def positions(cls):
    nw = THREADS // WAVE
    counts = np.zeros((nw, CLASSES), np.int64)
    rank = np.zeros(THREADS, np.int64)
    for w in range(nw):
        lanes = cls[w * WAVE:(w + 1) * WAVE]
        for k in range(CLASSES):
            ballot = lanes == k                         # __ballot(cls == k)
            counts[w, k] = ballot.sum()                 # __popcll(m)
            below = np.cumsum(ballot) - ballot          # __popcll(m & lanes below)
            rank[w * WAVE:(w + 1) * WAVE][ballot] = below[ballot]
    pos = np.empty(THREADS, np.int64)
    for t in range(THREADS):
        w, c = t // WAVE, cls[t]
        p = rank[t]
        for k in range(CLASSES):
            for v in range(nw):
                if k < c or (k == c and v < w):
                    p += counts[v, k]
        pos[t] = p
    return pos


@pytest.mark.parametrize("seed", range(6))
def test_shade_sort_is_a_stable_permutation(seed):
    rng = np.random.default_rng(seed)
    n_active = [THREADS, 1, 77, 200, 255, 0][seed]      # a batch may end inside the workgroup
    hit_fraction = [0.5, 0.0, 1.0, 0.8, 0.18, 0.5][seed]
    cls = np.where(rng.random(THREADS) < hit_fraction, 0, 1)
    cls[n_active:] = 2                                   # threads past the queue's end: no path
    pos = positions(cls)
    assert sorted(pos.tolist()) == list(range(THREADS))
    order = np.empty(THREADS, np.int64)
    order[pos] = np.arange(THREADS)                      # s_src[pos] = threadIdx.x
    assert np.array_equal(order, np.argsort(cls, kind="stable"))
    # the threads that receive a path are exactly the first n_active
    assert (cls[order][:n_active] != 2).all() and (cls[order][n_active:] == 2).all()
