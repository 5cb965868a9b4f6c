"""Shared test helpers (deterministic data as in the reference's tests)."""
import numpy as np


def lcg_samples(n, dim, state):
    """Deterministic LCG data (src/test.rs:973-982): ((state >> 33) as f32 / u32::MAX as f32) - 0.5."""
    out = np.empty(n * dim, np.float32)
    s = int(state)
    mask = (1 << 64) - 1
    umax = np.float32(4294967295.0)
    for i in range(n * dim):
        s = (s * 6364136223846793005 + 1442695040888963407) & mask
        out[i] = np.float32(np.float32(s >> 33) / umax) - np.float32(0.5)
    return out.reshape(n, dim)


def lcg_cloud(n, dim, state):
    """Arena test cloud (src/tsne/arena.rs:482-492): (state >> 40) as f32 / 2^24."""
    out = np.empty(n * dim, np.float32)
    s = int(state)
    mask = (1 << 64) - 1
    for i in range(n * dim):
        s = (s * 6364136223846793005 + 1442695040888963407) & mask
        out[i] = np.float32(s >> 40) / np.float32(1 << 24)
    return out.reshape(n, dim)


def f32_seq_dist(x):
    """All-pairs Euclidean distances with the reference metric's f32 arithmetic (sequential
    accumulation over dimensions, no FMA, then sqrt), vectorised over pairs."""
    x = np.asarray(x, np.float32)
    n, d = x.shape
    s = np.zeros((n, n), np.float32)
    for t in range(d):
        diff = (x[:, None, t] - x[None, :, t]).astype(np.float32)
        s = (s + (diff * diff).astype(np.float32)).astype(np.float32)
    return np.sqrt(s).astype(np.float32)


def two_clusters(n_per, dim, seed_state=42, sep=30.0):
    """Two separable clusters (src/test.rs:1055-1080 shape)."""
    base = lcg_samples(2 * n_per, dim, seed_state) * np.float32(6.0)
    base[n_per:] += np.float32(sep)
    return base


def same_cluster_fraction(emb, n_per):
    n = emb.shape[0]
    d = ((emb[:, None, :] - emb[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d, np.inf)
    nn = d.argmin(1)
    lab = np.arange(n) < n_per
    return float((lab == lab[nn]).mean())
