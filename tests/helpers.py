"""Dense FP64 restatements used to pin the oracle and the HIP path.

``tsne_reference_probabilities`` / ``tsne_reference_forces`` restate the
reference test suite's independent oracle
(tests/testthat/helper-tsne-reference.R:1-116) with dense numpy arrays;
``exact_knn`` restates tests/testthat/helper-knn.R:1-37 (exclude_self=TRUE).
"""
import numpy as np


def exact_knn(x, k, query=None):
    """Exact Euclidean KNN; returns 0-based (indices, distances), n x k."""
    x = np.asarray(x, np.float64)
    same = query is None
    q = x if same else np.asarray(query, np.float64)
    d2 = (q * q).sum(1)[:, None] + (x * x).sum(1)[None, :] - 2.0 * q @ x.T
    d2[d2 < 0] = 0
    if same:
        np.fill_diagonal(d2, np.inf)
    idx = np.argsort(d2, axis=1, kind="stable")[:, :k]
    dist = np.sqrt(np.take_along_axis(d2, idx, 1))
    return idx.astype(np.int32), dist


def tsne_reference_probabilities(indices, distances, perplexity):
    indices = np.asarray(indices)
    distances = np.asarray(distances, np.float64)
    n, k = indices.shape
    directed = np.zeros((n, n))
    target = np.log(perplexity)
    eps = np.finfo(np.float64).eps
    for i in range(n):
        d2 = distances[i] ** 2
        shifted = d2 - d2.min()
        if shifted.max() <= eps * max(1.0, d2.max()):
            directed[i, indices[i]] = 1.0 / k
            continue
        beta, beta_min, beta_max = 1.0, -np.inf, np.inf
        for _ in range(200):
            un = np.exp(-beta * shifted)
            mass = un.sum()
            entropy = beta * (shifted * un).sum() / mass + np.log(mass)
            diff = entropy - target
            if abs(diff) < 1e-5:
                break
            if diff > 0:
                beta_min = beta
                beta = 2 * beta if np.isinf(beta_max) else 0.5 * (beta + beta_max)
            else:
                beta_max = beta
                beta = 0.5 * beta if np.isinf(beta_min) else 0.5 * (beta + beta_min)
        prob = un / un.sum()
        for j in range(k):
            nb = indices[i, j]
            if nb != i:
                directed[i, nb] += prob[j]
    sym = directed + directed.T
    return sym / sym.sum()


def tsne_reference_forces(probability, layout, exaggeration=1.0):
    y = np.asarray(layout, np.float64)
    diff = y[:, None, :] - y[None, :, :]
    d2 = (diff ** 2).sum(-1)
    num = 1.0 / (1.0 + d2)
    np.fill_diagonal(num, 0.0)
    sum_q = num.sum()
    attractive = (exaggeration * probability * num)[:, :, None] * diff
    repulsive = (-(num ** 2) / sum_q)[:, :, None] * diff
    attractive = attractive.sum(1)
    repulsive = repulsive.sum(1)
    grad = attractive + repulsive
    pos = probability > 0
    nq = num / sum_q
    kl = float((probability[pos] * np.log(probability[pos] / nq[pos])).sum())
    return dict(attractive_force=attractive, repulsive_force=repulsive,
                optimizer_gradient=grad, objective_gradient=4 * grad, sum_q=sum_q, kl=kl)


def relative_l2(observed, reference):
    observed = np.asarray(observed, np.float64)
    reference = np.asarray(reference, np.float64)
    return float(np.sqrt(((observed - reference) ** 2).sum()) /
                 max(np.sqrt((reference ** 2).sum()), np.finfo(float).eps))


def gaussian_mixture(n, d, n_clusters, seed, center_sd=5.0, dirichlet=False):
    """Seeded Gaussian mixture (centres N(0, center_sd^2 I), within N(0, I))."""
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, center_sd, size=(n_clusters, d))
    if dirichlet:
        sizes = rng.multinomial(n, rng.dirichlet(np.ones(n_clusters)))
    else:
        sizes = np.full(n_clusters, n // n_clusters)
        sizes[: n - sizes.sum()] += 1
    labels = np.repeat(np.arange(n_clusters), sizes)
    x = centres[labels] + rng.normal(size=(n, d))
    return x.astype(np.float32), labels


class DeviceArray:
    """A device copy of a numpy array through the HIP runtime the backend
    library links (ctypes; no torch), for device-input tests."""

    def __init__(self, a):
        import ctypes

        from fastembedr_amd import _lib

        _lib.lib()  # the backend's runtime is loaded first
        self._hip = ctypes.CDLL("libamdhip64.so.7")
        a = np.ascontiguousarray(a)
        self.ptr = ctypes.c_void_p()
        assert self._hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(a.nbytes, 1))) == 0
        assert self._hip.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p),
                                   ctypes.c_size_t(a.nbytes), 1) == 0  # host to device
        self.nbytes = a.nbytes

    def to_host(self, dtype, shape):
        import ctypes

        out = np.zeros(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        assert self._hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                   ctypes.c_size_t(out.nbytes), 2) == 0  # device to host
        return out

    def free(self):
        if self.ptr:
            self._hip.hipFree(self.ptr)
            self.ptr = None
