"""numpy restatement of the warm-start network (laph_eigsys_amd/csrc/net.hip)
in FP64: forward pass, loss and a hand-derived backward pass.

TEST INFRASTRUCTURE ONLY.  The backward pass is pinned by central finite
differences (tests/test_gcn_oracle.py); the GPU kernels are then checked
against it.

Layout: vectors (n, K) with row 3*site + colour and column j = 16*g + c
(group g, channel c).  Parameters: layers*7*16*16 complex weights stored as
(re, im) float pairs, then layers*16 gate biases.  Hop t: 0 self,
1+2mu = U_mu(x) psi(x+mu), 2+2mu = U_mu(x-mu)^H psi(x-mu).
"""
import numpy as np

from oracle.lapheig_np import shift

NC, NT = 16, 7


def unpack(params, layers):
    p = np.asarray(params, dtype=np.float64)
    nw = layers * NT * NC * NC * 2
    W = p[:nw].reshape(layers, NT, NC, NC, 2)
    W = W[..., 0] + 1j * W[..., 1]
    b = p[nw:nw + layers * NC].reshape(layers, NC)
    return W, b


def transports(U, a):
    """U (L,L,L,3,3,3); a (L,L,L,3,G,NC) -> (NT, L,L,L,3,G,NC)."""
    T = [a]
    for mu in range(3):
        Um = U[..., mu, :, :]
        T.append(np.einsum("...ab,...bgc->...agc", Um, shift(a, mu, 1)))
        T.append(np.einsum("...ba,...bgc->...agc", np.conj(shift(Um, mu, -1)), shift(a, mu, -1)))
    return np.stack(T)


def transports_adj(U, g):
    """Adjoint of transports: sum_t T_t^H g_t for g (NT, ...)."""
    out = g[0].copy()
    for mu in range(3):
        Um = U[..., mu, :, :]
        # (T_{1+2mu})^H h(x) = U_mu(x-mu)^H h(x-mu)
        out += np.einsum("...ba,...bgc->...agc", np.conj(shift(Um, mu, -1)), shift(g[1 + 2 * mu], mu, -1))
        # (T_{2+2mu})^H h(x) = U_mu(x) h(x+mu)
        out += np.einsum("...ab,...bgc->...agc", Um, shift(g[2 + 2 * mu], mu, 1))
    return out


def to_field(v, L):
    n, K = v.shape
    return v.reshape(L, L, L, 3, K // NC, NC)


def forward(params, layers, U, x):
    """Returns (y (n,K), cache)."""
    L = U.shape[0]
    W, b = unpack(params, layers)
    a = to_field(np.asarray(x, dtype=np.complex128), L)
    cache = []
    for l in range(layers):
        T = transports(U, a)
        z = np.einsum("tcd,t...gd->...gc", W[l], T)
        if l < layers - 1:
            r = np.sqrt(np.sum(np.abs(z) ** 2, axis=3, keepdims=True))
            rb = r + b[l][None, None, None, None, None, :]
            with np.errstate(divide="ignore", invalid="ignore"):
                sc = np.where((r > 0) & (rb > 0), rb / np.where(r > 0, r, 1), 0.0)
            cache.append((a, T, z, r, sc))
            a = z * sc
        else:
            cache.append((a, T, z, None, None))
            a = z
    return a.reshape(-1, a.shape[-2] * NC), cache


def loss_fn(y, V):
    """loss = -mean_j |V^H y_j|^2 / |y_j|^2 and its conjugate gradient."""
    K = y.shape[1]
    S = V.conj().T @ y
    ny = np.sum(np.abs(y) ** 2, axis=0)
    f = np.sum(np.abs(S) ** 2, axis=0) / ny
    g = -(V @ S - f * y) / ny / K
    return -np.mean(f), g


def loss_grad(params, layers, U, x, V):
    """(loss, y, grad) with grad the real gradient in the params layout."""
    L = U.shape[0]
    W, b = unpack(params, layers)
    y, cache = forward(params, layers, U, x)
    loss, gy = loss_fn(y, V)
    g = to_field(gy, L)
    gW = np.zeros_like(W)
    gb = np.zeros_like(b)
    for l in range(layers - 1, -1, -1):
        a, T, z, r, sc = cache[l]
        if l < layers - 1:
            # a_out = z * (1 + b/r) where active
            dot = np.real(np.sum(np.conj(g) * z, axis=3, keepdims=True))
            act = sc > 0
            with np.errstate(divide="ignore", invalid="ignore"):
                rr = np.where(r > 0, r, 1)
                s1 = np.where(act, 1 + b[l] / rr, 0.0)
                s2 = np.where(act, -b[l] / rr ** 3 * dot, 0.0)
                gb[l] = np.sum(np.where(act, 2 * dot / rr, 0.0), axis=(0, 1, 2, 3, 4))
            g = s1 * g + s2 * z
        gW[l] = np.einsum("xyzkgc,txyzkgd->tcd", g, np.conj(T))
        gT = np.einsum("tcd,...gc->t...gd", np.conj(W[l]), g)
        g = transports_adj(U, gT)
    grad = np.concatenate([np.stack([2 * gW.real, 2 * gW.imag], axis=-1).reshape(-1), gb.reshape(-1)])
    return loss, y, grad


def init_params(layers, seed):
    """Same initial network as laph_net_create: (1/2)(1 + H/6) per channel plus a small mixing.
    (Values differ from the library's generator; only the structure matters here.)"""
    rng = np.random.default_rng(seed)
    W = np.zeros((layers, NT, NC, NC), dtype=np.complex128)
    for t in range(NT):
        W[:, t] += np.eye(NC) * (0.5 if t == 0 else 1.0 / 12.0)
    W += 0.02 * (rng.random(W.shape) - 0.5) + 0.02j * (rng.random(W.shape) - 0.5)
    b = 0.1 * (rng.random((layers, NC)) - 0.5)
    return np.concatenate([np.stack([W.real, W.imag], axis=-1).reshape(-1), b.reshape(-1)])
