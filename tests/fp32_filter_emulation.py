"""CPU emulation (numpy) of the batched solver's Chebyshev filter followed
by its FP64 Gram, in complex64 and in complex128 arithmetic, to see what
makes the filtered block's Gram approach numerical singularity.

The batched solver filters each Ritz pair (x, theta) of B = c D + a in
correction form: T_m(B) x = T_m(theta) x + e, with e computed in FP32.  A
column whose Ritz value lies in the damped interval keeps its own direction
with weight |T_m(theta)| <= 1, while its components along the wanted
eigenvectors grow by up to A = T_m(theta_max)^P.  Two candidates could push
the equilibrated Gram's smallest eigenvalue towards the level where an FP64
Cholesky of K columns fails (about K * 2^-53): the amplification itself
(columns drawn towards the few most amplified eigenvectors), or the FP32
rounding of e (about 2^-24 |e| in random directions).  Emulated at 8^3, both
arithmetics give the same Gram spectra in every cycle of bounded
amplification; in the uncapped four-pass cycle (amplification about 2e21)
the Gram is singular at working precision with either arithmetic (smallest
eigenvalue 1e-16 with the FP32 filter, -1.6e-15 with the FP64 one, whose
Cholesky fails): the loss of rank comes from the amplification, and FP32
rounding adds nothing measurable.

Two filter schedules are emulated:
  "uncapped": two single-pass warm-up cycles of T_16(B) from a Gaussian
              block, then T_16(B)^4 per cycle with no bound on A (the solver
              before the amplification cap);
  "current":  one warm-up cycle of T_16(B) from the start window (3 x the
              K-th level of the free Laplacian), then one correction pass per
              cycle of the highest degree d (multiple of 4, <= 128) with
              T_d(theta_max) <= 1e10, applied as the product over the roots
              of (T_d(y) - T_d(theta)) / (y - theta) in their Leja order.

Test infrastructure: tests/test_fp32_filter_conditioning.py runs it at a
small size.  Run directly for the table at a larger size:
    python tests/fp32_filter_emulation.py [L] [K] [nev]
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import lapheig_np as N  # noqa: E402

LAM_L = 12.0


def block_D(U, X, dtype):
    """D X for a block X (n, K) in the given complex dtype (links cast too)."""
    L = U.shape[0]
    K = X.shape[1]
    Uc = U.astype(dtype)
    psi = X.reshape(L, L, L, 3, K)
    out = (-6.0 * psi).astype(dtype)
    for mu in range(3):
        Um = Uc[..., mu, :, :]
        out += np.einsum("xyzab,xyzbk->xyzak", Um, N.shift(psi, mu, 1))
        out += np.einsum("xyzba,xyzbk->xyzak", np.conj(N.shift(Um, mu, -1)), N.shift(psi, mu, -1))
    return out.reshape(-1, K)


def free_level(L, K):
    s = 4.0 * np.sin(np.pi * np.arange(L) / L) ** 2
    ev = np.sort((s[:, None, None] + s[None, :, None] + s[None, None, :]).reshape(-1))
    return ev[min((K - 1) // 3, ev.size - 1)]


def cheb(m, th):
    th = np.asarray(th, dtype=np.complex128)
    return np.real(np.cos(m * np.arccos(th)))


def correction_root_order(m):
    z = np.cos(2.0 * np.pi * np.arange(1, m) / m + 0.1j)
    order, used = [], np.zeros(m - 1, bool)
    for n in range(m - 1):
        if n == 0:
            score = np.abs(z)
        else:
            with np.errstate(divide="ignore"):
                score = np.array([np.sum(np.log(np.abs(z[i] - z[np.array(order)]))) for i in range(m - 1)])
        score[used] = -np.inf
        best = int(np.argmax(score))
        used[best] = True
        order.append(best)
    return [k + 1 for k in order]


def filter_q_form(U, X, theta, c, a, m, dtype):
    """e_m / s_m for every column: the product over the m-1 roots of
    q_theta(y) = (T_m(y) - T_m(theta)) / (y - theta) applied to r = (B - theta) X,
    one factor per step with per-column complex shifts, in `dtype`; returns
    (gamma, w) with T_m(B) X / s_m = gamma X + w."""
    r = (c * block_D(U, X, np.complex128) + (a - theta)[None, :] * X)
    tm = cheb(m, theta)
    sm = np.maximum(np.abs(tm), 1.0)
    phi = np.arccos(theta.astype(np.complex128))
    sphi = np.sin(phi)
    h = r.astype(dtype)
    qk = np.ones_like(theta, dtype=np.complex128)
    sig = np.ones_like(theta)
    for st, k in enumerate(correction_root_order(m), start=1):
        x = theta * np.cos(2 * np.pi * k / m) - sphi * np.sin(2 * np.pi * k / m)
        qk = qk * 2.0 * (theta - x)
        sn = sm if st == m - 1 else np.maximum(np.abs(qk), 1.0)
        a1 = (2.0 * sig / sn).astype(np.float64)
        Bh = c * block_D(U, h, dtype) + np.asarray(a, dtype=dtype) * h
        h = (a1.astype(dtype)[None, :] * Bh - (a1 * x).astype(dtype)[None, :] * h).astype(dtype)
        sig = sn
    return tm / sm, h


def filter_three_term(U, X, theta, c, a, m, passes, dtype):
    """Correction pass e_{k+1} = 2 B e_k - e_{k-1} + 2 T_k(theta) r (scaled by
    s_k = max(|T_k(theta)|, 1)), then passes-1 plain passes
    w_p = T_m(B) w_{p-1} / s_m + gamma^(p-1) w_1, in `dtype`; returns
    (gamma^passes, w_passes) with T_m(B)^P X / s_m^P = gamma^P X + w_P."""
    B = lambda v: (c * block_D(U, v, dtype) + np.asarray(a, dtype=dtype) * v).astype(dtype)  # noqa: E731
    r = (c * block_D(U, X, np.complex128) + (a - theta)[None, :] * X).astype(dtype)
    tau = [np.ones_like(theta), theta.copy()]
    for k in range(1, m):
        tau.append(2 * theta * tau[k] - tau[k - 1])
    s = [np.maximum(np.abs(t), 1.0) for t in tau]
    e_old = np.zeros_like(r)
    e = r  # e_1 (s_1 scaling folded below)
    e = (e / s[1].astype(np.float32)[None, :]).astype(dtype)
    for k in range(1, m):
        nxt = (2 * s[k] / s[k + 1]).astype(np.float32)[None, :] * B(e) \
            - (s[k - 1] / s[k + 1]).astype(np.float32)[None, :] * e_old \
            + (2 * tau[k] / s[k + 1]).astype(np.float32)[None, :] * r
        e_old, e = e, nxt.astype(dtype)  # e_k = E_k / s_k, each at its own scale
    gamma = tau[m] / s[m]
    w1 = e
    w = w1
    for p in range(2, passes + 1):
        h_old, h = w, (B(w) / np.float32(1.0)).astype(dtype)
        for k in range(1, m):
            h_old, h = h, (2 * B(h) - h_old).astype(dtype)
        w = (h / s[m].astype(np.float32)[None, :] + (gamma ** (p - 1)).astype(np.float32)[None, :] * w1).astype(dtype)
    return gamma ** passes, w


def rayleigh_ritz(U, Y):
    """Ritz values mu (ascending, of -D) and vectors of the span of Y (FP64,
    CholQR2 then eigh of the projected matrix)."""
    for _ in range(2):
        Y = Y / np.linalg.norm(Y, axis=0)[None, :]
        R = np.linalg.cholesky(Y.conj().T @ Y).conj().T
        Y = np.linalg.solve(R.T, Y.T).T
    M = -(Y.conj().T @ block_D(U, Y, np.complex128))
    mu, Z = np.linalg.eigh(0.5 * (M + M.conj().T))
    return mu, Y @ Z


def gram_health(Y):
    """Smallest eigenvalue of the equilibrated FP64 Gram S Y^H Y S and whether
    LAPACK's Cholesky of it succeeds."""
    G = Y.conj().T @ Y
    d = 1.0 / np.sqrt(np.real(np.diag(G)))
    Ge = d[:, None] * G * d[None, :]
    lam = float(np.linalg.eigvalsh(0.5 * (Ge + Ge.conj().T))[0])
    try:
        np.linalg.cholesky(Ge)
        ok = True
    except np.linalg.LinAlgError:
        ok = False
    return lam, ok


def run(schedule, L=8, K=64, nev=48, cycles=4, seed=7, eps=0.3, dtype=np.complex64):
    """Emulated cycles of one schedule from a Gaussian block: per cycle the
    degree, passes, largest amplification A = T_m(theta_max)^P, and the
    equilibrated Gram's smallest eigenvalue and Cholesky success for the
    filtered block."""
    cfg = N.synth_config(seed, eps, L, 1)
    U = N.hyp(N.links_from_config(cfg, L), 0.62, 0.62, 3)
    rng = np.random.default_rng(seed)
    n = 3 * L ** 3
    X = rng.standard_normal((n, K)) + 1j * rng.standard_normal((n, K))
    rows = []
    if schedule == "uncapped":
        mu, X = rayleigh_ritz(U, X)
        lc = min(mu[-1], 0.75 * LAM_L)
    else:  # the current solver: no start Rayleigh-Ritz, window from the free levels
        X = X / np.linalg.norm(X, axis=0)[None, :]
        lc = min(3.0 * free_level(L, K), 0.75 * LAM_L)
        mu = np.full(K, lc)
    for it in range(cycles):
        c = 2.0 / (LAM_L - lc)
        a = 1.0 + c * lc
        theta = 1.0 + c * (lc - mu)
        th_max = 1.0 + 2.0 * lc / (LAM_L - lc)
        if schedule == "uncapped":
            m, passes = 16, (1 if it < 2 else 4)
            g, w = filter_three_term(U, X, theta, c, a, m, passes, dtype)
        else:
            passes = 1
            if it == 0:
                m = 16
            else:
                m = int(min(np.floor(np.arccosh(1e10) / np.arccosh(th_max)), 128)) // 4 * 4
                m = max(m, 4)
            g, w = filter_q_form(U, X, theta, c, a, m, dtype)
        amp = float(cheb(m, th_max)) ** passes
        Y = g[None, :] * X + w.astype(np.complex128)
        lam, ok = gram_health(Y)
        rows.append({"cycle": it, "degree": m, "passes": passes, "amplification": amp, "gram_min_eig": lam,
                     "cholesky_ok": ok})
        try:
            mu, X = rayleigh_ritz(U, Y)
        except np.linalg.LinAlgError:
            rows[-1]["rr_failed"] = True
            break
        lc = min(mu[-1], 0.75 * LAM_L)
    return rows


def main():
    L = int(sys.argv[1]) if len(sys.argv) > 1 else 10
    K = int(sys.argv[2]) if len(sys.argv) > 2 else 128
    nev = int(sys.argv[3]) if len(sys.argv) > 3 else 96
    for sched in ("uncapped", "current"):
        for dt, name in ((np.complex64, "FP32"), (np.complex128, "FP64")):
            print(f"{sched} schedule, filter in {name}, L={L} K={K}")
            for r in run(sched, L, K, nev, cycles=5, dtype=dt):
                print("  " + "  ".join(f"{k}={v:.3g}" if isinstance(v, float) else f"{k}={v}" for k, v in r.items()),
                      flush=True)


if __name__ == "__main__":
    main()
