"""Independent numpy restatement of the LapH_EigSys path, used to cross-check
the C oracle (oracle/lapheig_oracle.c) and to build golden fixtures.

TEST INFRASTRUCTURE ONLY (never imported by the product package).

Everything is written with whole-lattice numpy operations (np.roll for the
periodic neighbours, batched 3x3 algebra, numpy.linalg.eigh for the matrix
square root), so it shares no code path with the C oracle's per-site loops.

It also holds the definition of the synthetic gauge field:
  U = exp(i eps H),  H = sum_{a=1..8} xi_a lambda_a / 2,  xi_a ~ N(0,1),
with the Gaussians drawn from a counter-based hash keyed by
(seed, t, x, y, z, mu) -- the same definition the product's generator
implements (laph_synth_config in laph_eigsys_amd/csrc/host_io.cpp), so both can
be compared.
"""
import numpy as np

# ---------------------------------------------------------------------------
# counter-based RNG: splitmix64 finaliser of seed*golden + (ctr+1)*odd
M64 = (1 << 64) - 1
GOLD = 0x9E3779B97F4A7C15
ODD = 0xD1B54A32D192ED03


def _mix(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def hash_u01(seed, ctr):
    """Uniform (0,1) doubles from (seed, counter array)."""
    with np.errstate(over="ignore"):
        h = np.uint64((seed * GOLD) & M64) + (np.asarray(ctr, dtype=np.uint64) + np.uint64(1)) * np.uint64(ODD)
    z = _mix(h)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def gell_mann():
    lam = np.zeros((8, 3, 3), dtype=np.complex128)
    lam[0][0, 1] = lam[0][1, 0] = 1
    lam[1][0, 1] = -1j
    lam[1][1, 0] = 1j
    lam[2][0, 0] = 1
    lam[2][1, 1] = -1
    lam[3][0, 2] = lam[3][2, 0] = 1
    lam[4][0, 2] = -1j
    lam[4][2, 0] = 1j
    lam[5][1, 2] = lam[5][2, 1] = 1
    lam[6][1, 2] = -1j
    lam[6][2, 1] = 1j
    lam[7] = np.diag([1, 1, -2]) / np.sqrt(3.0)
    return lam


def synth_config(seed, eps, L, T):
    """io.cpp-layout config: [t][x][y][z][72] doubles; slot mu'*18+6a+2b+{re,im},
    mu'=0 is U_t, 1..3 are U_x, U_y, U_z.  The RNG counter of link (t,x,y,z,mu)
    with ILDG mu (0..3 = x,y,z,t) is ((((t*L+x)*L+y)*L+z)*4+mu); Gaussian j of
    that link uses counters 16*c+2j and 16*c+2j+1 (Box-Muller, cosine branch)."""
    t, x, y, z, mu = np.meshgrid(np.arange(T), np.arange(L), np.arange(L), np.arange(L),
                                 np.arange(4), indexing="ij")
    c = ((((t * L + x) * L + y) * L + z) * 4 + mu).astype(np.uint64)
    xi = np.empty(c.shape + (8,))
    for j in range(8):
        u1 = hash_u01(seed, c * np.uint64(16) + np.uint64(2 * j))
        u2 = hash_u01(seed, c * np.uint64(16) + np.uint64(2 * j + 1))
        xi[..., j] = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    H = np.einsum("...a,aij->...ij", xi, gell_mann()) / 2.0
    w, Q = np.linalg.eigh(H)
    U = np.einsum("...ik,...k,...jk->...ij", Q, np.exp(1j * eps * w), Q.conj())
    # ILDG mu 0..3 = x,y,z,t -> config slot mu' = 1,2,3,0
    out = np.empty((T, L, L, L, 4, 3, 3), dtype=np.complex128)
    out[..., 1:4, :, :] = U[..., 0:3, :, :]
    out[..., 0, :, :] = U[..., 3, :, :]
    return out.view(np.float64).reshape(-1).copy()


# ---------------------------------------------------------------------------
# LIME writer (ILDG binary record, big-endian doubles, file order t,z,y,x)
def _lime_header(rtype, nbytes, mb, me):
    import struct
    flags = (0x8000 if mb else 0) | (0x4000 if me else 0)
    h = struct.pack(">IHHQ", 0x456789AB, 1, flags, nbytes)
    return h + rtype.encode().ljust(128, b"\0")


def write_lime(path, config, T, L, extra_first=True):
    cfg = np.asarray(config, dtype=np.float64).reshape(T, L, L, L, 4, 3, 3, 2)  # t,x,y,z,mu',a,b,ri
    # file order: t, z, y, x; mu = x,y,z,t ; element (i,j) row-major
    ildg = np.empty((T, L, L, L, 4, 3, 3, 2))
    f = cfg.transpose(0, 3, 2, 1, 4, 5, 6, 7)  # t,z,y,x
    ildg[..., 0:3, :, :, :] = f[..., 1:4, :, :, :]
    ildg[..., 3, :, :, :] = f[..., 0, :, :, :]
    payload = ildg.astype(">f8").tobytes()
    xml = (f"<?xml version=\"1.0\" encoding=\"UTF-8\"?><ildgFormat><version>1.0</version>"
           f"<field>su3gauge</field><precision>64</precision><lx>{L}</lx><ly>{L}</ly>"
           f"<lz>{L}</lz><lt>{T}</lt></ildgFormat>").encode()
    with open(path, "wb") as fh:
        recs = []
        if extra_first:
            recs.append(("ildg-format", xml))
        recs.append(("ildg-binary-data", payload))
        for i, (rt, data) in enumerate(recs):
            fh.write(_lime_header(rt, len(data), i == 0, i == len(recs) - 1))
            fh.write(data)
            pad = (-len(data)) % 8
            fh.write(b"\0" * pad)


# ---------------------------------------------------------------------------
# links as (L,L,L,3,3,3): [x,y,z,dir,a,b]
def links_from_config(ts_config, L):
    c = np.asarray(ts_config, dtype=np.float64).reshape(L, L, L, 4, 3, 3, 2)
    c = c[..., 0] + 1j * c[..., 1]
    return np.ascontiguousarray(c[:, :, :, 1:4])


def shift(a, d, sgn):
    """a(x + sgn*e_d) for a field with leading (x,y,z) axes."""
    return np.roll(a, -sgn, axis=d)


def dag(m):
    return np.conj(np.swapaxes(m, -1, -2))


def proj_su3(M):
    """M (M^H M)^{-1/2} det(.)^{-1/3} (principal root), batched."""
    H = dag(M) @ M
    w, Q = np.linalg.eigh(H)
    Sinv = (Q * (1.0 / np.sqrt(w))[..., None, :]) @ dag(Q)
    W = M @ Sinv
    det = np.linalg.det(W)
    return W * (1.0 / det ** (1.0 / 3.0))[..., None, None]


def hyp(U, a1, a2, iters):
    U = U.copy()
    for _ in range(iters):
        dec = {}
        for mu in range(3):
            for rho in range(3):
                if rho == mu:
                    continue
                Um, Ur = U[..., mu, :, :], U[..., rho, :, :]
                st = Ur @ shift(Um, rho, 1) @ dag(shift(Ur, mu, 1))
                Urm = shift(Ur, rho, -1)
                st = st + dag(Urm) @ shift(Um, rho, -1) @ shift(shift(Ur, rho, -1), mu, 1)
                dec[(mu, rho)] = proj_su3((1 - a2) * Um + (a2 / 2) * st)
        new = np.empty_like(U)
        for mu in range(3):
            acc = np.zeros_like(U[..., mu, :, :])
            for nu in range(3):
                if nu == mu:
                    continue
                rho = 3 - mu - nu
                A = dec[(nu, rho)]
                B = dec[(mu, rho)]
                acc = acc + A @ shift(B, nu, 1) @ dag(shift(A, mu, 1))
                Am = shift(A, nu, -1)
                acc = acc + dag(Am) @ shift(B, nu, -1) @ shift(Am, mu, 1)
            new[..., mu, :, :] = proj_su3((1 - a1) * U[..., mu, :, :] + (a1 / 4) * acc)
        U = new
    return U


def apply_D(U, psi):
    """psi: (L,L,L,3).  D psi = sum_mu [U_mu psi(x+mu) + U_mu^H(x-mu) psi(x-mu)] - 6 psi."""
    out = -6.0 * psi
    for mu in range(3):
        Um = U[..., mu, :, :]
        out = out + np.einsum("...ab,...b->...a", Um, shift(psi, mu, 1))
        out = out + np.einsum("...ba,...b->...a", np.conj(shift(Um, mu, -1)), shift(psi, mu, -1))
    return out


def dense_D(U):
    L = U.shape[0]
    n = 3 * L ** 3
    D = np.zeros((n, n), dtype=np.complex128)
    for j in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[j] = 1
        D[:, j] = apply_D(U, e.reshape(L, L, L, 3)).reshape(-1)
    return D


def sparse_minus_D(U):
    """-D = 6 - H as a scipy CSR matrix assembled entry by entry from the
    links (config_utils.cpp:18-58: diagonal 6; row 3x+a, column 3(x+mu)+b:
    -U_mu(x)[a,b]; column 3(x-mu)+b: -U_mu(x-mu)^H[a,b]); site x = (x*L+y)*L+z,
    periodic neighbours (navigation.cpp:27-88)."""
    import scipy.sparse as sp
    L = U.shape[0]
    V3 = L ** 3
    idx = np.arange(V3).reshape(L, L, L)
    rows, cols, vals = [np.arange(3 * V3)], [np.arange(3 * V3)], [np.full(3 * V3, 6.0 + 0j)]
    a = np.arange(3)
    for mu in range(3):
        up = shift(idx, mu, 1).reshape(-1)
        dn = shift(idx, mu, -1).reshape(-1)
        Um = U[..., mu, :, :].reshape(V3, 3, 3)
        Udn = Um[dn]  # U_mu(x - mu)
        r = (3 * np.arange(V3)[:, None, None] + a[None, :, None]) * np.ones((1, 1, 3), dtype=np.int64)
        rows += [r.reshape(-1), r.reshape(-1)]
        cols += [(3 * up[:, None, None] + a[None, None, :] + 0 * r).reshape(-1),
                 (3 * dn[:, None, None] + a[None, None, :] + 0 * r).reshape(-1)]
        vals += [(-Um).reshape(-1), (-np.conj(np.swapaxes(Udn, 1, 2))).reshape(-1)]
    n = 3 * V3
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def cheb_filter(U, psi, lamL, lamC, deg):
    c = 2.0 / (lamL - lamC)
    a = 1.0 + c * lamC
    B = lambda v: c * apply_D(U, v) + a * v  # noqa: E731
    t_old, t_cur = psi, B(psi)
    for _ in range(2, deg + 1):
        t_old, t_cur = t_cur, 2 * B(t_cur) - t_old
    return t_cur


def gauge_transform(U, Om):
    """U'_mu(x) = Om(x)^H U_mu(x) Om(x+mu)   (timeslice.cpp:372-388)."""
    out = np.empty_like(U)
    for mu in range(3):
        out[..., mu, :, :] = dag(Om) @ U[..., mu, :, :] @ shift(Om, mu, 1)
    return out


# ---------------------------------------------------------------------------
# source shape (sourceshape_funcs.cpp:31-136)
def source_shape(V, L):
    """[((r1, r2, r3), |r|, psi)] for r in range(0, L, 2)^3 (r3 fastest):
    psi = sum_x sqrt(tr(M M^H)), M = V_{x+r} V_x^H over the nev eigenvectors
    (rows of V, file layout 3*site + colour)."""
    nev = V.shape[0]
    W = V.reshape(nev, L, L, L, 3)
    out = []
    for r1 in range(0, L, 2):
        for r2 in range(0, L, 2):
            for r3 in range(0, L, 2):
                Wr = np.roll(W, shift=(-r1, -r2, -r3), axis=(1, 2, 3))  # W(x + r)
                M = np.einsum("kxyza,kxyzb->xyzab", Wr, W.conj())
                psi = np.sqrt((np.abs(M) ** 2).sum(axis=(-1, -2))).sum()
                out.append(((r1, r2, r3), float(np.sqrt(r1 * r1 + r2 * r2 + r3 * r3)), float(psi)))
    return out


def avg_radii(results):
    """sourceshape_funcs.cpp:102-136: psi averaged over equal |r| (radii equal
    within 10e-12), in order of first appearance -> [(r, psi_avg)]."""
    acc = []
    for _, rad, psi in results:
        copied = False
        for a in acc:
            if abs(rad - a[1]) < 10e-12:
                a[0] += 1
                a[2] += psi
                copied = True
        if not copied:
            acc.append([1, rad, psi])
    return [(a[1], a[2] / a[0]) for a in acc]
