"""Independent numpy formulations used to pin the oracle (test-only).

Written from the mathematics, not from the oracle's code: fields are complex
arrays of shape (N0, N1, N2, N3, 4, 3, 3), neighbours come from np.roll, HYP
follows the Hasenfratz-Knechtli definition (hep-lat/0103029) with explicitly
labelled decorated links, and Wilson loops are walked link by link.
"""
import numpy as np

import oracle as O


def to_grid(u, dims):
    """oracle layout (V, 4, 18) -> complex (N0,N1,N2,N3,4,3,3)."""
    return O.to_complex(u).reshape(tuple(dims) + (4, 3, 3))


def from_grid(g):
    dims = g.shape[:4]
    return np.ascontiguousarray(O.from_complex(g).reshape(int(np.prod(dims)), 4, 18))


def up(a, mu, n=1):
    """a(x + n*mu)"""
    return np.roll(a, -n, axis=mu)


def dn(a, mu, n=1):
    """a(x - n*mu)"""
    return np.roll(a, n, axis=mu)


def dag(a):
    return np.conj(np.swapaxes(a, -1, -2))


def mm(*ms):
    out = ms[0]
    for m in ms[1:]:
        out = out @ m
    return out


def project(y, iters):
    """hep-lat/0506008: scale, then X = Y(3 - Y^dag Y)/2, Y = (1 - i Im det X / 3) X."""
    y = np.array(y, dtype=complex)
    d = np.sum(np.abs(y) ** 2, axis=(-1, -2))
    small = d < np.sqrt(18.0) * np.finfo(float).eps
    sc = np.where(small, 1.0, 1.0 / np.sqrt(np.where(small, 1.0, d) / 3.0))
    w = y * sc[..., None, None]
    eye = np.eye(3)
    for _ in range(iters):
        x = w @ (3 * eye - dag(w) @ w) / 2
        z = 1 - 1j * np.imag(np.linalg.det(x)) / 3
        w = z[..., None, None] * x
    return np.where(small[..., None, None], y, w)


def staple(A, B, mu, nu):
    """A_nu(x) B_mu(x+nu) A_nu(x+mu)^dag + A_nu(x-nu)^dag B_mu(x-nu) A_nu(x-nu+mu)."""
    an, bm = A[..., nu, :, :], B[..., mu, :, :]
    upst = mm(an, up(bm, nu), dag(up(an, mu)))
    dnst = mm(dag(dn(an, nu)), dn(bm, nu), up(dn(an, nu), mu))
    return upst + dnst


def ape_step(U, alpha, piter, dirs=(1, 2, 3)):
    out = U.copy()
    for mu in dirs:
        s = sum(staple(U, U, mu, eta) for eta in dirs if eta != mu)
        y = U[..., mu, :, :] + alpha * s
        out[..., mu, :, :] = project(y, piter) if piter > 0 else y
    return out


def _link(field, mu):
    return field[..., mu, :, :]


def hyp_textbook(U, a1, a2, a3, piter, dirs=(0, 1, 2, 3)):
    """Hasenfratz-Knechtli HYP with fat links labelled by their excluded directions."""
    P = (lambda y: project(y, piter)) if piter > 0 else (lambda y: y)
    dims4 = range(4)
    # Vbar[(mu, frozenset{nu,rho})]: decorated only in the remaining direction eta
    vbar = {}
    for mu in dims4:
        for nu in dims4:
            for rho in dims4:
                if len({mu, nu, rho}) < 3 or nu > rho:
                    continue
                eta = ({0, 1, 2, 3} - {mu, nu, rho}).pop()
                s = staple(U, U, mu, eta)
                vbar[(mu, frozenset((nu, rho)))] = P((1 - a3) * _link(U, mu) + a3 / 2 * s)

    def fat1(mu, excl):
        return vbar[(mu, frozenset(excl))]

    # Vtilde[(mu, nu)]: sum over rho != mu, nu of staples of Vbar_{rho;nu mu}, Vbar_{mu;rho nu}
    vt = {}
    for mu in dims4:
        for nu in dims4:
            if nu == mu:
                continue
            s = 0
            for rho in dims4:
                if rho in (mu, nu):
                    continue
                arho = fat1(rho, (nu, mu))
                bmu = fat1(mu, (rho, nu))
                s = s + mm(arho, up(bmu, rho), dag(up(arho, mu))) \
                    + mm(dag(dn(arho, rho)), dn(bmu, rho), up(dn(arho, rho), mu))
            vt[(mu, nu)] = P((1 - a2) * _link(U, mu) + a2 / 4 * s)
    out = U.copy()
    for mu in dirs:
        s = 0
        for nu in dims4:
            if nu == mu:
                continue
            anu = vt[(nu, mu)]
            bmu = vt[(mu, nu)]
            s = s + mm(anu, up(bmu, nu), dag(up(anu, mu))) \
                + mm(dag(dn(anu, nu)), dn(bmu, nu), up(dn(anu, nu), mu))
        out[..., mu, :, :] = P((1 - a1) * _link(U, mu) + a1 / 6 * s)
    return out


def wilson_loops_walk(U, levels, mwlt, mwlr):
    """W[x0, R-1, T-1, i, j] by walking each R x T loop link by link:
    from x along k with level i for R steps, up T steps in time, back along k
    at time x0+T with level j, down to x.  Normalised by 9 N1 N2 N3."""
    N0, N1, N2, N3 = U.shape[:4]
    msl = len(levels)
    W = np.zeros((N0, mwlr, mwlt, msl, msl))
    U0 = U[..., 0, :, :]
    eye = np.broadcast_to(np.eye(3, dtype=complex), U0.shape)
    for k in (1, 2, 3):
        for R in range(1, mwlr + 1):
            for T in range(1, mwlt + 1):
                # time line from y upwards: U0(y) U0(y+0) ... U0(y+(T-1)0)
                for i in range(msl):
                    for j in range(msl):
                        M = eye.copy()
                        for q in range(R):                       # out along k, level i
                            M = M @ up(levels[i][..., k - 1, :, :], k, q)
                        for q in range(T):                       # up in time at x+Rk
                            M = M @ up(up(U0, k, R), 0, q)
                        for q in reversed(range(R)):             # back along k, level j
                            M = M @ dag(up(up(levels[j][..., k - 1, :, :], 0, T), k, q))
                        for q in reversed(range(T)):             # down in time at x
                            M = M @ dag(up(U0, 0, q))
                        tr = np.real(np.trace(M, axis1=-2, axis2=-1))
                        W[:, R - 1, T - 1, i, j] += tr.sum(axis=(1, 2, 3))
    return W / (9.0 * N1 * N2 * N3)


def abelian_flux_field(dims, n12, n13, n23):
    """Spatial links U_k = diag(e^{i th_k}, e^{-i th_k}, 1) with a constant
    plaquette angle F_kl in each spatial plane (th_1 = 0, th_2 = F12 x1,
    th_3 = F13 x1 + F23 x2, F12 = 2 pi n12 / N1, F13 = 2 pi n13 / N1,
    F23 = 2 pi n23 / N2, so the field is periodic); time links 1.
    Returns the field [V][4][18] and {(k, l): F_kl}."""
    f12, f13, f23 = (2 * np.pi * n12 / dims[1], 2 * np.pi * n13 / dims[1], 2 * np.pi * n23 / dims[2])
    x = np.indices(dims)
    th = {1: np.zeros(dims), 2: f12 * x[1], 3: f13 * x[1] + f23 * x[2]}
    g = np.zeros(dims + (4, 3, 3), dtype=complex)
    g[..., 0, :, :] = np.eye(3)
    for k in (1, 2, 3):
        g[..., k, 0, 0] = np.exp(1j * th[k])
        g[..., k, 1, 1] = np.exp(-1j * th[k])
        g[..., k, 2, 2] = 1.0
    F = {(1, 2): f12, (1, 3): f13, (2, 3): f23}
    F.update({(b, a): v for (a, b), v in list(F.items())})
    return from_grid(g), F


def hyp3_abelian_factors(F, a1, a2, wiring):
    """3-d HYP without projection of an abelian constant-flux field: every
    smeared link is U_k times a real diagonal factor, per colour charge
    q = (1, -1, 0).  The staple sum of a plane with links scaled by a (nu) and
    b (mu) is a^2 b U_mu 2 cos(q F); level 1 decorates U_k in the plane (k,l)
    with g(F_kl) = (1 - a2) + a2 cos(q F_kl).  Level 2 takes, for the staple
    of link mu in the plane (mu,nu),
      wiring "reference": nu-links decorated in the plane (nu,mu) and mu-links
        decorated in (mu,nu) (hyp_spatial_links.c:187-234, dlnks_idx :116-121);
      wiring "textbook": nu-links decorated in (nu,rho) and mu-links in
        (mu,rho), rho the third direction (hep-lat/0103029).
    Returns {k: factor per colour (3,)}."""
    q = np.array([1.0, -1.0, 0.0])

    def g(f):
        return (1 - a2) + a2 * np.cos(q * f)

    out = {}
    for mu in (1, 2, 3):
        s = np.zeros(3)
        for nu in (1, 2, 3):
            if nu == mu:
                continue
            rho = 6 - mu - nu
            if wiring == "reference":
                a, b = g(F[(nu, mu)]), g(F[(mu, nu)])
            else:
                a, b = g(F[(nu, rho)]), g(F[(mu, rho)])
            s += a * a * b * np.cos(q * F[(mu, nu)])
        out[mu] = (1 - a1) + (a1 / 2) * s
    return out


def scale_links(u, dims, factors):
    """U_k -> U_k diag(factors[k]) for k = 1, 2, 3"""
    g = to_grid(u, dims).copy()
    for k, f in factors.items():
        g[..., k, :, :] = g[..., k, :, :] * f[None, :]
    return from_grid(g)
