"""Element stiffness restated from ``JaxSSO/element.py`` (oracle; test infrastructure only).

All functions are vectorised over a leading element axis and are written with
analytic operations only (``anorm`` instead of ``abs``-based norms, a tie-
averaging ``min``), so that the same code evaluated in complex128 gives
complex-step derivatives (Squire & Trapp 1998): d f/dx = Im f(x + i h)/h.
"""
from __future__ import annotations

import numpy as np


# ---------------------------------------------------------------------------
# analytic helpers
# ---------------------------------------------------------------------------

def anorm(v, axis=-1):
    """Euclidean norm sqrt(sum v^2) -- analytic continuation of ``jnp.linalg.norm``."""
    return np.sqrt(np.sum(v * v, axis=axis))


def aabs(x):
    """|x| continued analytically: x * sign(Re x)."""
    return x * np.sign(np.real(x))


TIE_RTOL = 1e-10


def tie_min(vals, axis=-1, rtol=None):
    """Minimum along ``axis``; derivative is the mean over the tied minima.

    JAX differentiates ``jnp.min`` (``JaxSSO/element.py:978``) by averaging the
    tangents of entries exactly equal to the minimum *in its own rounding*.  On
    symmetric (e.g. rectangular) elements the candidates tie in exact
    arithmetic and the rounded values split the ties by last-ulp accidents, so
    here entries within ``rtol`` (relative) of the minimum count as tied: JAX's
    rule applied to the exact values, independent of rounding.  The value
    returned is the minimum itself (the reference's value); the imaginary part
    (complex step) is the mean over the tie set.
    """
    rtol = TIE_RTOL if rtol is None else rtol
    re = np.real(vals)
    m = re.min(axis=axis, keepdims=True)
    mask = (re - m) <= rtol * np.abs(m)
    mean = np.sum(np.where(mask, vals, 0), axis=axis) / mask.sum(axis=axis)
    return np.squeeze(m, axis=axis) + 1j * np.imag(mean) if np.iscomplexobj(vals) else np.squeeze(m, axis=axis)


# ---------------------------------------------------------------------------
# Beam-column, JaxSSO/element.py:66-149
# ---------------------------------------------------------------------------

def beam_dircos(crds):
    """Direction-cosine matrix (N,3,3) of ``BeamCol.T`` (element.py:66-105).

    Returns (dirCos, yflag) where yflag marks members with Cxz == 0 exactly
    (parallel to global Y), for which the reference uses the special branch
    ``[[0,Cy,0],[-Cy,0,-1],[-Cy,0,0]]`` (element.py:92-94).
    """
    x1, y1, z1, x2, y2, z2 = [crds[:, k] for k in range(6)]
    length = np.sqrt((x1 - x2) ** 2 + (y1 - y2) ** 2 + (z1 - z2) ** 2)
    Cx = (x2 - x1) / length
    Cy = (y2 - y1) / length
    Cz = (z2 - z1) / length
    Cxz = anorm(np.stack([Cx, np.zeros_like(Cx), Cz], -1))
    yflag = np.real(Cxz) == 0
    safe = np.where(yflag, 1.0, Cxz)
    zero = np.zeros_like(Cx)
    one = np.ones_like(Cx)
    gen = np.stack([np.stack([Cx, Cy, Cz], -1),
                    np.stack([Cz / safe, zero, -Cx / safe], -1),
                    np.stack([-Cx * Cy / safe, Cxz, -Cy * Cz / safe], -1)], -2)
    spec = np.stack([np.stack([zero, Cy, zero], -1),
                     np.stack([-Cy, zero, -one], -1),
                     np.stack([-Cy, zero, zero], -1)], -2)
    return np.where(yflag[:, None, None], spec, gen), yflag


def blockdiag(d, nblk):
    n = d.shape[0]
    T = np.zeros((n, 3 * nblk, 3 * nblk), dtype=d.dtype)
    for b in range(nblk):
        T[:, 3 * b:3 * b + 3, 3 * b:3 * b + 3] = d
    return T


def beam_K_local(crds, E, G, Iy, Iz, J, A):
    """Local 12x12 Euler-Bernoulli stiffness of ``BeamCol.K_local`` (element.py:107-128)."""
    x1, y1, z1, x2, y2, z2 = [crds[:, k] for k in range(6)]
    L = np.sqrt((x1 - x2) ** 2 + (y1 - y2) ** 2 + (z1 - z2) ** 2)
    n = crds.shape[0]
    dt = np.result_type(crds, E, G, Iy, Iz, J, A)
    k = np.zeros((n, 12, 12), dtype=dt)
    ax = A * E / L
    bz3, bz2, bz1 = 12 * E * Iz / L ** 3, 6 * E * Iz / L ** 2, E * Iz / L
    by3, by2, by1 = 12 * E * Iy / L ** 3, 6 * E * Iy / L ** 2, E * Iy / L
    tj = G * J / L
    ent = {
        (0, 0): ax, (0, 6): -ax, (6, 6): ax,
        (1, 1): bz3, (1, 5): bz2, (1, 7): -bz3, (1, 11): bz2,
        (5, 5): 4 * bz1, (5, 7): -bz2, (5, 11): 2 * bz1,
        (7, 7): bz3, (7, 11): -bz2, (11, 11): 4 * bz1,
        (2, 2): by3, (2, 4): -by2, (2, 8): -by3, (2, 10): -by2,
        (4, 4): 4 * by1, (4, 8): by2, (4, 10): 2 * by1,
        (8, 8): by3, (8, 10): by2, (10, 10): 4 * by1,
        (3, 3): tj, (3, 9): -tj, (9, 9): tj,
    }
    for (i, j), v in ent.items():
        k[:, i, j] = v
        k[:, j, i] = v
    return k


def beam_Ke(crds, props):
    """Global 12x12 stiffness ``solve(T, K_local) @ T`` (element.py:130-139).

    crds (N,6) = [x1,y1,z1,x2,y2,z2]; props (N,6) = [E,G,Iy,Iz,J,A].
    """
    E, G, Iy, Iz, J, A = [props[:, k] for k in range(6)]
    kl = beam_K_local(crds, E, G, Iy, Iz, J, A)
    d, _ = beam_dircos(crds)
    T = blockdiag(d, 4)
    return np.linalg.solve(T, kl) @ T


# ---------------------------------------------------------------------------
# MITC-4 quad, JaxSSO/element.py:471-1106
# ---------------------------------------------------------------------------

GP = 1 / 3 ** 0.5
GAUSS = ((GP, GP), (-GP, GP), (-GP, -GP), (GP, -GP))
KB_DOFS = np.array([2, 3, 4, 8, 9, 10, 14, 15, 16, 20, 21, 22])
KM_DOFS = np.array([0, 1, 6, 7, 12, 13, 18, 19])
DRILL_DOFS = np.array([5, 11, 17, 23])
ROT_DIAG = np.array([1, 2, 4, 5, 7, 8, 10, 11])


def _vecs(crds):
    X = [crds[:, 3 * a:3 * a + 3] for a in range(4)]
    v32 = X[1] - X[2]
    v31 = X[0] - X[2]
    v34 = X[3] - X[2]
    v42 = X[1] - X[3]
    return v31, v32, v34, v42


def quad_frame(crds):
    """Unit local axes (x, y, z) of ``Quad.loc_crds``/``Quad.T`` (element.py:488-539, 643-696)."""
    v31, v32, v34, v42 = _vecs(crds)
    x_axis = v31
    z_axis = np.cross(x_axis, v42)
    y_axis = np.cross(z_axis, x_axis)
    x_axis = x_axis / anorm(x_axis)[:, None]
    y_axis = y_axis / anorm(y_axis)[:, None]
    z_axis = z_axis / anorm(z_axis)[:, None]
    return x_axis, y_axis, z_axis


def quad_loc_crds(crds):
    """Local in-plane coordinates [x1,y1,x2,y2,x3,y3,x4,y4] with node 3 at the origin."""
    v31, v32, v34, _ = _vecs(crds)
    xa, ya, _ = quad_frame(crds)
    dot = lambda a, b: np.sum(a * b, -1)
    zero = np.zeros(crds.shape[0], dtype=crds.dtype)
    return np.stack([dot(v31, xa), dot(v31, ya), dot(v32, xa), dot(v32, ya),
                     zero, zero, dot(v34, xa), dot(v34, ya)], -1)


def quad_J(xy, r, s):
    """2x2 Jacobian of ``Quad.J`` (element.py:698-709)."""
    x1, y1, x2, y2, x3, y3, x4, y4 = [xy[:, k] for k in range(8)]
    j00 = x1 * (s + 1) - x2 * (s + 1) + x3 * (s - 1) - x4 * (s - 1)
    j01 = y1 * (s + 1) - y2 * (s + 1) + y3 * (s - 1) - y4 * (s - 1)
    j10 = x1 * (r + 1) - x2 * (r - 1) + x3 * (r - 1) - x4 * (r + 1)
    j11 = y1 * (r + 1) - y2 * (r - 1) + y3 * (r - 1) - y4 * (r + 1)
    return 0.25 * np.stack([np.stack([j00, j01], -1), np.stack([j10, j11], -1)], -2)


def det2(J):
    return J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]


def quad_dH(xy, r, s):
    """Shape-function derivatives J^-1 dN/d(r,s) (element.py:721-722, 811-812)."""
    J = quad_J(xy, r, s)
    dN = 0.25 * np.array([[1 + s, -1 - s, -1 + s, 1 - s],
                          [1 + r, 1 - r, -1 + r, -1 - r]])
    rhs = np.broadcast_to(dN, (xy.shape[0], 2, 4)).astype(J.dtype)
    return np.linalg.solve(J, rhs)


def quad_B_kappa(xy, r, s):
    """Bending B (N,3,12), per node [w, theta_x, theta_y] (element.py:711-733)."""
    dH = quad_dH(xy, r, s)
    n = xy.shape[0]
    B = np.zeros((n, 3, 12), dtype=dH.dtype)
    for a in range(4):
        B[:, 0, 3 * a + 2] = -dH[:, 0, a]
        B[:, 1, 3 * a + 1] = dH[:, 1, a]
        B[:, 2, 3 * a + 1] = dH[:, 0, a]
        B[:, 2, 3 * a + 2] = -dH[:, 1, a]
    return B


def quad_B_gamma(xy, r, s):
    """MITC-4 assumed transverse-shear B (N,2,12) of ``Quad.B_gamma_MITC4`` (element.py:735-801)."""
    x1, y1, x2, y2, x3, y3, x4, y4 = [xy[:, k] for k in range(8)]
    Ax = x1 - x2 - x3 + x4
    Bx = x1 - x2 + x3 - x4
    Cx = x1 + x2 - x3 - x4
    Ay = y1 - y2 - y3 + y4
    By = y1 - y2 + y3 - y4
    Cy = y1 + y2 - y3 - y4
    zero = np.zeros_like(x1)
    r_axis = np.stack([(x1 + x4) / 2 - (x2 + x3) / 2, (y1 + y4) / 2 - (y2 + y3) / 2, zero], -1)
    s_axis = np.stack([(x1 + x2) / 2 - (x3 + x4) / 2, (y1 + y2) / 2 - (y3 + y4) / 2, zero], -1)
    r_axis = r_axis / anorm(r_axis)[:, None]
    s_axis = s_axis / anorm(s_axis)[:, None]
    det_J = det2(quad_J(xy, r, s))
    gr = ((Cx + r * Bx) ** 2 + (Cy + r * By) ** 2) ** 0.5 / (8 * det_J)
    gs = ((Ax + s * Bx) ** 2 + (Ay + s * By) ** 2) ** 0.5 / (8 * det_J)
    g_rz = gr[:, None] * np.stack([
        (1 + s) / 2 + zero, -(y1 - y2) * (1 + s) / 4, (x1 - x2) * (1 + s) / 4,
        -(1 + s) / 2 + zero, -(y1 - y2) * (1 + s) / 4, (x1 - x2) * (1 + s) / 4,
        -(1 - s) / 2 + zero, -(y4 - y3) * (1 - s) / 4, (x4 - x3) * (1 - s) / 4,
        (1 - s) / 2 + zero, -(y4 - y3) * (1 - s) / 4, (x4 - x3) * (1 - s) / 4], -1)
    g_sz = gs[:, None] * np.stack([
        (1 + r) / 2 + zero, -(y1 - y4) * (1 + r) / 4, (x1 - x4) * (1 + r) / 4,
        (1 - r) / 2 + zero, -(y2 - y3) * (1 - r) / 4, (x2 - x3) * (1 - r) / 4,
        -(1 - r) / 2 + zero, -(y2 - y3) * (1 - r) / 4, (x2 - x3) * (1 - r) / 4,
        -(1 + r) / 2 + zero, -(y1 - y4) * (1 + r) / 4, (x1 - x4) * (1 + r) / 4], -1)
    x_axis = np.array([1.0, 0.0, 0.0])
    cos_a = r_axis @ x_axis
    cos_b = s_axis @ x_axis
    sin_a = -anorm(np.cross(r_axis, x_axis))   # sign as element.py:795
    sin_b = anorm(np.cross(s_axis, x_axis))
    row0 = g_rz * sin_b[:, None] - g_sz * sin_a[:, None]
    row1 = -g_rz * cos_b[:, None] + g_sz * cos_a[:, None]
    return np.stack([row0, row1], -2)


def quad_B_m(xy, r, s):
    """Membrane B (N,3,8), per node [u, v] (element.py:803-818)."""
    dH = quad_dH(xy, r, s)
    n = xy.shape[0]
    B = np.zeros((n, 3, 8), dtype=dH.dtype)
    for a in range(4):
        B[:, 0, 2 * a] = dH[:, 0, a]
        B[:, 1, 2 * a + 1] = dH[:, 1, a]
        B[:, 2, 2 * a] = dH[:, 1, a]
        B[:, 2, 2 * a + 1] = dH[:, 0, a]
    return B


def quad_Cb(nu, E, h):
    c = E * h ** 3 / (12 * (1 - nu ** 2))
    z = np.zeros_like(c)
    o = np.ones_like(c)
    return c[:, None, None] * np.stack([np.stack([o, nu + z, z], -1),
                                        np.stack([nu + z, o, z], -1),
                                        np.stack([z, z, (1 - nu) / 2 + z], -1)], -2)


def quad_Cs(nu, E, h):
    c = E * h * (5 / 6) / (2 * (1 + nu))
    n = c.shape[0]
    return c[:, None, None] * np.broadcast_to(np.eye(2), (n, 2, 2))


def quad_Cm(nu, E, kx, ky):
    Ex = E * kx
    Ey = E * ky
    G = E / (2 * (1 + nu))
    z = np.zeros_like(Ex)
    m = np.stack([np.stack([Ex, nu * Ex, z], -1),
                  np.stack([nu * Ey, Ey, z], -1),
                  np.stack([z, z, (1 - nu * nu) * G], -1)], -2)
    return m / (1 - nu * nu)[:, None, None]


def quad_k12(crds, t, E, nu):
    """Bending + MITC-4 shear 12x12 (before the drilling spring), element.py:923-975."""
    xy = quad_loc_crds(crds)
    Cb = quad_Cb(nu, E, t)
    Cs = quad_Cs(nu, E, t)
    k = 0
    for (r, s) in GAUSS:
        dJ = det2(quad_J(xy, r, s))[:, None, None]
        Bk = quad_B_kappa(xy, r, s)
        k = k + np.swapaxes(Bk, 1, 2) @ (Cb @ Bk) * dJ
    k2 = 0
    for (r, s) in GAUSS:
        dJ = det2(quad_J(xy, r, s))[:, None, None]
        Bg = quad_B_gamma(xy, r, s)
        k2 = k2 + np.swapaxes(Bg, 1, 2) @ (Cs @ Bg) * dJ
    return k + k2


def quad_k_b(crds, t, E, nu):
    """24x24 bending block with drilling spring min|diag|/1000 (element.py:923-995)."""
    k = quad_k12(crds, t, E, nu)
    diag = k[:, ROT_DIAG, ROT_DIAG]
    k_rz = tie_min(aabs(diag)) / 1000
    n = crds.shape[0]
    kx = np.zeros((n, 24, 24), dtype=k.dtype)
    kx[:, KB_DOFS[:, None], KB_DOFS[None, :]] = k
    kx[:, DRILL_DOFS, DRILL_DOFS] = k_rz[:, None]
    return kx


def quad_k_m(crds, t, E, nu, kx_mod, ky_mod):
    """24x24 membrane block (element.py:1035-1071)."""
    xy = quad_loc_crds(crds)
    Cm = quad_Cm(nu, E, kx_mod, ky_mod)
    k = 0
    for (r, s) in GAUSS:
        B = quad_B_m(xy, r, s)
        dJ = det2(quad_J(xy, r, s))[:, None, None]
        k = k + np.swapaxes(B, 1, 2) @ (Cm @ B) * dJ
    k = t[:, None, None] * k
    n = crds.shape[0]
    kx = np.zeros((n, 24, 24), dtype=k.dtype)
    kx[:, KM_DOFS[:, None], KM_DOFS[None, :]] = k
    return kx


def quad_K_local(crds, props):
    t, E, nu, kxm, kym = [props[:, k] for k in range(5)]
    return quad_k_m(crds, t, E, nu, kxm, kym) + quad_k_b(crds, t, E, nu)


def quad_Ke(crds, props):
    """Global 24x24 stiffness ``solve(T, k_m + k_b) @ T`` (element.py:1073-1084).

    crds (N,12) = nodes i,j,m,n each [X,Y,Z]; props (N,5) = [t,E,nu,kx_mod,ky_mod].
    Note ``k_b`` takes kx_mod = ky_mod = 1 (element.py:1080); only the membrane
    uses the modifiers.
    """
    k = quad_K_local(crds, props)
    xa, ya, za = quad_frame(crds)
    d = np.stack([xa, ya, za], -2)
    T = blockdiag(d, 8)
    return np.linalg.solve(T, k) @ T


def quad_area(crds):
    """Surface area of ``Quad.A`` (element.py:471-487)."""
    X = [crds[:, 3 * a:3 * a + 3] for a in range(4)]
    e12 = X[1] - X[0]
    e14 = X[3] - X[0]
    e32 = X[1] - X[2]
    e34 = X[3] - X[2]
    return anorm(np.cross(e12, e14)) * 0.5 + anorm(np.cross(e34, e32)) * 0.5


def element_dofs(cnct):
    """Global DOF list per element: 6*tag .. 6*tag+5 per node in connectivity order."""
    c = np.asarray(cnct, dtype=np.int64)
    return (6 * c[:, :, None] + np.arange(6)[None, None, :]).reshape(c.shape[0], -1)


def gather_crds(crds, cnct):
    return np.asarray(crds)[np.asarray(cnct)].reshape(len(cnct), -1)
