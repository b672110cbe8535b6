"""Seeded synthetic structures used by the tests and the benchmark.

Every generator returns a :class:`Structure`: plain numpy arrays in the layout
``JaxSSO/model.py:221-338`` freezes a model into (``crds`` (n,3) f64,
``cnct_*`` int32, ``prop_beamcols`` (n,6) ``[E,G,Iy,Iz,J,A]``, ``prop_quads``
(n,5) ``[t,E,nu,kx_mod,ky_mod]``, ``known_id`` int32 DOF ids ``6*tag+k``,
``nodal_loads`` (6n,) f64).

Grid conventions: node id = j*(n+1)+i with x = i*Lx/n, y = j*Ly/n; quads are
``[BL, BR, TR, TL]`` row by row, the formula of
``Test/shells_fea_validation.ipynb`` cell 1; beams are every x-direction member
row by row, then every y-direction member, i-node at the lower id.
Random draws come from ``numpy.random.default_rng(SeedSequence([cfg, design]))``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

STEEL_E = 210e9
STEEL_SECTION = dict(A=3.0e-3, Iy=5.0e-6, Iz=5.0e-6, J=1.0e-5)


@dataclass
class Structure:
    crds: np.ndarray
    cnct_beamcols: np.ndarray
    prop_beamcols: np.ndarray
    cnct_quads: np.ndarray
    prop_quads: np.ndarray
    known_id: np.ndarray
    nodal_loads: np.ndarray
    name: str = ""
    design_nodes: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    extra_loads: list = field(default_factory=list)  # additional load cases (6n,)

    @property
    def n_node(self) -> int:
        return self.crds.shape[0]

    @property
    def ndof(self) -> int:
        return 6 * self.crds.shape[0]

    @property
    def n_beamcol(self) -> int:
        return self.cnct_beamcols.shape[0]

    @property
    def n_quad(self) -> int:
        return self.cnct_quads.shape[0]

    def load_cases(self) -> np.ndarray:
        """All load cases stacked as (n_case, 6n)."""
        return np.stack([self.nodal_loads] + list(self.extra_loads))


def _rng(cfg: int, design: int = 0) -> np.random.Generator:
    return np.random.default_rng(np.random.SeedSequence([cfg, design]))


def grid_nodes(n: int, lx: float, ly: float):
    i = np.tile(np.arange(n + 1), n + 1)
    j = np.repeat(np.arange(n + 1), n + 1)
    return i * (lx / n), j * (ly / n), i, j


def grid_quads(n: int) -> np.ndarray:
    j, i = np.divmod(np.arange(n * n), n)
    bl = j * (n + 1) + i
    return np.stack([bl, bl + 1, bl + n + 2, bl + n + 1], axis=1).astype(np.int32)


def grid_beams(n: int) -> np.ndarray:
    jx, ix = np.divmod(np.arange(n * (n + 1)), n)          # x-members: n per row, n+1 rows
    a = jx * (n + 1) + ix
    bx = np.stack([a, a + 1], 1)
    iy, jy = np.divmod(np.arange(n * (n + 1)), n)          # y-members, column i, segment j
    a = jy * (n + 1) + iy
    by = np.stack([a, a + n + 1], 1)
    # y-members listed row by row as well (segment j outer, column i inner)
    order = np.lexsort((by[:, 0] % (n + 1), by[:, 0] // (n + 1)))
    return np.concatenate([bx, by[order]]).astype(np.int32)


def _support_ids(nodes, pattern) -> np.ndarray:
    pattern = np.asarray(pattern)
    active = np.nonzero(pattern == 1)[0]
    nodes = np.asarray(nodes, dtype=np.int64)
    return (6 * nodes[:, None] + active[None, :]).ravel().astype(np.int32)


def _loads(n_node: int, nodes, fz: float) -> np.ndarray:
    f = np.zeros(6 * n_node)
    f[6 * np.asarray(nodes, dtype=np.int64) + 2] = fz
    return f


def _beam_props(n_el: int) -> np.ndarray:
    E = STEEL_E
    G = E / 2.6
    s = STEEL_SECTION
    row = np.array([E, G, s["Iy"], s["Iz"], s["J"], s["A"]])
    return np.tile(row, (n_el, 1))


def gridshell(n: int, span: float = 10.0, rise: float = 2.0, noise: float = 0.05,
              cfg: int = 0, design: int = 0, half_span_cases: bool = False) -> Structure:
    """Beam-column gridshell on an n x n bay grid (configurations 0 and 3)."""
    rng = _rng(cfg, design)
    x, y, i, j = grid_nodes(n, span, span)
    z = rise * np.sin(np.pi * x / span) * np.sin(np.pi * y / span)
    z = z + rng.normal(0.0, noise, size=z.shape)
    crds = np.stack([x, y, z], 1)
    cnct = grid_beams(n)
    perim = (i == 0) | (i == n) | (j == 0) | (j == n)
    free = np.nonzero(~perim)[0]
    known = _support_ids(np.nonzero(perim)[0], [1, 1, 1, 0, 0, 0])
    loads = _loads(crds.shape[0], free, -1000.0)
    extra = []
    if half_span_cases:
        h = span / 2
        for sel in (x[free] >= h, x[free] <= h, y[free] >= h):
            extra.append(_loads(crds.shape[0], free[sel], -1000.0))
    return Structure(crds, cnct, _beam_props(cnct.shape[0]), np.zeros((0, 4), np.int32),
                     np.zeros((0, 5)), known, loads, name=f"gridshell{n}x{n}",
                     design_nodes=free.astype(np.int32), extra_loads=extra)


def cfg0() -> Structure:
    """10x10-bay beam gridshell over 10 m x 10 m: 121 nodes, 220 beam-columns."""
    return gridshell(10, cfg=0)


def cfg1(design: int = 0) -> Structure:
    """40x40 MITC-4 shell on 10 m x 10 m, four fully fixed corners, -2 kN/node.

    ``design`` > 0 draws another seeded thickness field of the same shell."""
    n, L = 40, 10.0
    rng = _rng(1, design)
    x, y, i, j = grid_nodes(n, L, L)
    z = 1.5 * np.sin(np.pi * x / L) * np.sin(np.pi * y / L)
    cnct = grid_quads(n)
    t = 0.10 * rng.uniform(0.9, 1.1, size=cnct.shape[0])
    props = np.stack([t, np.full_like(t, 30e9), np.full_like(t, 0.2),
                      np.ones_like(t), np.ones_like(t)], 1)
    corners = np.array([0, n, n * (n + 1), (n + 1) ** 2 - 1])
    free = np.setdiff1d(np.arange((n + 1) ** 2), corners)
    return Structure(np.stack([x, y, z], 1), np.zeros((0, 2), np.int32), np.zeros((0, 6)), cnct,
                     props, _support_ids(corners, [1] * 6), _loads((n + 1) ** 2, free, -2000.0),
                     name="shell40x40", design_nodes=free.astype(np.int32))


def _fourier_thickness(rng, xc, yc, L, lo=0.08, hi=0.15, modes=8):
    field_ = np.zeros_like(xc)
    for _ in range(modes):
        c = rng.normal()
        while True:
            p, q = rng.integers(0, 4, size=2)
            if p or q:
                break
        phi = rng.uniform(0, 2 * np.pi)
        field_ += c * np.cos(2 * np.pi * (p * xc + q * yc) / L + phi)
    f0, f1 = field_.min(), field_.max()
    return lo + (hi - lo) * (field_ - f0) / (f1 - f0)


def shell_pinned_edges(n: int, L: float, rise: float, cfg: int = 2, design: int = 0,
                       load: float = -1000.0) -> Structure:
    """n x n MITC-4 shell, all edge nodes pinned, smooth Fourier thickness (configuration 2)."""
    rng = _rng(cfg, design)
    x, y, i, j = grid_nodes(n, L, L)
    z = rise * np.sin(np.pi * x / L) * np.sin(np.pi * y / L)
    cnct = grid_quads(n)
    xc = x[cnct].mean(1)
    yc = y[cnct].mean(1)
    t = _fourier_thickness(rng, xc, yc, L)
    props = np.stack([t, np.full_like(t, 30e9), np.full_like(t, 0.2),
                      np.ones_like(t), np.ones_like(t)], 1)
    edge = (i == 0) | (i == n) | (j == 0) | (j == n)
    free = np.nonzero(~edge)[0]
    return Structure(np.stack([x, y, z], 1), np.zeros((0, 2), np.int32), np.zeros((0, 6)), cnct,
                     props, _support_ids(np.nonzero(edge)[0], [1, 1, 1, 0, 0, 0]),
                     _loads((n + 1) ** 2, free, load), name=f"shell{n}x{n}",
                     design_nodes=free.astype(np.int32))


def cfg2() -> Structure:
    """200x200 MITC-4 shell over 20 m x 20 m (40,401 nodes, 242,406 DOF)."""
    return shell_pinned_edges(200, 20.0, 3.0)


def cfg3(n_designs: int = 256):
    """Batch of uneven beam gridshells (8..32 bays), 4 load cases each."""
    rng = _rng(3, 10**6)
    sizes = rng.integers(8, 33, size=n_designs)
    return [gridshell(int(nb), noise=0.1, cfg=3, design=d, half_span_cases=True)
            for d, nb in enumerate(sizes)]


def _gaussian_smooth(a: np.ndarray, sigma: float) -> np.ndarray:
    """Separable Gaussian filter with reflect boundaries (truncate = 4 sigma)."""
    r = int(4 * sigma + 0.5)
    k = np.exp(-0.5 * (np.arange(-r, r + 1) / sigma) ** 2)
    k /= k.sum()
    out = a
    for ax in (0, 1):
        pad = [(0, 0), (0, 0)]
        pad[ax] = (r, r)
        p = np.pad(out, pad, mode="symmetric")
        out = np.apply_along_axis(lambda v: np.convolve(v, k, mode="valid"), ax, p)
    return out


def simp_shell(design: int, n: int = 100, L: float = 15.0) -> Structure:
    """One SIMP shell design of configuration 4 (100x100 quads over 15 m x 15 m)."""
    rng = _rng(4, design)
    x, y, i, j = grid_nodes(n, L, L)
    z = np.zeros_like(x)
    for _ in range(6):
        h = rng.uniform(0, 2)
        cx, cy = rng.uniform(0, L, size=2)
        s = rng.uniform(1.5, 4)
        z += h * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
    cnct = grid_quads(n)
    s = _gaussian_smooth(rng.normal(size=(n, n)), 3.0).ravel()
    rho = np.clip(0.5 + 0.25 * s / s.std(), 0.01, 1.0)
    E = rho ** 3 * 30e9
    props = np.stack([np.full_like(E, 0.12), E, np.full_like(E, 0.2),
                      np.ones_like(E), np.ones_like(E)], 1)
    m = n // 2
    sup = [0, n, n * (n + 1), (n + 1) ** 2 - 1, m, n * (n + 1) + m, m * (n + 1), m * (n + 1) + n]
    free = np.setdiff1d(np.arange((n + 1) ** 2), sup)
    st = Structure(np.stack([x, y, z], 1), np.zeros((0, 2), np.int32), np.zeros((0, 6)), cnct,
                   props, _support_ids(sup, [1, 1, 1, 0, 0, 0]),
                   _loads((n + 1) ** 2, free, -1500.0), name=f"simp{n}x{n}",
                   design_nodes=free.astype(np.int32))
    st.rho = rho
    return st


def cfg4(n_designs: int = 64):
    return [simp_shell(d) for d in range(n_designs)]


# ---------------------------------------------------------------------------
# Cases recorded in the reference's validation notebooks
# ---------------------------------------------------------------------------

def notebook_arch() -> Structure:
    """Parabolic arch of 99 beam-columns, ``Test/beamcols_fea_validation.ipynb`` cell 1."""
    n_node, Q, rise, span = 100, 500, 5, 10
    x = np.linspace(0, span, n_node)
    y = np.zeros(n_node)
    z = -(rise / (span ** 2 / 4)) * ((x - span / 2) ** 2 - span ** 2 / 4)
    z[0] = 0
    z[n_node - 1] = 0
    design = np.arange(1, n_node - 1)
    cnct = np.stack([np.arange(n_node - 1), np.arange(1, n_node)], 1).astype(np.int32)
    E = 1.999e08
    G = E / (2 * (1 + 0.3))
    Iy, Iz = 6.572e-05, 3.301e-06
    props = np.tile([E, G, Iy, Iz, Iy + Iz, 4.265e-03], (n_node - 1, 1))
    known = _support_ids([0, n_node - 1], [1, 1, 1, 1, 0, 1])
    loads = _loads(n_node, design, -Q)
    return Structure(np.stack([x, y, z], 1), cnct, props, np.zeros((0, 4), np.int32),
                     np.zeros((0, 5)), known, loads, name="arch", design_nodes=design.astype(np.int32))


def notebook_barrel(support=(1, 1, 1, 0, 0, 0)) -> Structure:
    """19x19 MITC-4 barrel vault, ``Test/shells_fea_validation.ipynb`` cells 1 and 3."""
    n_col = n_row = 19
    span = 19
    xs = np.tile(np.linspace(0, span, n_col + 1), n_row + 1)
    ys = np.tile(np.linspace(0, span, n_row + 1), (n_col + 1, 1)).T.reshape(-1)
    zs = 0.05 * (-(np.linspace(0, span, n_col + 1) - span / 2) ** 2 + (span / 2) ** 2)
    zs = np.tile(zs, n_row + 1)
    zs = np.where(xs == 0, 0, zs)
    zs = np.where(xs == span, 0, zs)
    n_node = (n_col + 1) * (n_row + 1)
    n_ele = n_col * n_row
    cnct = np.zeros((n_ele, 4), dtype=np.int32)
    for e in range(n_ele):
        i_row = e // n_col
        cnct[e, 3] = e + i_row + n_col + 1
        cnct[e, 2] = cnct[e, 3] + 1
        cnct[e, 1] = cnct[e, 2] - (n_col + 1)
        cnct[e, 0] = cnct[e, 1] - 1
    t, E, nu = 0.25, 24855578.0 * 1e-3, 0.2
    design = np.nonzero((xs != 0) & (xs != span))[0]
    fixed = np.nonzero(~((xs != 0) & (xs != span)))[0]
    Q = 500 * 400 / n_node
    props = np.tile([t, E, nu, 1.0, 1.0], (n_ele, 1))
    return Structure(np.stack([xs, ys, zs], 1), np.zeros((0, 2), np.int32), np.zeros((0, 6)),
                     cnct, props, _support_ids(fixed, support), _loads(n_node, design, -Q),
                     name="barrel", design_nodes=design.astype(np.int32))


def notebook_frames(n: int, m: int = 100) -> Structure:
    """Multi-span arch chain ``f(n, m)`` of ``Test/Frames_speed.ipynb`` cell 1."""
    Q, rise, span = 50000, 10, 30
    n_node = m * n + 1
    x = np.linspace(0, span * m, n_node)
    y = np.zeros(n_node)
    z = -(rise / (span ** 2 / 4)) * ((x % span - span / 2) ** 2 - span ** 2 / 4)
    ids = np.arange(n_node)
    design = ids[ids % n != 0]
    fixed = ids[ids % n == 0]
    cnct = np.stack([ids[:-1], ids[1:]], 1).astype(np.int32)
    E = 1.999e08
    G = E / (2 * (1 + 0.3))
    Iy, Iz = 6.572e-05, 3.301e-06
    props = np.tile([E, G, Iy, Iz, Iy + Iz, 4.265e-03], (n_node - 1, 1))
    return Structure(np.stack([x, y, z], 1), cnct, props, np.zeros((0, 4), np.int32),
                     np.zeros((0, 5)), _support_ids(fixed, [1, 1, 1, 1, 0, 1]),
                     _loads(n_node, design, -Q / (m * (n - 1))), name=f"frames{n}x{m}",
                     design_nodes=design.astype(np.int32))
