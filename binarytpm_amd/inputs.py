"""Seeded synthetic inputs of the Thermal stage (shapes, illumination,
eclipse masks, sparse view factors, crater model).

These stand in for the reference's geometry and view-factor producers
(asteroidgeometry / astshadow / surfacegeometry / Shape_Selfheating), which
are upstream of the conduction stage: they only need to produce inputs of the
right shapes and physical ranges.  Everything is deterministic in its seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

SIGMA = 5.67e-8
SOLAR_1AU = 1367.0          # W m^-2, T3905/main.cpp:905 (S = 1367 / r^2)
NS = 100


# --------------------------------------------------------------------------
# meshes
# --------------------------------------------------------------------------

def icosphere(subdiv: int):
    """Unit icosphere: 20 * 4**subdiv triangles (outward winding)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    V = np.array(v, dtype=np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array(f, dtype=np.int64)
    for _ in range(subdiv):
        verts = list(V)
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        V = np.array(verts)
        F = np.array(nf, dtype=np.int64)
    return V, F


def ellipsoid(subdiv: int, axes=(1.0, 1.0, 1.0)):
    V, F = icosphere(subdiv)
    return V * np.asarray(axes, dtype=np.float64), F


def facet_geometry(V, F):
    """Unit normals, centroids, areas and an orthonormal local frame
    (x along the first edge, z = normal) per facet."""
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    cr = np.cross(b - a, c - a)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    n = cr / np.linalg.norm(cr, axis=1, keepdims=True)
    cen = (a + b + c) / 3.0
    x = b - a
    x -= (x * n).sum(1, keepdims=True) * n
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    y = np.cross(n, x)
    frame = np.stack([x, y, n], axis=1)       # rows = local axes
    return n, cen, area, frame


def _rotz(phi):
    c, s = np.cos(phi), np.sin(phi)
    R = np.zeros(phi.shape + (3, 3))
    R[..., 0, 0], R[..., 0, 1] = c, -s
    R[..., 1, 0], R[..., 1, 1] = s, c
    R[..., 2, 2] = 1.0
    return R


# --------------------------------------------------------------------------
# single body: spin axis normal to the orbit, Sun in the equatorial plane
# --------------------------------------------------------------------------

def sun_body_frame(ntau: int, nrev: float = 1.0):
    """Sun direction in the body frame at each step (body spins about +z;
    Sun fixed along inertial +x)."""
    phi = 2.0 * np.pi * nrev * np.arange(ntau) / ntau
    return np.stack([np.cos(phi), -np.sin(phi), np.zeros_like(phi)], axis=1)


def direct_illumination(normals, sun, solar, mask=None):
    """D[f][s] = solar * max(0, n_f . u_s) * mask."""
    cosi = np.clip(normals @ sun.T, 0.0, None)
    D = solar * cosi
    if mask is not None:
        D = D * mask
    return D.astype(np.float32)


# --------------------------------------------------------------------------
# binary: eclipses by analytic ray casting against the partner body
# --------------------------------------------------------------------------

def _ray_hits_ellipsoid(P, d, C, R, axes):
    """Rays P + t d (t > 0) against ellipsoids (centre C, rotation R whose
    rows are the body axes in inertial coordinates, semi-axes `axes`)."""
    q = np.einsum("...ij,...j->...i", R, P - C) / axes
    w = np.einsum("...ij,...j->...i", R, d) / axes
    A = (w * w).sum(-1)
    B = 2.0 * (q * w).sum(-1)
    Cc = (q * q).sum(-1) - 1.0
    disc = B * B - 4 * A * Cc
    ok = disc > 0
    sq = np.sqrt(np.where(ok, disc, 0.0))
    t2 = (-B + sq) / (2 * A)
    return ok & (t2 > 1e-9)


@dataclass
class BinaryIllum:
    D_primary: np.ndarray      # [nf_p][ntau]
    D_secondary: np.ndarray    # [nf_s][ntau]
    mask_primary: np.ndarray   # [nf_p][ntau] 1 = sunlit (no eclipse)
    mask_secondary: np.ndarray


def binary_illumination(Vp, Fp, axes_p, Vs, Fs, radius_s, sep, ntau, nrev_p, nrev_s, solar):
    """Edge-on circular mutual orbit in the xy plane, Sun along +x; the
    secondary is tidally locked (spin = orbit).  Per step, a facet is in
    eclipse if the ray from its centroid toward the Sun meets the partner."""
    np_, cp, _, _ = facet_geometry(Vp, Fp)
    ns_, cs, _, _ = facet_geometry(Vs, Fs)
    u = np.array([1.0, 0.0, 0.0])
    s = np.arange(ntau)
    php = 2 * np.pi * nrev_p * s / ntau
    ths = 2 * np.pi * nrev_s * s / ntau
    Rp = _rotz(php)                       # body -> inertial
    Rs = _rotz(ths)
    Cs = np.stack([sep * np.cos(ths), sep * np.sin(ths), np.zeros_like(ths)], 1)
    Dp = np.zeros((len(Fp), ntau), np.float32)
    Ds = np.zeros((len(Fs), ntau), np.float32)
    Mp = np.ones((len(Fp), ntau), np.float32)
    Ms = np.ones((len(Fs), ntau), np.float32)
    for k in range(ntau):
        n_in = np_ @ Rp[k].T
        p_in = cp @ Rp[k].T
        cosi = np.clip(n_in @ u, 0.0, None)
        hit = _ray_hits_ellipsoid(p_in, u[None, :], Cs[k][None, :], np.eye(3)[None],
                                  np.array([radius_s] * 3))
        Mp[:, k] = np.where(hit, 0.0, 1.0)
        Dp[:, k] = solar * cosi * Mp[:, k]
        n_sin = ns_ @ Rs[k].T
        p_sin = cs @ Rs[k].T + Cs[k]
        cosi = np.clip(n_sin @ u, 0.0, None)
        hit = _ray_hits_ellipsoid(p_sin, u[None, :], np.zeros(3)[None], Rp[k].T[None],
                                  np.asarray(axes_p))
        Ms[:, k] = np.where(hit, 0.0, 1.0)
        Ds[:, k] = solar * cosi * Ms[:, k]
    return BinaryIllum(Dp, Ds, Mp, Ms)


def binary_shape(subdiv_p: int = 2, subdiv_s: int = 1, axes_p=(0.6, 0.5, 0.45),
                 radius_s: float = 0.3, sep: float = 1.2, dent: float = 0.0):
    """A binary as one vertex list for the geometry producer: the primary
    ellipsoid (vertices first; with dent > 0 its radius is modulated by
    1 - dent * max(0, sin(3 lon) cos(lat))^2, which makes concave troughs that
    shadow their neighbours) and a spherical secondary offset by sep along x
    in its own equatorial frame (its orbit).  Returns (equpos [nv][3] km,
    facets [nf][3], nv_primary)."""
    Vp, Fp = ellipsoid(subdiv_p, axes_p)
    if dent > 0.0:
        lon = np.arctan2(Vp[:, 1], Vp[:, 0])
        lat = np.arcsin(np.clip(Vp[:, 2] / np.linalg.norm(Vp, axis=1), -1.0, 1.0))
        Vp = Vp * (1.0 - dent * np.maximum(0.0, np.sin(3.0 * lon) * np.cos(lat)) ** 2)[:, None]
    Vs, Fs = ellipsoid(subdiv_s, (radius_s, radius_s, radius_s))
    Vs = Vs + np.array([sep, 0.0, 0.0])
    equ = np.vstack([Vp, Vs]).astype(np.float64)
    fv = np.vstack([Fp, Fs + len(Vp)]).astype(np.int32)
    return equ, fv, len(Vp)


def occlusion_scene(nx: int = 6, nblock: int = 16, seed: int = 0, gap: float = 1.0):
    """A floor (z = 0, facing up) and a ceiling (z = gap, facing down), each
    an nx x nx grid of triangle pairs over the unit square with vertex
    heights jittered by 1e-3, and nblock small one-sided triangles in
    between (half facing up, half down, tilted up to 20 degrees) that hide
    parts of the floor from the ceiling and back: a scene whose facet view
    factors need the occlusion cull.  Returns (vertices [nv][3], facets
    [nf][3])."""
    rng = np.random.default_rng(seed)
    V, F = [], []

    def grid(z, up, off):
        base = len(V)
        g = (np.arange(nx + 1) / nx)
        for y in g:
            for x in g:
                V.append((x + off[0], y + off[1], z + 1e-3 * rng.standard_normal()))
        for j in range(nx):
            for i in range(nx):
                a = base + j * (nx + 1) + i
                b, c, d = a + 1, a + nx + 1, a + nx + 2
                for t in ((a, b, d), (a, d, c)):
                    F.append(t if up else (t[0], t[2], t[1]))

    grid(0.0, True, (0.0, 0.0))
    grid(gap, False, tuple(0.1 * rng.random(2)))
    for k in range(nblock):
        c = np.array([0.2 + 0.6 * rng.random(), 0.2 + 0.6 * rng.random(),
                      gap * (0.35 + 0.3 * rng.random())])
        phi = 2.0 * math.pi * rng.random()
        tilt = math.radians(20.0) * rng.random()
        ang = phi + np.array([0.0, 2.0 * math.pi / 3.0, 4.0 * math.pi / 3.0])
        pts = np.stack([np.cos(ang), np.sin(ang), np.zeros(3)], axis=1) * 0.12
        pts[:, 2] = np.tan(tilt) * pts[:, 0]
        base = len(V)
        V.extend(map(tuple, pts + c))
        F.append((base, base + 1, base + 2) if k % 2 == 0 else (base, base + 2, base + 1))
    return np.asarray(V, np.float64), np.asarray(F, np.int32)


# --------------------------------------------------------------------------
# sparse facet view factors
# --------------------------------------------------------------------------

def sparse_viewfactors(nf: int, density: float = 0.03, vmax: float = 0.2, seed: int = 1,
                       rowsum_max: float = 1.0):
    """Seeded CSR view factors: about density*nf entries per row (no self
    term), values ~U[0, vmax], each row scaled so its sum is <= a seeded
    target in [0.2, rowsum_max].  Also returns a unit direction per entry
    in the upper hemisphere of the receiving facet's local frame."""
    rng = np.random.default_rng(seed)
    k = max(1, int(round(density * nf)))
    rowptr = np.zeros(nf + 1, dtype=np.int32)
    cols, vals = [], []
    for f in range(nf):
        c = rng.choice(nf - 1, size=min(k, nf - 1), replace=False)
        c = np.where(c >= f, c + 1, c)
        c.sort()
        v = rng.uniform(0.0, vmax, size=c.size)
        target = rng.uniform(0.2, rowsum_max)
        ssum = v.sum()
        if ssum > target:
            v *= target / ssum
        cols.append(c)
        vals.append(v)
        rowptr[f + 1] = rowptr[f] + c.size
    colidx = np.concatenate(cols).astype(np.int32)
    vf = np.concatenate(vals).astype(np.float32)
    nnz = colidx.size
    z = rng.uniform(0.05, 1.0, nnz)
    ph = rng.uniform(0, 2 * np.pi, nnz)
    rxy = np.sqrt(1 - z * z)
    dirs = np.stack([rxy * np.cos(ph), rxy * np.sin(ph), z], 1)
    return (rowptr, colidx, vf), dirs


# --------------------------------------------------------------------------
# crater: spherical cap of half-angle gamma, 100 triangles
# --------------------------------------------------------------------------

@dataclass
class CraterGeom:
    V: np.ndarray         # [61, 3] vertices (unit sphere, cap around -z)
    F: np.ndarray         # [100, 3]
    normals: np.ndarray   # [100, 3] unit normals pointing into the cavity
    area: np.ndarray      # [100]
    samples: np.ndarray   # [100, m*m, 3] sample points per sub-facet
    z_rim: float
    gamma: float
    vf: np.ndarray        # [100*100] vf[i*100+j] = VF(j->i)


def crater_cap(gamma_deg: float = 45.0, rings: int = 5, m: int = 10) -> CraterGeom:
    """Spherical-cap crater: rings of 4k vertices at polar angles
    gamma*k/rings from the bottom point; 4*rings^2 = 100 triangles for 5
    rings.  View factors by the spherical-cavity identity
    VF(j->i) = A_i / (4 pi R^2), zero diagonal."""
    g = math.radians(gamma_deg)
    V = [np.array([0.0, 0.0, -1.0])]
    ring_idx = [[0]]
    for k in range(1, rings + 1):
        th = g * k / rings
        n = 4 * k
        idx = []
        for i in range(n):
            ph = 2 * np.pi * i / n
            V.append(np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph),
                               -math.cos(th)]))
            idx.append(len(V) - 1)
        ring_idx.append(idx)
    V = np.array(V)

    def ang(i):
        return math.atan2(V[i, 1], V[i, 0]) % (2 * np.pi)

    F = []
    for k in range(1, rings + 1):
        inner, outer = ring_idx[k - 1], ring_idx[k]
        if k == 1:
            for i in range(len(outer)):
                F.append((inner[0], outer[i], outer[(i + 1) % len(outer)]))
            continue
        # zipper between two rings in angular order
        ni, no = len(inner), len(outer)
        a_i = [ang(v) for v in inner] + [ang(inner[0]) + 2 * np.pi]
        a_o = [ang(v) for v in outer] + [ang(outer[0]) + 2 * np.pi]
        i = o = 0
        while i < ni or o < no:
            if o < no and (i >= ni or a_o[o + 1] <= a_i[i + 1]):
                F.append((inner[i % ni], outer[o], outer[(o + 1) % no]))
                o += 1
            else:
                F.append((inner[i % ni], outer[o % no], inner[(i + 1) % ni]))
                i += 1
    F = np.array(F, dtype=np.int64)
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    cr = np.cross(b - a, c - a)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    n = cr / np.linalg.norm(cr, axis=1, keepdims=True)
    cen = (a + b + c) / 3.0
    flip = (n * cen).sum(1) > 0                   # normals toward the centre
    n[flip] *= -1
    # m*m sample points per triangle (centroids of the uniform subdivision)
    pts = []
    for i in range(m):
        for j in range(m - i):
            pts.append(((i + 1 / 3) / m, (j + 1 / 3) / m))
            if i + j < m - 1:
                pts.append(((i + 2 / 3) / m, (j + 2 / 3) / m))
    w = np.array(pts)                             # barycentric (w1, w2)
    samples = (a[:, None, :] + w[None, :, 0:1] * (b - a)[:, None, :]
               + w[None, :, 1:2] * (c - a)[:, None, :])
    nsf = len(F)
    vf = np.repeat((area / (4 * np.pi))[:, None], nsf, axis=1)   # [i][j] = A_i/(4 pi)
    np.fill_diagonal(vf, 0.0)
    return CraterGeom(V, F, n, area, samples, -math.cos(g), g,
                      vf.astype(np.float32).reshape(-1))


def crater_visibility(cg: CraterGeom, d: np.ndarray) -> np.ndarray:
    """Fraction of each sub-facet's sample points from which a ray toward the
    unit direction d (local frame, d_z > 0) leaves the cap through its
    opening; d is [nd, 3], result [nd, 100].  A ray from P inside the unit
    sphere exits the sphere at its far root; it escapes iff that exit point
    lies above the rim plane (the chord is straight, so a chord with both
    ends below the rim stays below it)."""
    P = cg.samples                                        # [100, S, 3]
    pd = np.einsum("fsk,dk->dfs", P, d)                   # [nd, 100, S]
    pp = (P * P).sum(-1)[None]                            # [1, 100, S]
    t = -pd + np.sqrt(np.maximum(pd * pd - (pp - 1.0), 0.0))
    qz = P[None, :, :, 2] + t * d[:, None, None, 2]
    lit = (qz >= cg.z_rim - 1e-12) & (d[:, None, None, 2] > 0)
    return lit.mean(-1)


class CraterTable:
    """Visibility of every sub-facet on a grid of local directions
    (elevation x azimuth), bilinearly interpolated."""

    def __init__(self, cg: CraterGeom, d_el: float = 2.0, d_az: float = 2.0):
        self.cg = cg
        self.el = np.radians(np.arange(0.0, 90.0 + 1e-9, d_el))
        self.az = np.radians(np.arange(0.0, 360.0, d_az))
        E, A = np.meshgrid(self.el, self.az, indexing="ij")
        d = np.stack([np.cos(E) * np.cos(A), np.cos(E) * np.sin(A), np.sin(E)], -1).reshape(-1, 3)
        d[:, 2] = np.maximum(d[:, 2], 1e-6)
        vis = np.concatenate([crater_visibility(cg, d[i:i + 512]) for i in range(0, len(d), 512)])
        self.vis = vis.reshape(len(self.el), len(self.az), -1).astype(np.float32)

    def lookup(self, d: np.ndarray) -> np.ndarray:
        """Visibility [n, 100] for unit directions d [n, 3] (0 if d_z <= 0)."""
        el = np.arcsin(np.clip(d[:, 2], -1, 1))
        az = np.arctan2(d[:, 1], d[:, 0]) % (2 * np.pi)
        fe = np.clip(el / (self.el[1] - self.el[0]), 0, len(self.el) - 1 - 1e-9)
        fa = az / (self.az[1] - self.az[0])
        ie = np.floor(fe).astype(np.int64)
        ia = np.floor(fa).astype(np.int64) % len(self.az)
        we = (fe - ie)[:, None].astype(np.float32)
        wa = (fa - np.floor(fa))[:, None].astype(np.float32)
        ie1 = np.minimum(ie + 1, len(self.el) - 1)
        ia1 = (ia + 1) % len(self.az)
        v = ((1 - we) * ((1 - wa) * self.vis[ie, ia] + wa * self.vis[ie, ia1])
             + we * ((1 - wa) * self.vis[ie1, ia] + wa * self.vis[ie1, ia1]))
        return np.where(d[:, 2:3] > 0, v, 0.0).astype(np.float32)


def crater_illumination(frames, sun, facet_mask, table: CraterTable, chunk: int = 65536):
    """shadow[f][s][j] (sunlit fraction, times the facet's own shadow/eclipse
    mask) and cosang[f][s][j] = max(0, n_j . u_local)."""
    nf, nt = frames.shape[0], sun.shape[0]
    loc = np.einsum("fij,sj->fsi", frames, sun).reshape(-1, 3)   # [nf*nt, 3]
    mask = facet_mask.reshape(-1) if facet_mask is not None else None
    shadow = np.empty((nf * nt, NS), np.float32)
    cosang = np.empty((nf * nt, NS), np.float32)
    n = table.cg.normals.astype(np.float32)
    for i in range(0, nf * nt, chunk):
        d = loc[i:i + chunk]
        v = table.lookup(d)
        if mask is not None:
            v *= mask[i:i + chunk, None]
        shadow[i:i + chunk] = v
        cosang[i:i + chunk] = np.clip(d.astype(np.float32) @ n.T, 0.0, None)
    return shadow.reshape(nf, nt, NS), cosang.reshape(nf, nt, NS)


def crater_sun_local(frames, sun, facet_mask=None) -> np.ndarray:
    """Sun direction in every facet's crater frame, [nf][ntau][3] float32,
    scaled by the facet's own sunlit fraction (the GPU-side crater
    illumination input; see btpm_crater.sun_local)."""
    loc = np.einsum("fij,sj->fsi", frames, sun)
    if facet_mask is not None:
        loc = loc * np.asarray(facet_mask)[:, :, None]
    return np.ascontiguousarray(loc, dtype=np.float32)


def shape_to_crater_vf(csr, dirs, table: CraterTable):
    """drf[k][j] = clamp(VF_k * vis_j(d_k) * max(0, n_j . d_k) / d_k,z, 0, 1)
    per CSR entry k (cf. calculate_surface_global_viewfactors,
    T3905/main.cpp:1054-1181)."""
    vfk = np.asarray(csr[2], np.float64)
    vis = table.lookup(dirs)
    cosj = np.clip(dirs @ table.cg.normals.T, 0.0, None)
    drf = vfk[:, None] * vis * cosj / np.maximum(dirs[:, 2:3], 1e-3)
    return np.clip(drf, 0.0, 1.0).astype(np.float32)


# --------------------------------------------------------------------------
# benchmark configurations (sizes of the stage's workloads)
# --------------------------------------------------------------------------

P_PRIMARY = 4.3 * 3600.0


def solar_at(r_au: float) -> float:
    return SOLAR_1AU / (r_au * r_au)


@dataclass
class SmoothCase:
    illum: np.ndarray           # [nf][ntau] direct (or total) illumination
    albedo: float
    ntau: int
    rot_period: float
    numrevs: Optional[np.ndarray] = None
    csr: Optional[tuple] = None
    dirs: Optional[np.ndarray] = None
    frames: Optional[np.ndarray] = None
    sun: Optional[np.ndarray] = None
    mask: Optional[np.ndarray] = None
    scatter: int = 0


def config0(subdiv: int = 4, ntau: int = 650, r_au: float = 1.1) -> SmoothCase:
    """Single body: subdivision-4 icosphere (5120 facets) scaled to a
    1.2 x 1.0 x 0.9 km ellipsoid, spin axis normal to the orbit, P = 4.3 h,
    650 steps, r = 1.1 AU, A = 0.1; no roughness, no self-heating."""
    V, F = ellipsoid(subdiv, (0.6, 0.5, 0.45))
    n, _, _, frames = facet_geometry(V, F)
    sun = sun_body_frame(ntau)
    D = direct_illumination(n, sun, solar_at(r_au))
    return SmoothCase(D, 0.1, ntau, P_PRIMARY, frames=frames, sun=sun)


def config1(ntau_per_primary_rev: int = 650, nrev_p: int = 15, nrev_s: int = 4, seed: int = 1):
    """Binary: config-0 primary + 1280-facet secondary (r = 0.3 km) on a
    circular 3.0 km orbit seen edge-on, tidally locked, 15:4 super-cycle
    (secondary period 16.125 h); eclipse masks by ray casting; 3 %-dense
    seeded facet view factors on the primary.  Facet numbering: primary
    first, then secondary."""
    ntau = ntau_per_primary_rev * nrev_p
    axes_p = (0.6, 0.5, 0.45)
    Vp, Fp = ellipsoid(4, axes_p)
    Vs, Fs = ellipsoid(3, (0.3, 0.3, 0.3))
    bi = binary_illumination(Vp, Fp, axes_p, Vs, Fs, 0.3, 3.0, ntau, nrev_p, nrev_s,
                             solar_at(1.1))
    nfp, nfs = len(Fp), len(Fs)
    (rp, ci, vf), dirs = sparse_viewfactors(nfp, 0.03, 0.2, seed)
    rowptr = np.concatenate([rp, np.full(nfs, rp[-1], np.int32)])
    D = np.concatenate([bi.D_primary, bi.D_secondary])
    numrevs = np.concatenate([np.full(nfp, float(nrev_p)), np.full(nfs, float(nrev_s))])
    # one cycle = the super-cycle; the primary period sets P (omega uses n/P)
    rot_period = P_PRIMARY * nrev_p
    return SmoothCase(D, 0.1, ntau, rot_period, numrevs=numrevs.astype(np.float32),
                      csr=(rowptr, ci, vf), dirs=dirs,
                      mask=np.concatenate([bi.mask_primary, bi.mask_secondary]), scatter=2)


def rough_primary(subdiv: int = 4, ntau: int = 650, r_au: float = 1.1, seed: int = 1,
                  gamma_deg: float = 45.0, nfac: Optional[int] = None, density: float = 0.03):
    """Config-2 shape: the config-1 primary alone (single-body cycle of 650
    steps), its seeded facet view factors, and a 45-degree spherical-cap
    crater (100 sub-facets) on every facet.  nfac (optional) keeps only the
    first nfac facets as crater parents (view factors still span the body).
    Returns (SmoothCase, crater inputs dict)."""
    V, F = ellipsoid(subdiv, (0.6, 0.5, 0.45))
    n, _, _, frames = facet_geometry(V, F)
    sun = sun_body_frame(ntau)
    solar = solar_at(r_au)
    D = direct_illumination(n, sun, solar)
    nf = len(F)
    csr, dirs = sparse_viewfactors(nf, density, 0.2, seed)
    cg = crater_cap(gamma_deg)
    table = CraterTable(cg)
    drf = shape_to_crater_vf(csr, dirs, table)
    case = SmoothCase(D, 0.1, ntau, P_PRIMARY, csr=csr, dirs=dirs, frames=frames, sun=sun,
                      scatter=2)
    k = nf if nfac is None else nfac
    shadow, cosang = crater_illumination(frames[:k], sun, np.ones((k, ntau), np.float32), table)
    crater = dict(vf=cg.vf, shadow=shadow, cosang=cosang, solar=solar, drf=drf, geom=cg,
                  table=table)
    return case, crater


def rough_binary_primary(ntau_per_primary_rev: int = 650, nrev_p: int = 15, nrev_s: int = 4,
                         seed: int = 1, gamma_deg: float = 45.0):
    """The sweep's shape on the config-1 binary: its 5120-facet primary over
    the 15:4 super-cycle (9750 steps, numrevs 15) with the eclipse masks the
    1280-facet secondary casts, the primary's seeded facet view factors and a
    45-degree cap crater on every facet.  The secondary only eclipses (its
    own temperatures are not part of these solves).  Crater sunlight comes as
    sun_local (each facet's Sun direction in its crater frame, scaled by its
    eclipse mask) with the sunlit-fraction table, which the GPU evaluates in
    its scattering kernel, so no [nf][ntau][100] arrays are built.
    Returns (SmoothCase of the primary, crater inputs dict)."""
    ntau = ntau_per_primary_rev * nrev_p
    axes_p = (0.6, 0.5, 0.45)
    Vp, Fp = ellipsoid(4, axes_p)
    Vs, Fs = ellipsoid(3, (0.3, 0.3, 0.3))
    bi = binary_illumination(Vp, Fp, axes_p, Vs, Fs, 0.3, 3.0, ntau, nrev_p, nrev_s,
                             solar_at(1.1))
    _, _, _, frames = facet_geometry(Vp, Fp)
    nf = len(Fp)
    csr, dirs = sparse_viewfactors(nf, 0.03, 0.2, seed)
    cg = crater_cap(gamma_deg)
    table = CraterTable(cg)
    drf = shape_to_crater_vf(csr, dirs, table)
    sun = sun_body_frame(ntau, nrev_p)
    case = SmoothCase(bi.D_primary, 0.1, ntau, P_PRIMARY * nrev_p,
                      numrevs=np.full(nf, float(nrev_p), np.float32), csr=csr, dirs=dirs,
                      frames=frames, sun=sun, mask=bi.mask_primary, scatter=2)
    crater = dict(vf=cg.vf, sun_local=crater_sun_local(frames, sun, bi.mask_primary),
                  solar=solar_at(1.1), drf=drf, geom=cg, table=table)
    return case, crater


def serving_shape(subdiv: int, seed: int, ntau: int = 650):
    """Config-4 request shape at 1 AU: a seeded icosphere-ellipsoid (semi-axes
    a = 0.3..1.2 km, b = a*U[0.7,1], c = b*U[0.7,1]; 20*4^subdiv facets), spin
    axis normal to the orbit, P = 4.3 h, A = 0.1, no facet view factors.
    Other heliocentric distances r are requests with flux scale 1/r^2.
    Returns (SmoothCase with the direct illumination, sun_local [nf][ntau][3])."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.3, 1.2)
    b = a * rng.uniform(0.7, 1.0)
    c = b * rng.uniform(0.7, 1.0)
    V, F = ellipsoid(subdiv, (a, b, c))
    n, _, _, frames = facet_geometry(V, F)
    sun = sun_body_frame(ntau)
    D = direct_illumination(n, sun, solar_at(1.0))
    case = SmoothCase(D, 0.1, ntau, P_PRIMARY, frames=frames, sun=sun)
    return case, crater_sun_local(frames, sun)


def sweep_inertias(n: int = 32, lo: float = 10.0, hi: float = 2500.0):
    """Log-spaced thermal inertias Gamma_k = lo * (hi/lo)^(k/(n-1)) (n = 1: lo)."""
    if n <= 1:
        return np.full(max(n, 0), lo)
    return lo * (hi / lo) ** (np.arange(n) / (n - 1))
