"""Seeded synthetic meshes and per-frame motion for the collision workloads.

All randomness comes from SplitMix64 streams (integer arithmetic only) and
floats are built from 24-bit mantissa draws, so the same seed gives the same
bits on every IEEE machine. Geometry is computed in float64 and rounded once to
float32; only the UV sphere, the cloth field and the deformation use sin/cos
(numpy's float64 sin/cos), so inputs that feed committed fixtures are stored in
the fixture itself rather than regenerated.

Meshes are returned as ``(xyz, faces)``: ``xyz`` float32 of shape (V, 3),
``faces`` uint32 of shape (F, 3) -- the flattened form that
``OibvhTree::setup`` builds from a ``Mesh`` (oibvhTree.cu:157-170).
"""
from __future__ import annotations

import functools
import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n SplitMix64 outputs of stream ``seed`` starting at draw ``offset``."""
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """float64 values k * 2**-24, k uniform in [0, 2**24) (exact in float32)."""
    return (splitmix64(seed, n, offset) >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))


def uniform(seed: int, n: int, lo: float, hi: float, offset: int = 0) -> np.ndarray:
    return lo + (hi - lo) * uniform01(seed, n, offset)


class Rng:
    """Sequential draws from one SplitMix64 stream."""

    def __init__(self, seed: int):
        self.seed = seed
        self.pos = 0

    def u01(self, n: int) -> np.ndarray:
        out = uniform01(self.seed, n, self.pos)
        self.pos += n
        return out

    def uniform(self, n: int, lo: float, hi: float) -> np.ndarray:
        return lo + (hi - lo) * self.u01(n)

    def integers(self, n: int, lo: int, hi_incl: int) -> np.ndarray:
        z = splitmix64(self.seed, n, self.pos)
        self.pos += n
        span = np.uint64(hi_incl - lo + 1)
        return (z % span).astype(np.int64) + lo


# --------------------------------------------------------------------------
# icosphere
# --------------------------------------------------------------------------
def _normalize(v: np.ndarray) -> np.ndarray:
    return v / np.sqrt((v * v).sum(axis=1))[:, None]


@functools.lru_cache(maxsize=None)
def icosphere_unit(subdiv: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere in float64: V = 10*4^s + 2, F = 20*4^s (cached: treat
    the arrays as read-only)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array(
        [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
         [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array(
        [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
         [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
         [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
         [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v = _normalize(v)
    for _ in range(subdiv):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.minimum(e[:, 0], e[:, 1]) * nv + np.maximum(e[:, 0], e[:, 1])
        uniq, inv = np.unique(key, return_inverse=True)
        a, b = uniq // nv, uniq % nv
        mid = _normalize((v[a] + v[b]) * 0.5)
        m = inv.reshape(3, -1) + nv  # ab, bc, ca per face
        ab, bc, ca = m[0], m[1], m[2]
        f0, f1, f2 = f[:, 0], f[:, 1], f[:, 2]
        f = np.concatenate([
            np.stack([f0, ab, ca], 1), np.stack([f1, bc, ab], 1),
            np.stack([f2, ca, bc], 1), np.stack([ab, bc, ca], 1)])
        v = np.concatenate([v, mid])
    return v, f


def icosphere(subdiv: int, radius: float = 1.0, noise: float = 0.0, seed: int = 0,
              center=(0.0, 0.0, 0.0)) -> tuple[np.ndarray, np.ndarray]:
    """Icosphere whose vertices are scaled radially by (1 + u), u ~ U[-noise, noise]."""
    v, f = icosphere_unit(subdiv)
    if noise:
        v = v * (1.0 + uniform(seed, v.shape[0], -noise, noise))[:, None]
    xyz = (v * radius + np.asarray(center, dtype=np.float64)).astype(np.float32)
    return xyz, f.astype(np.uint32)


# --------------------------------------------------------------------------
# UV sphere
# --------------------------------------------------------------------------
def uv_sphere(nlon: int, nstack: int, radius: float = 1.0,
              center=(0.0, 0.0, 0.0)) -> tuple[np.ndarray, np.ndarray]:
    """UV sphere: V = nlon*(nstack-1) + 2, F = 2*nlon*(nstack-1)."""
    k = np.arange(1, nstack, dtype=np.float64)
    j = np.arange(nlon, dtype=np.float64)
    th = np.pi * k / nstack
    ph = 2.0 * np.pi * j / nlon
    st, ct = np.sin(th)[:, None], np.cos(th)[:, None]
    ring = np.stack([st * np.cos(ph)[None, :], st * np.sin(ph)[None, :],
                     np.broadcast_to(ct, (nstack - 1, nlon))], axis=-1).reshape(-1, 3)
    v = np.concatenate([[[0.0, 0.0, 1.0]], ring, [[0.0, 0.0, -1.0]]])
    nr = nstack - 1
    south = 1 + nr * nlon
    jj = np.arange(nlon)
    jn = (jj + 1) % nlon
    top = np.stack([np.zeros(nlon, np.int64), 1 + jj, 1 + jn], 1)
    r0 = np.arange(nr - 1)[:, None] * nlon + 1
    a = (r0 + jj[None, :]).ravel()
    b = (r0 + jn[None, :]).ravel()
    c = a + nlon
    d = b + nlon
    quads = np.concatenate([np.stack([a, c, b], 1), np.stack([b, c, d], 1)])
    last = 1 + (nr - 1) * nlon
    bot = np.stack([np.full(nlon, south, np.int64), last + jn, last + jj], 1)
    f = np.concatenate([top, quads, bot])
    xyz = (v * radius + np.asarray(center, dtype=np.float64)).astype(np.float32)
    return xyz, f.astype(np.uint32)


# --------------------------------------------------------------------------
# grids, boxes, cloth
# --------------------------------------------------------------------------
def _grid_faces(nu: int, nv: int) -> np.ndarray:
    """Two triangles per cell of an nu x nv vertex grid (row-major)."""
    i = np.arange(nu - 1)[:, None]
    j = np.arange(nv - 1)[None, :]
    a = (i * nv + j).ravel()
    b = a + 1
    c = a + nv
    d = c + 1
    f = np.empty((2 * a.size, 3), np.int64)
    f[0::2] = np.stack([a, b, d], 1)
    f[1::2] = np.stack([a, d, c], 1)
    return f


def plane(n: int, size: float = 1.0) -> tuple[np.ndarray, np.ndarray]:
    """n x n cells in [0,size]^2 at z = 0: V = (n+1)^2, F = 2n^2."""
    u = np.arange(n + 1, dtype=np.float64) * (size / n)
    x, y = np.meshgrid(u, u, indexing="ij")
    v = np.stack([x.ravel(), y.ravel(), np.zeros(x.size)], 1)
    return v.astype(np.float32), _grid_faces(n + 1, n + 1).astype(np.uint32)


def box(n: int, half: float = 0.5, noise: float = 0.0, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """Axis-aligned cube [-half, half]^3, each side an n x n grid with its own
    vertices: V = 6(n+1)^2, F = 12 n^2. With noise > 0 every vertex moves
    along its side's normal by U[-noise*half, noise*half]."""
    u = np.linspace(-half, half, n + 1)
    a, b = np.meshgrid(u, u, indexing="ij")
    a, b = a.ravel(), b.ravel()
    g = _grid_faces(n + 1, n + 1)
    vs, fs = [], []
    for k, (axis, sign) in enumerate([(0, 1), (0, -1), (1, 1), (1, -1), (2, 1), (2, -1)]):
        s = np.empty((a.size, 3))
        others = [i for i in range(3) if i != axis]
        s[:, axis] = sign * half
        s[:, others[0]] = a if sign > 0 else b
        s[:, others[1]] = b if sign > 0 else a
        if noise:
            s[:, axis] += uniform(seed + k, a.size, -noise * half, noise * half)
        vs.append(s)
        fs.append(g + k * a.size)
    return np.concatenate(vs).astype(np.float32), np.concatenate(fs).astype(np.uint32)


def cloth_positions(nv: int, amplitude: float = 0.2, wavelength: float = 0.05, seed: int = 0,
                    fold: float = 1.5, phase: float = 0.0) -> np.ndarray:
    """Vertex positions of cloth(); `phase` advances the height field's first
    phase (a deformation over time)."""
    a, b = uniform(seed, 2, 0.0, 2.0 * np.pi)
    a = a + phase
    t = np.arange(nv, dtype=np.float64) / (nv - 1)
    U, Vv = np.meshgrid(t, t, indexing="ij")
    w = 2.0 * np.pi / wavelength
    su = np.sin(w * U + a)
    x = U + fold / w * su
    z = amplitude * su * np.sin(w * Vv + b)
    return np.stack([x.ravel(), Vv.ravel(), z.ravel()], 1).astype(np.float32)


def cloth(nv: int, amplitude: float = 0.2, wavelength: float = 0.05, seed: int = 0,
          fold: float = 1.5, phase: float = 0.0) -> tuple[np.ndarray, np.ndarray]:
    """nv x nv vertex cloth over [0,1]^2, F = 2(nv-1)^2.

    z = A sin(2 pi u / wl + a) sin(2 pi v / wl + b) is the height field; the
    x coordinate is folded as x = u + fold * wl / (2 pi) * sin(2 pi u / wl + a),
    which turns back on itself whenever fold > 1, so the sheet self-intersects.
    Phases a, b come from the seed."""
    v = cloth_positions(nv, amplitude, wavelength, seed, fold, phase)
    return v, _grid_faces(nv, nv).astype(np.uint32)


# --------------------------------------------------------------------------
# motion
# --------------------------------------------------------------------------
def random_rotation(rng: Rng) -> np.ndarray:
    """Uniform rotation from a unit quaternion drawn by rejection in the
    4-ball (no trigonometry)."""
    while True:
        q = rng.uniform(4, -1.0, 1.0)
        n2 = float((q * q).sum())
        if 1e-6 < n2 <= 1.0:
            break
    w, x, y, z = q / math.sqrt(n2)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


class NormalWave:
    """Deformation v(t) = v0 + n0 * amp * sin(omega * t + k . v0) with the rest
    normal n0 (the unit radial direction of a sphere centred at `center`)."""

    def __init__(self, xyz0: np.ndarray, center, seed: int, amp: float = 0.05):
        p = xyz0.astype(np.float64)
        c = np.asarray(center, dtype=np.float64)
        d = p - c
        self.p = p
        self.n = d / np.sqrt((d * d).sum(axis=1))[:, None]
        r = Rng(seed)
        self.omega = float(r.uniform(1, 0.05, 0.25)[0])
        self.k = r.uniform(3, -8.0, 8.0)
        self.phase = p @ self.k
        self.amp = amp

    def frame(self, t: int) -> np.ndarray:
        s = self.amp * np.sin(self.omega * t + self.phase)
        return (self.p + self.n * s[:, None]).astype(np.float32)


# --------------------------------------------------------------------------
# workloads
# --------------------------------------------------------------------------
CFG1_SEED = 0x0B1B0001
CFG2_SEED = 0x0B1B0002
CFG3_SEED = 0x0B1B0003
CFG4_SEED = 0x0B1B0004
CFG5_SEED = 0x0B1B0005


def interpenetrating_icospheres(subdiv: int = 6, seed: int = CFG1_SEED, offset: float = 1.6,
                                noise: float = 0.02):
    """Two noisy unit icospheres, B offset by `offset` radii along x (the
    plain, CPU-runnable workload at subdiv 6: 81,920 triangles each)."""
    a = icosphere(subdiv, 1.0, noise, seed)
    b = icosphere(subdiv, 1.0, noise, seed + 1, center=(offset, 0.0, 0.0))
    return [a, b]


def deforming_uv_pair(nlon: int = 1000, nstack: int = 501, offset: float = 0.9,
                      seed: int = CFG2_SEED):
    """Two UV spheres (1,000,000 triangles each at 1000 x 501) at `offset`
    radii along x, each with its own NormalWave."""
    a = uv_sphere(nlon, nstack)
    b = uv_sphere(nlon, nstack, center=(offset, 0.0, 0.0))
    waves = [NormalWave(a[0], (0.0, 0.0, 0.0), seed), NormalWave(b[0], (offset, 0.0, 0.0), seed + 1)]
    return [a, b], waves


def _axis_rotation(axis: np.ndarray, angle: float) -> np.ndarray:
    """Rodrigues rotation (float64)."""
    a = axis / math.sqrt(float((axis * axis).sum()))
    K = np.array([[0.0, -a[2], a[1]], [a[2], 0.0, -a[0]], [-a[1], a[0], 0.0]])
    return np.eye(3) + math.sin(angle) * K + (1.0 - math.cos(angle)) * (K @ K)


class RigidScenes:
    """Batched workload: `n_scenes` scenes of U{lo..hi} rigid meshes, noisy
    icospheres or boxes, each with a triangle count drawn log-uniform in
    [fmin, fmax] (icospheres come in 20 * 4^s triangles, s = 3..6, boxes in
    12 n^2; a target an icosphere level cannot reach within its range becomes
    a box). Meshes are instances of a small library of base shapes (kind,
    size, noise variant) scaled to `radius`; every instance has its own rigid
    motion: a random initial rotation, a translation uniform in the unit cube,
    a spin about a random axis and a slow drift. transforms(f) gives the
    (n, 4, 4) float32 model matrices of frame f."""

    def __init__(self, n_scenes: int = 512, seed: int = CFG3_SEED, meshes=(2, 16), faces=(1000, 200000),
                 radius: float = 0.25, variants: int = 2, noise: float = 0.02):
        r = Rng(seed)
        self.library: list[tuple[np.ndarray, np.ndarray]] = []
        self.kinds: list[tuple] = []
        index: dict[tuple, int] = {}
        self.instances: list[tuple[int, int]] = []  # (scene, library index)
        lo, hi = math.log(faces[0]), math.log(faces[1])
        for s in range(n_scenes):
            k = int(r.integers(1, meshes[0], meshes[1])[0])
            for _ in range(k):
                target = math.exp(float(r.uniform(1, lo, hi)[0]))
                ico = r.u01(1)[0] < 0.5
                variant = int(r.integers(1, 0, variants - 1)[0])
                key = None
                if ico:
                    sub = int(round(math.log(target / 20.0) / math.log(4.0)))
                    if 3 <= sub <= 6 and faces[0] <= 20 * 4 ** sub <= faces[1] * 2:
                        key = ("ico", sub, variant)
                if key is None:
                    key = ("box", max(3, int(round(math.sqrt(target / 12.0)))), variant)
                if key not in index:
                    vs = seed * 131 + len(self.library) * 7 + 1
                    if key[0] == "ico":
                        xyz, f = icosphere(key[1], 1.0, noise, vs)
                    else:
                        xyz, f = box(key[1], 0.6, noise, vs)
                    index[key] = len(self.library)
                    self.library.append((xyz, f))
                    self.kinds.append(key)
                self.instances.append((s, index[key]))
        n = len(self.instances)
        self.n_scenes = n_scenes
        self.seed = seed
        self.noise = noise
        self.radius = radius
        self.R0 = np.stack([random_rotation(r) for _ in range(n)])
        self.t0 = r.uniform(3 * n, 0.0, 1.0).reshape(n, 3)
        self.axis = r.uniform(3 * n, -1.0, 1.0).reshape(n, 3)
        self.omega = r.uniform(n, 0.005, 0.03)
        self.vel = r.uniform(3 * n, -0.002, 0.002).reshape(n, 3)

    @property
    def triangles(self) -> int:
        return int(sum(self.library[b][1].shape[0] for _, b in self.instances))

    def unique_mesh(self, i: int) -> tuple[np.ndarray, np.ndarray]:
        """Instance i with geometry of its own: the base shape of its library
        entry (same kind and size) with a noise draw seeded by the instance
        index, so that no two meshes share vertex data."""
        key = self.kinds[self.instances[i][1]]
        vs = self.seed * 131 + 0x9E37 * (i + 1)
        if key[0] == "ico":
            return icosphere(key[1], 1.0, self.noise, vs)
        return box(key[1], 0.6, self.noise, vs)

    def transforms(self, frame: int) -> np.ndarray:
        n = len(self.instances)
        out = np.zeros((n, 4, 4), np.float64)
        for i in range(n):
            R = _axis_rotation(self.axis[i], self.omega[i] * frame) @ self.R0[i]
            out[i, :3, :3] = R * self.radius
            out[i, :3, 3] = self.t0[i] + self.vel[i] * frame
            out[i, 3, 3] = 1.0
        return out.astype(np.float32)


def precision_scenes(n_scenes: int = 64, seed: int = CFG4_SEED, n_plane: int = 354, n_box: int = 144):
    """Precision stress: scenes of two meshes (~250k triangles each at the
    default sizes: planes 2 * 354^2 = 250,632, boxes 12 * 144^2 = 248,832)
    placed coplanar or edge-touching with seeded perturbations of
    10^U[-7,-5] of the unit extent, so that many SAT comparisons are
    near-ties. Kinds cycle through: coplanar planes (B shifted in-plane, every
    vertex of B lifted by its own +-eps so that B's triangles straddle A's
    plane), planes meeting along an edge (B stood up at x = 1 + eps), boxes
    face to face (gap or overlap eps), boxes edge to edge."""
    r = Rng(seed)
    scenes = []
    pa = plane(n_plane)
    ba = box(n_box, 0.5)
    for s in range(n_scenes):
        kind = s % 4
        eps = 10.0 ** float(r.uniform(1, -7.0, -5.0)[0]) * (1.0 if r.u01(1)[0] < 0.5 else -1.0)
        if kind == 0:
            x, f = pa
            shift = r.uniform(2, 0.1, 0.5)
            nvx = x.shape[0]
            jit = 10.0 ** r.uniform(nvx, -7.0, -5.0) * np.where(r.u01(nvx) < 0.5, -1.0, 1.0)
            b = x.astype(np.float64) + np.array([shift[0], shift[1], 0.0])
            b[:, 2] = jit
            scenes.append([pa, (b.astype(np.float32), f)])
        elif kind == 1:
            x, f = pa
            p = x.astype(np.float64)
            b = np.stack([1.0 + eps + p[:, 2], p[:, 1] + float(r.uniform(1, -0.3, 0.3)[0]), p[:, 0] - 0.5], 1)
            scenes.append([pa, (b.astype(np.float32), f)])
        elif kind == 2:
            x, f = ba
            p = x.astype(np.float64)
            off = r.uniform(2, -0.4, 0.4)
            b = p + np.array([1.0 + eps, off[0], off[1]])
            scenes.append([ba, (b.astype(np.float32), f)])
        else:
            x, f = ba
            p = x.astype(np.float64)
            R = _axis_rotation(np.array([0.0, 0.0, 1.0]), math.pi / 4)
            b = p @ R.T + np.array([0.5 + math.sqrt(0.5) + eps, float(r.uniform(1, -0.3, 0.3)[0]), 0.0])
            scenes.append([ba, (b.astype(np.float32), f)])
    return scenes


def cloth_instances(n: int = 8, nv: int = 2048, seed: int = CFG5_SEED, amplitude: float = 0.2,
                    wavelength: float = 0.05):
    """Self-collision workload: n cloth sheets of nv x nv vertices, each with
    its own seeded phases (the perturbation between instances)."""
    return [cloth(nv, amplitude, wavelength, seed + i) for i in range(n)]
