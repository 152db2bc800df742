"""Seeded random ellipse / ellipsoid phantoms and their exact line integrals.

The chord length through a rotated ellipse follows
``CTLAB::Ellipsoid::intersectionLength`` (src/Ellipsoid.h:209-247): move the
line into the ellipse frame, then the two roots of the quadratic give the
chord.  It is generalised here to ellipsoids rotated about z.

Random numbers come from SplitMix64 (Steele, Lea, Flood, OOPSLA 2014) so that
every consumer (tests, bench, training data) builds the same phantom from a
seed.  Draw order per object: centre x, y, (z), semi-axis a, b, (c), angle,
attenuation; uniform doubles are the top 53 bits of each output.
"""
import numpy as np

_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = int(seed) & _M64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def uniform_array(self, n, lo=0.0, hi=1.0):
        return np.array([self.uniform(lo, hi) for _ in range(n)], dtype=np.float64)


def random_ellipsoids(seed, n, XN, YN, ZN, dx, dz, semi_xy=(5.0, 100.0), semi_z=(2.0, 32.0),
                      mu=(0.0, 0.04)):
    """n ellipsoids: centres uniform over the volume box, in-plane semi-axes in
    voxels, z semi-axis in slices, rotation about z in [0, pi), attenuation /mm."""
    r = SplitMix64(seed)
    out = []
    for _ in range(n):
        cx = r.uniform(-XN * dx / 2, XN * dx / 2)
        cy = r.uniform(-YN * dx / 2, YN * dx / 2)
        cz = r.uniform(-ZN * dz / 2, ZN * dz / 2) if ZN > 1 else 0.0
        a = r.uniform(*semi_xy) * dx
        b = r.uniform(*semi_xy) * dx
        c = r.uniform(*semi_z) * dz if ZN > 1 else np.inf
        ang = r.uniform(0.0, np.pi)
        m = r.uniform(*mu)
        out.append((cx, cy, cz, a, b, c, ang, m))
    return np.array(out, dtype=np.float64)


def random_ellipses(seed, n, XN, YN, dx, semi=(5.0, 100.0), mu=(0.0, 0.04)):
    return random_ellipsoids(seed, n, XN, YN, 1, dx, 1.0, semi_xy=semi, mu=mu)


def voxelize(ells, XN, YN, ZN, dx, dz, cx=0.0, cy=0.0, cz=0.0):
    """Sum of ellipsoid indicators sampled at voxel centres, (Z,X,Y) order:
    returns float32 array of shape (YN, XN, ZN)."""
    xs = cx + (np.arange(XN) - (XN - 1) / 2.0) * dx
    ys = cy + (np.arange(YN) - (YN - 1) / 2.0) * dx
    zs = cz + (np.arange(ZN) - (ZN - 1) / 2.0) * dz
    vol = np.zeros((YN, XN, ZN), dtype=np.float64)
    for (ex, ey, ez, a, b, c, ang, m) in ells:
        r = max(a, b)
        ix = np.nonzero(np.abs(xs - ex) <= r)[0]
        iy = np.nonzero(np.abs(ys - ey) <= r)[0]
        iz = np.arange(ZN) if np.isinf(c) else np.nonzero(np.abs(zs - ez) <= c)[0]
        if ix.size == 0 or iy.size == 0 or iz.size == 0:
            continue
        Y, X = np.meshgrid(ys[iy], xs[ix], indexing="ij")
        ca, sa = np.cos(ang), np.sin(ang)
        u = (X - ex) * ca + (Y - ey) * sa
        w = -(X - ex) * sa + (Y - ey) * ca
        q = (u / a) ** 2 + (w / b) ** 2                     # (ny, nx)
        if np.isinf(c):
            inside = np.broadcast_to((q <= 1.0)[:, :, None], q.shape + (iz.size,))
        else:
            inside = (q[:, :, None] + ((zs[iz][None, None, :] - ez) / c) ** 2) <= 1.0
        vol[iy[0]:iy[-1] + 1, ix[0]:ix[-1] + 1, iz[0]:iz[-1] + 1] += m * inside
    return vol.astype(np.float32)


def line_integral(ells, p0, p1):
    """Exact integral of the ellipsoid phantom along the segment direction of the
    infinite line through p0 and p1 (3-vectors, arrays of shape (..., 3))."""
    p0 = np.asarray(p0, dtype=np.float64)
    p1 = np.asarray(p1, dtype=np.float64)
    d = p1 - p0
    d = d / np.linalg.norm(d, axis=-1, keepdims=True)
    total = np.zeros(p0.shape[:-1])
    for (ex, ey, ez, a, b, c, ang, m) in ells:
        ca, sa = np.cos(ang), np.sin(ang)
        sx, sy, sz = p0[..., 0] - ex, p0[..., 1] - ey, p0[..., 2] - ez
        ox, oy = sx * ca + sy * sa, -sx * sa + sy * ca
        dx_, dy_ = d[..., 0] * ca + d[..., 1] * sa, -d[..., 0] * sa + d[..., 1] * ca
        dz_ = d[..., 2]
        ic = 0.0 if np.isinf(c) else 1.0 / (c * c)
        A = dx_ ** 2 / a ** 2 + dy_ ** 2 / b ** 2 + dz_ ** 2 * ic
        B = 2.0 * (dx_ * ox / a ** 2 + dy_ * oy / b ** 2 + dz_ * sz * ic)
        Cc = ox ** 2 / a ** 2 + oy ** 2 / b ** 2 + sz ** 2 * ic - 1.0
        disc = B * B - 4.0 * A * Cc
        chord = np.where(disc > 0, np.sqrt(np.maximum(disc, 0.0)) / np.maximum(A, 1e-300), 0.0)
        total += m * chord
    return total


# CTLAB's SheppLoganPhantom (src/Ellipsoid.h:252-317): ten ellipses on the unit
# square [-1/2, 1/2]^2 (the classic table with lengths and positions halved),
# as (centre x, centre y, semi-axis x, semi-axis y, rotation in degrees,
# intensity); the intensities add where ellipses overlap.
_SHEPP_LOGAN = (
    (0.0, 0.0, 0.345, 0.46, 0.0, 1.0),
    (0.0, -0.0092, 0.3312, 0.437, 0.0, -0.8),
    (0.11, 0.0, 0.055, 0.155, -18.0, -0.2),
    (-0.11, 0.0, 0.08, 0.205, 18.0, -0.2),
    (0.0, 0.175, 0.105, 0.125, 0.0, 0.1),
    (0.0, 0.05, 0.023, 0.023, 0.0, 0.1),
    (0.0, -0.05, 0.023, 0.023, 0.0, 0.1),
    (-0.04, -0.3025, 0.023, 0.0115, 0.0, 0.1),
    (0.0, -0.303, 0.0115, 0.0115, 0.0, 0.1),
    (0.03, -0.3025, 0.0115, 0.023, 0.0, 0.1),
)


def shepp_logan(size_mm, mu=1.0, height_mm=None):
    """CTLAB's Shepp-Logan phantom scaled to a square of side size_mm, intensities
    times mu (e.g. 0.02 /mm for water-like soft tissue), in the ellipsoid format
    of this module.  Without height_mm the ellipses are infinite cylinders along
    z (a 2-D phantom on every slice, as the reference's); with it, every
    ellipse becomes an ellipsoid of z semi-axis height_mm / 2 about z = 0."""
    c = np.inf if height_mm is None else 0.5 * float(height_mm)
    return np.array([(x * size_mm, y * size_mm, 0.0, a * size_mm, b * size_mm, c, np.deg2rad(r), m * mu)
                     for (x, y, a, b, r, m) in _SHEPP_LOGAN], dtype=np.float64)


def add_poisson_noise(prj, I0, seed):
    """counts ~ Poisson(I0 exp(-p)); returns -log(max(counts,1)/I0) as float32."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = rng.poisson(I0 * np.exp(-np.asarray(prj, dtype=np.float64)))
    return (-np.log(np.maximum(counts, 1) / I0)).astype(np.float32)
