"""Seeded scene generators for the benchmark and parity workloads.

Every generator takes a uint64 seed and draws from SplitMix64, so a scene is
reproducible bit for bit from (seed, instance id) on any host.  Scenes are
returned as ABI structs (abi.efl_settings / efl_source / efl_obstacle /
efl_force) that both the HIP library and the CPU oracle accept.

World coordinates follow the reference: pixels, origin at the grid centre,
y up (update_solid.wgsl:80-85, update_fluid_source.wgsl:34-38).  Obstacle
records follow construct_rigid_body_buffer_for_gpu (obstacle.rs:89-160);
parry's ShapeType ids are 0 ball, 1 cuboid, 2 capsule, 4 triangle.
"""
import math

from . import abi

SEED_BASE = 0xE0F1D000
MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea & Flood 2014); stream = instance id."""

    def __init__(self, seed, stream=0):
        self.state = (seed + 0x9E3779B97F4A7C15 * (stream + 1)) & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))


def projection(kind="gauss_seidel", gs_iterations=20, pre=2, post=2, coarsest=3,
               max_cycles=1, rel_tol=0.0, check_every=1):
    kinds = {"jacobi": abi.PROJECTION_JACOBI, "gauss_seidel": abi.PROJECTION_GAUSS_SEIDEL,
             "multigrid": abi.PROJECTION_MULTIGRID, "mgpcg": abi.PROJECTION_MGPCG,
             "learned": abi.PROJECTION_LEARNED}
    p = abi.efl_projection()
    p.kind = kinds[kind]
    p.gs_iterations = gs_iterations
    p.mg_pre_iterations = pre
    p.mg_post_iterations = post
    p.mg_coarsest_iterations = coarsest
    p.max_cycles = max_cycles
    p.rel_tol = rel_tol
    p.check_every = check_every
    return p


def reinit(kind="fast_iterative", fim_iterations=10):
    r = abi.efl_reinit()
    r.kind = abi.REINIT_FAST_ITERATIVE if kind == "fast_iterative" else abi.REINIT_JUMP_FLOODING
    r.fim_iterations = fim_iterations
    return r


def settings(width, height, rho=99.7, gravity=(0.0, 9.8), length_unit=10.0, proj=None, rein=None,
             transform=None):
    s = abi.efl_settings()
    s.width = width
    s.height = height
    s.rho = rho
    s.gravity[0], s.gravity[1] = gravity
    s.length_unit = length_unit
    t = transform if transform is not None else identity()
    for i in range(16):
        s.transform[i] = t[i]
    s.projection = proj if proj is not None else projection()
    s.reinit = rein if rein is not None else reinit()
    return s


def identity():
    return [1.0 if i % 5 == 0 else 0.0 for i in range(16)]


def source(center, shape="aabb", half_size=(1.0, 1.0), radius=1.0, velocity=(0.0, 0.0), sink=False):
    s = abi.efl_source()
    s.center[0], s.center[1] = center
    if shape == "aabb":
        s.shape = 1
        s.data[0], s.data[1] = half_size
    else:
        s.shape = 0
        s.data[0], s.data[1] = radius, 0.0
    s.velocity[0], s.velocity[1] = velocity
    s.mode = 1 if sink else 0
    return s


def force(f, position):
    e = abi.efl_force()
    e.force[0], e.force[1] = f
    e.position[0], e.position[1] = position
    return e


def rigid_transform(angle, tx, ty):
    """Column-major Mat4 of a 2D rotation by `angle` then translation."""
    c, s = math.cos(angle), math.sin(angle)
    m = [c, s, 0.0, 0.0, -s, c, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, tx, ty, 0.0, 1.0]
    inv_tx = -(c * tx + s * ty)
    inv_ty = -(-s * tx + c * ty)
    inv = [c, -s, 0.0, 0.0, s, c, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, inv_tx, inv_ty, 0.0, 1.0]
    return m, inv


def obstacle(entity_id, shape, values, angle, pos, lin_vel=(0.0, 0.0), ang_vel=0.0):
    o = abi.efl_obstacle()
    o.entity_id = entity_id
    o.shape = shape
    for i in range(6):
        o.values[i] = values[i] if i < len(values) else 0.0
    m, inv = rigid_transform(angle, pos[0], pos[1])
    for i in range(16):
        o.transform[i] = m[i]
        o.inverse_transform[i] = inv[i]
    o.linear_velocity[0], o.linear_velocity[1] = lin_vel
    o.angular_velocity = ang_vel
    return o


def random_body(rng, kind):
    """(shape id, values, bounding radius, area) for one body of `kind`."""
    if kind == "circle":
        r = rng.uniform(8.0, 40.0)
        return 0, [r], r, math.pi * r * r
    if kind == "rectangle":
        a, b = rng.uniform(16.0, 80.0), rng.uniform(16.0, 80.0)
        return 1, [0.5 * a, 0.5 * b], 0.5 * math.hypot(a, b), a * b
    if kind == "capsule":
        r = rng.uniform(6.0, 20.0)
        seg = rng.uniform(20.0, 60.0)
        return 2, [0.0, -0.5 * seg, 0.0, 0.5 * seg, r], 0.5 * seg + r, math.pi * r * r + 2 * r * seg
    # triangle with edges in [20, 70], rejection-sampled, centroid at the origin
    while True:
        a, b, c = (rng.uniform(20.0, 70.0) for _ in range(3))
        if a + b > c and a + c > b and b + c > a:
            break
    # vertices: P0 = (0, 0), P1 = (c, 0), P2 from the law of cosines
    x2 = (b * b + c * c - a * a) / (2 * c)
    y2 = math.sqrt(max(b * b - x2 * x2, 0.0))
    cx, cy = (0.0 + c + x2) / 3.0, (0.0 + 0.0 + y2) / 3.0
    pts = [(-cx, -cy), (c - cx, -cy), (x2 - cx, y2 - cy)]
    rad = max(math.hypot(px, py) for px, py in pts)
    area = 0.5 * c * y2
    return 4, [pts[0][0], pts[0][1], pts[1][0], pts[1][1], pts[2][0], pts[2][1]], rad, area


def shape_moments(shape, values, h=0.25):
    """(area, second moment of area about the local origin) of a body shape,
    by midpoint sampling of its inside on an h-spaced lattice (cells^2,
    cells^4).  Shapes are centred on their local origin."""
    import numpy as np
    v = list(values) + [0.0] * (6 - len(values))
    if shape == 0:
        R = v[0]
    elif shape == 1:
        R = math.hypot(v[0], v[1])
    elif shape == 2:
        R = max(math.hypot(v[0], v[1]), math.hypot(v[2], v[3])) + v[4]
    else:
        R = max(math.hypot(v[0], v[1]), math.hypot(v[2], v[3]), math.hypot(v[4], v[5]))
    t = np.arange(-R - h, R + h, h) + 0.5 * h
    X, Y = np.meshgrid(t, t)
    if shape == 0:
        inside = X * X + Y * Y < v[0] ** 2
    elif shape == 1:
        inside = (np.abs(X) < v[0]) & (np.abs(Y) < v[1])
    elif shape == 2:
        ax, ay, bx, by, r = v[:5]
        abx, aby = bx - ax, by - ay
        tt = np.clip(((X - ax) * abx + (Y - ay) * aby) / (abx * abx + aby * aby), 0, 1)
        inside = np.hypot(X - (ax + tt * abx), Y - (ay + tt * aby)) < r
    else:
        p = [(v[0], v[1]), (v[2], v[3]), (v[4], v[5])]
        s = [np.sign((Y - p[i][1]) * (p[(i + 1) % 3][0] - p[i][0]) - (X - p[i][0]) * (p[(i + 1) % 3][1] - p[i][1]))
             for i in range(3)]
        inside = ((s[0] > 0) & (s[1] > 0) & (s[2] > 0)) | ((s[0] < 0) & (s[1] < 0) & (s[2] < 0))
    area = inside.sum() * h * h
    inertia = ((X * X + Y * Y) * inside).sum() * h * h
    return float(area), float(inertia)


def bodies_for(obstacles, densities, dynamic=True):
    """efl_body records (density relative units, area, inertia) for obstacles."""
    out = []
    for o, d in zip(obstacles, densities):
        b = abi.efl_body()
        b.density = d
        b.area, b.inertia = shape_moments(o.shape, list(o.values))
        b.dynamic = 1 if dynamic else 0
        out.append(b)
    return out


def random_bodies(rng, width, height, counts, rho, clearance=4.0, max_tries=2000):
    """Non-overlapping bodies (bounding circles) inside the walls.

    counts: {"circle": n, "rectangle": n, "capsule": n, "triangle": n}.
    Returns (obstacles, densities)."""
    placed = []
    obstacles = []
    densities = []
    eid = 0
    for kind in ("circle", "rectangle", "capsule", "triangle"):
        for _ in range(counts.get(kind, 0)):
            shape, values, rad, _area = random_body(rng, kind)
            angle = rng.uniform(0.0, 2 * math.pi)
            dens = rng.uniform(0.5, 2.0) * rho
            for _try in range(max_tries):
                m = rad + clearance + 2.0
                if 2 * m >= min(width, height):
                    break
                x = rng.uniform(-0.5 * width + m, 0.5 * width - m)
                y = rng.uniform(-0.5 * height + m, 0.5 * height - m)
                if all(math.hypot(x - px, y - py) > rad + pr + clearance for px, py, pr in placed):
                    placed.append((x, y, rad))
                    obstacles.append(obstacle(eid, shape, values, angle, (x, y)))
                    densities.append(dens)
                    eid += 1
                    break
    return obstacles, densities


def dam_break(size=256, seed=SEED_BASE + 1):
    """Config 1: dam break on a size^2 grid, GS(100) at weight 1.9, FIM(10).

    Fluid is an AABB with half size 0.5*(1.0, 0.6)*size centred at
    (0, -0.2)*size (README.md:45-62, examples/water_surface.rs:52-76)."""
    s = settings(size, size, proj=projection("gauss_seidel", gs_iterations=100))
    src = [source((0.0, -0.2 * size), "aabb", half_size=(0.5 * size, 0.3 * size))]
    return dict(settings=s, init_sources=src, sources=[], obstacles=[], dt=1.0 / 60.0, seed=seed)


def coupled_tank(size=1024, seed=SEED_BASE + 2, proj=None, rel_tol=1e-5, n_each=4):
    """Config 2: size^2 grid, fluid filling the lower 55%, 4 bodies of each
    shape at seeded positions, multigrid to ||r||/||b|| <= rel_tol."""
    rng = SplitMix64(seed, 0)
    if proj is None:
        proj = projection("multigrid", max_cycles=100, rel_tol=rel_tol, check_every=1)
    s = settings(size, size, proj=proj)
    src = [source((0.0, -0.225 * size), "aabb", half_size=(0.5 * size, 0.275 * size))]
    obs, dens = random_bodies(rng, size, size, {k: n_each for k in ("circle", "rectangle", "capsule", "triangle")},
                              rho=99.7)
    return dict(settings=s, init_sources=src, sources=[], obstacles=obs, densities=dens,
                dt=1.0 / 60.0, seed=seed)


def batch_instance(rng, proj=None, size=None):
    """One instance of config 3: side lengths log-uniform in [128, 1024]
    rounded to a multiple of 32 (or size x size), random gravity direction,
    0-8 bodies and a fluid AABB covering 30-70% of the area."""
    def side():
        v = math.exp(rng.uniform(math.log(128.0), math.log(1024.0)))
        return int(max(128, min(1024, round(v / 32.0) * 32)))
    W, H = side(), side()
    if size is not None:
        W = H = size
    ang = rng.uniform(0.0, 2 * math.pi)
    g = (9.8 * math.cos(ang), 9.8 * math.sin(ang))
    frac = rng.uniform(0.3, 0.7)
    fw = W * math.sqrt(frac) * rng.uniform(0.8, 1.0)
    fh = min(H * 0.98, frac * W * H / fw)
    cx = rng.uniform(-0.5 * W + 0.5 * fw, 0.5 * W - 0.5 * fw)
    cy = rng.uniform(-0.5 * H + 0.5 * fh, 0.5 * H - 0.5 * fh)
    nb = int(rng.uniform(0, 9))
    counts = {"circle": 0, "rectangle": 0, "capsule": 0, "triangle": 0}
    kinds = list(counts)
    for i in range(nb):
        counts[kinds[i % 4]] += 1
    obs, dens = random_bodies(rng, W, H, counts, rho=99.7)
    if proj is None:
        proj = projection("multigrid", max_cycles=1)
    s = settings(W, H, gravity=g, proj=proj)
    src = [source((cx, cy), "aabb", half_size=(0.5 * fw, 0.5 * fh))]
    return dict(settings=s, init_sources=src, sources=[], obstacles=obs, densities=dens, dt=1.0 / 60.0)


def batch(n=64, seed=SEED_BASE + 3, proj=None):
    """Config 3: n independent instances of uneven size."""
    return [batch_instance(SplitMix64(seed, i), proj=proj) for i in range(n)]


def droplet_levelset(rng, width, height, n_drops, r_lo=4.0, r_hi=32.0, pool=0.2):
    """Air level set of a pool filling the bottom `pool` of the grid plus
    n_drops circular droplets (radius U[r_lo, r_hi]) above it, in grid cells
    (row 0 = top, negative = fluid); for efl_init_levelset, which takes
    initial fluid beyond the 16-source cap.  Returns a list of rows of floats
    (numpy.float32 array when numpy is importable)."""
    import numpy as np
    ys = np.arange(height, dtype=np.float32)[:, None]
    xs = np.arange(width, dtype=np.float32)[None, :]
    surface = (1.0 - pool) * height
    ls = np.broadcast_to(np.float32(surface) - ys, (height, width)).copy()
    top = max(r_hi + 1.0, surface - r_hi - 2.0)
    for _ in range(n_drops):
        r = rng.uniform(r_lo, r_hi)
        cx = rng.uniform(r + 1.0, width - r - 1.0)
        cy = rng.uniform(r + 1.0, top)
        x0, x1 = int(max(0, cx - r - 2)), int(min(width, cx + r + 3))
        y0, y1 = int(max(0, cy - r - 2)), int(min(height, cy + r + 3))
        d = np.hypot(xs[:, x0:x1] - np.float32(cx), ys[y0:y1, :] - np.float32(cy)) - np.float32(r)
        ls[y0:y1, x0:x1] = np.minimum(ls[y0:y1, x0:x1], d.astype(np.float32))
    return ls


def random_forces(rng, width, height, n, f_max=50.0):
    """n LocalForces: magnitude U[0, f_max] m/s^2 in a uniform direction at
    uniform positions (pixels, as ForceToFluid takes them)."""
    out = []
    for _ in range(n):
        a = rng.uniform(0.0, 2 * math.pi)
        m = rng.uniform(0.0, f_max)
        out.append(force((m * math.cos(a), m * math.sin(a)), (rng.uniform(0, width), rng.uniform(0, height))))
    return out


def combined(size=4096, seed=SEED_BASE + 5, n_drops=200, n_each=6, n_forces=64, proj=None, rel_tol=1e-5):
    """Config 5: size^2 grid with a pool in the bottom 20% and n_drops
    droplets above it (an initial level set, not sources), 4 * n_each mixed
    bodies and n_forces local forces (resampled by the caller every 30
    steps); MG-preconditioned CG to rel_tol."""
    rng = SplitMix64(seed, 0)
    if proj is None:
        proj = projection("mgpcg", max_cycles=200, rel_tol=rel_tol, check_every=1)
    s = settings(size, size, proj=proj)
    ls = droplet_levelset(rng, size, size, n_drops)
    obs, dens = random_bodies(rng, size, size, {k: n_each for k in ("circle", "rectangle", "capsule", "triangle")},
                              rho=99.7)
    frc = random_forces(rng, size, size, n_forces)
    return dict(settings=s, init_sources=[], sources=[], levelset=ls, obstacles=obs, densities=dens, forces=frc,
                rng=rng, dt=1.0 / 60.0, seed=seed)
