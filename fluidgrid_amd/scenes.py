"""Seeded synthetic scenes for the SPH step.

Dam-break blocks follow the reference emitter's particle order: y-layer
outermost, then z, then x innermost (Emitter.cpp:16-42).  Each lattice point
is jittered by (2u-1)*0.01h per axis with u drawn from a SplitMix64 stream
(seed = 20261202 + config id), three draws per particle in id order.

`appmain_scene()` restates the reference application's hard-coded scene
(AppMain.cpp:78-133 with BlockEmitter, Emitter.cpp:16-42) in float32
arithmetic; it is the parity scene for the REF_XOR key mode.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

SEED_BASE = 20261202
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, offset=0):
    """The first `count` outputs (after skipping `offset`) of SplitMix64."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, count, offset=0):
    """Uniform floats in [0,1) built from the top 24 bits of SplitMix64."""
    return (splitmix64(seed, count, offset) >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))


def kernels_cache(h):
    """KernelsCache as HeterogeneousSimulator.cpp:33-39 computes it: the
    float PI constant, pow(h, n) in double, one rounding to float."""
    pi_f = float(np.float32(3.14159265))
    h = float(np.float32(h))
    kc = abi.KernelsCache()
    kc.kPoly6 = float(np.float32(315.0 / (float(np.float32(64.0 * pi_f)) * h ** 9)))
    kc.kSpikyGradient = float(np.float32(-(45.0 / (pi_f * h ** 6))))
    kc.kViscosityLaplacian = float(np.float32(45.0 / (pi_f * h ** 6)))
    return kc


def key_bits(num_buckets):
    return max(1, int(num_buckets - 1).bit_length())


@dataclass
class Scene:
    name: str
    pos: np.ndarray            # (N,4) float32, w = active flag
    vel: np.ndarray            # (N,4) float32
    fluid: abi.FluidParams
    common: abi.CommonFluidParams
    kernels: abi.KernelsCache
    key_mode: int
    steps: int
    seed: int = 0
    info: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.pos.shape[0]

    @property
    def num_buckets(self):
        return self.common.num_buckets()


# Dam-break configurations: (lattice nx, ny, nz), h, box extents, dt, steps.
CONFIGS = {
    0: dict(lattice=(32, 16, 32), h=0.025, box=(1.0, 1.0, 1.0), dt=4e-4, steps=100),
    1: dict(lattice=(128, 64, 128), h=0.01, box=(1.5, 1.5, 1.5), dt=1.5e-4, steps=200),
    2: dict(lattice=(256, 128, 256), h=0.005, box=(1.5, 1.5, 1.5), dt=7.5e-5, steps=200),
    4: dict(lattice=(512, 256, 256), h=0.003, box=(2.0, 1.0, 1.0), dt=8e-5, steps=100),
}


def _cells(extent, h):
    return int(math.ceil(extent / h - 1e-9))


def dam_break(lattice, h, box, dt, steps=100, seed=SEED_BASE, name="dam_break",
              key_mode=abi.KEY_LINEAR, jitter=0.01, stiffness_k=3.0, mu=3.5, rho0=1000.0,
              velocity_limit=600.0, gravity=-9.81):
    """Jittered block of nx*ny*nz particles at spacing h/2 from the box corner."""
    nx, ny, nz = lattice
    n = nx * ny * nz
    half = 0.5 * h
    pos = np.zeros((n, 4), np.float32)
    # particle l = (iy nz + iz) nx + ix (y layers outermost, x innermost);
    # its jitter is draws 3l..3l+2 of the stream.  Built in chunks so that
    # large lattices stay in cache; every element is computed the same way.
    chunk = 1 << 21
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        l = np.arange(a, b, dtype=np.int64)
        lat = np.stack([l % nx, l // (nx * nz), (l // nx) % nz], axis=1).astype(np.float64)
        base = half * (1.0 + lat)
        if jitter:
            u = uniform01(seed, 3 * (b - a), offset=3 * a).reshape(b - a, 3)
            base = base + (2.0 * u - 1.0) * (jitter * h)
        pos[a:b, :3] = base.astype(np.float32)
    pos[:, 3] = 1.0
    vel = np.zeros((n, 4), np.float32)

    cp = abi.CommonFluidParams()
    hf = float(np.float32(h))
    cp.cell_size = abi.f3(hf, hf, hf)
    cells = tuple(_cells(e, h) for e in box)
    cp.bucket_size = abi.i3(*cells)
    cp.keyBits = key_bits(cells[0] * cells[1] * cells[2])
    cp.kernel_radius = hf
    cp.kernel_radius_squared = float(np.float32(hf) * np.float32(hf))
    cp.pointRadius = float(np.float32(half))
    cp.pointScale = 1.0
    cp.velocity_limit = velocity_limit
    cp.volume_min = abi.f3(0.0, 0.0, 0.0)
    cp.volume_max = abi.f3(*box)
    cp.global_external_force = abi.f3(0.0, gravity, 0.0)
    cp.simulation_scale = 1.0
    cp.deltaTimeIntegration = dt
    cp.fluid_to_srb_relation = abi.i3(1, 1, 1)

    fp = abi.FluidParams()
    fp.mass = float(np.float32(rho0 * half ** 3))
    fp.restDensity = rho0
    fp.viscosity = mu
    fp.stiffness = 0.0
    fp.particle_radius = float(np.float32(half))
    fp.gasConstant = stiffness_k
    fp.max_particles = n
    return Scene(name, pos, vel, fp, cp, kernels_cache(hf), key_mode, steps, seed,
                 dict(lattice=lattice, h=h, box=box, dt=dt))


def config_scene(cfg_id, key_mode=abi.KEY_LINEAR):
    c = CONFIGS[cfg_id]
    return dam_break(c["lattice"], c["h"], c["box"], c["dt"], c["steps"], seed=SEED_BASE + cfg_id,
                     name=f"dam_break_cfg{cfg_id}", key_mode=key_mode)


def with_inflow(scene, window_min, window_max, speed=0.5, layers=4, spacing=None):
    """Adds an inflow through the x-min face (normal +x) of the box: a lattice
    window [y0, y1] x [z0, z1] at x = volume_min.x + spacing, emitted layer by
    layer at `speed` (fg_inflow_desc).  The reserve (layers x nu x nv slots)
    is appended to the particle arrays as inactive particles (w = -1, zero
    velocity) parked one per cell in the top cell layers of the box, where no
    active particle reaches them; inactive particles still take part in the
    grid, as in the reference.  Returns the scene (arrays extended) with the
    descriptor in scene.info["inflow"]."""
    cp = scene.common
    h = cp.kernel_radius
    s = float(np.float32(spacing if spacing is not None else 0.5 * h))
    (y0, z0), (y1, z1) = window_min, window_max
    nu = int(math.floor((y1 - y0) / s + 1e-9)) + 1
    nv = int(math.floor((z1 - z0) / s + 1e-9)) + 1
    reserve = layers * nu * nv
    cs = cp.cell_size.x
    bx, by, bz = cp.bucket_size.x, cp.bucket_size.y, cp.bucket_size.z
    per_layer = bx * bz
    if reserve > per_layer * by // 4:
        raise ValueError("reserve does not fit in the top quarter of the box")
    k = np.arange(reserve)
    park = np.zeros((reserve, 4), np.float32)
    park[:, 0] = cp.volume_min.x + (k % bx + 0.5) * cs
    park[:, 2] = cp.volume_min.z + ((k // bx) % bz + 0.5) * cs
    park[:, 1] = cp.volume_max.y - (k // per_layer + 0.5) * cs
    park[:, 3] = -1.0
    first = scene.n
    scene.pos = np.concatenate([scene.pos, park])
    scene.vel = np.concatenate([scene.vel, np.zeros_like(park)])
    scene.fluid.max_particles = scene.n
    d = abi.InflowDesc()
    d.origin = abi.f3(cp.volume_min.x + s, y0, z0)
    d.axisU, d.axisV, d.normal = abi.f3(0, 1, 0), abi.f3(0, 0, 1), abi.f3(1, 0, 0)
    d.spacing, d.speed = s, float(np.float32(speed))
    d.nu, d.nv, d.firstSlot, d.numLayers = nu, nv, first, layers
    scene.info["inflow"] = d
    scene.info["active"] = first
    return scene


def small_dam_break(lattice=(12, 8, 10), h=0.05, box=(1.0, 1.0, 1.0), dt=4e-4, seed=SEED_BASE + 100,
                    key_mode=abi.KEY_LINEAR):
    return dam_break(lattice, h, box, dt, 10, seed=seed, name="small_dam_break", key_mode=key_mode)


def appmain_scene():
    """The reference application's scene (AppMain.cpp:84-123) in float32."""
    f = np.float32
    cp = abi.CommonFluidParams()
    cp.volume_min = abi.f3(-20.0, -20.0, -20.0)
    cp.volume_max = abi.f3(20.0, 85.0, 40.0)
    cp.global_external_force = abi.f3(0.0, float(f(-9.8)), 0.0)
    cp.simulation_scale = float(f(0.009))
    cp.deltaTimeIntegration = float(f(0.004))
    cp.bucket_size = abi.i3(64, 64, 64)
    cp.keyBits = 24
    cp.velocity_limit = 600.0
    cp.pointScale = float(f(800.0 / math.tan(90.0 * 0.5 * float(f(3.1415)) / 180.0)))
    cp.fluid_to_srb_relation = abi.i3(1, 1, 1)

    fp = abi.FluidParams()
    fp.max_particles = 24000
    mass = f(f(128 * 1024.0) / f(24000)) * f(0.0002)
    fp.mass = float(mass)
    fp.restDensity = 1000.0
    fp.stiffness = float(f(0.2))
    fp.viscosity = 1.0
    fp.gasConstant = 1.5
    fp.particle_radius = float(f(0.3))

    dist = np.power(f(mass / f(1000.0)), f(f(1.0) / f(3.0)))
    spacing = f(f(dist * f(0.87)) / f(cp.simulation_scale))
    kr = f(f(f(2.0) * dist) * f(0.87))
    cp.kernel_radius = float(kr)
    cp.kernel_radius_squared = float(f(kr * kr))
    cp.pointRadius = float(spacing)
    pv = f(kr / f(cp.simulation_scale))
    cp.cell_size = abi.f3(float(pv), float(pv), float(pv))

    pos = block_emitter((f(-20.0), f(-20.0), f(-20.0)), (f(20.0), f(85.0), f(0.0)), spacing, fp.max_particles)
    vel = np.zeros_like(pos)
    return Scene("appmain_ref_xor", pos, vel, fp, cp, kernels_cache(float(kr)), abi.KEY_REF_XOR, 100, 0,
                 dict(spacing=float(spacing)))


def block_emitter(bmin, bmax, spacing, count):
    """BlockEmitter::reset (Emitter.cpp:16-42) with float-accumulated loops."""
    f = np.float32
    spacing = f(spacing)
    xs, zs = [], []
    x = f(bmin[0])
    while x < f(bmax[0]):
        xs.append(x)
        x = f(x + spacing)
    z = f(bmin[2])
    while z < f(bmax[2]):
        zs.append(z)
        z = f(z + spacing)
    out = np.zeros((count, 4), np.float32)
    out[:, 3] = -1.0
    used = 0
    pos_y = f(bmin[1])
    per_layer = len(xs) * len(zs)
    if per_layer == 0:
        return out
    while used < count:
        for zz in zs:
            for xx in xs:
                if used < count:
                    out[used] = (xx, pos_y, zz, 1.0)
                    used += 1
        pos_y = f(pos_y + spacing)
    return out


# ---------------------------------------------------------------------------
# Rigid bodies
# ---------------------------------------------------------------------------

def _surface_voxels(sdf, lo, hi, d):
    """Centres of the voxels (spacing d) whose cube can meet the surface
    sdf = 0, i.e. |sdf(centre)| <= (sqrt(3)/2) d."""
    axes = [np.arange(lo[k], hi[k] + 0.5 * d, d) for k in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    sel = np.abs(sdf(X, Y, Z)) <= 0.5 * np.sqrt(3.0) * d
    return np.stack([X[sel], Y[sel], Z[sel]], axis=1)


def torus_spheres(R, r, d):
    """Surface-voxel sample of a torus (axis +y, centred at the origin)."""
    ext = R + r + d

    def sdf(x, y, z):
        return np.sqrt((np.sqrt(x * x + z * z) - R) ** 2 + y * y) - r

    return _surface_voxels(sdf, (-ext, -(r + d), -ext), (ext, r + d, ext), d)


def ball_spheres(radius, d):
    """Surface-voxel sample of a sphere of the given radius centred at the origin."""
    ext = radius + d

    def sdf(x, y, z):
        return np.sqrt(x * x + y * y + z * z) - radius

    return _surface_voxels(sdf, (-ext,) * 3, (ext,) * 3, d)


def box_spheres(edge, d):
    """Surface-voxel sample of a cube of the given edge centred at the origin."""
    a = 0.5 * edge
    ext = a + d

    def sdf(x, y, z):
        q = np.stack([np.abs(x) - a, np.abs(y) - a, np.abs(z) - a])
        outside = np.sqrt((np.maximum(q, 0.0) ** 2).sum(0))
        inside = np.minimum(q.max(0), 0.0)
        return outside + inside

    return _surface_voxels(sdf, (-ext,) * 3, (ext,) * 3, d)


def cylinder_spheres(radius, height, d):
    """Surface-voxel sample of a closed cylinder (axis +y, base at y = 0,
    centred on the y axis), the static obstacle of the obstacle scenes."""
    ext = radius + d

    def sdf(x, y, z):
        dr = np.sqrt(x * x + z * z) - radius
        dy = np.abs(y - 0.5 * height) - 0.5 * height
        outside = np.sqrt(np.maximum(dr, 0.0) ** 2 + np.maximum(dy, 0.0) ** 2)
        return outside + np.minimum(np.maximum(dr, dy), 0.0)

    return _surface_voxels(sdf, (-ext, -d, -ext), (ext, height + d, ext), d)


def quat_rotate(q, v):
    """Rotate vectors v (N,3) by unit quaternion q = (x, y, z, w)."""
    u = np.asarray(q[:3], np.float64)
    w = float(q[3])
    v = np.asarray(v, np.float64)
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


@dataclass
class BodySet:
    """Dynamic rigid bodies as world-space sphere sets (body-major order)."""
    world: np.ndarray          # (M,4) float32, w = body index (M may be 0)
    offsets: np.ndarray        # (B+1,) uint32: spheres of body b are [offsets[b], offsets[b+1])
    mass: np.ndarray           # (B,) float64
    sphere_radius: float
    rb: abi.CommonRigidBodyParams
    spring: float              # SimulationParams.springForce (kSpring = -spring)
    damping: float             # SimulationParams.dampingForce (kDamping = -damping)
    fluid_radius: float        # contact radius of a fluid particle
    names: list = field(default_factory=list)
    static: np.ndarray = None  # (K,4) float32 static sphere centres (w = 0), or None

    @property
    def m(self):
        return int(self.world.shape[0])

    @property
    def b(self):
        return int(self.mass.shape[0])


def make_bodies(sets, masses, h, common, spring=2000.0, damping=50.0, names=None, buckets=(64, 64, 64),
                static_sets=None, st_buckets=(64, 64, 64)):
    """Pack sphere sets (list of (K,3) world centres) into a BodySet with the
    coupling parameters used by the coupled configurations; static_sets (list
    of (K,3) centres) become the static spheres, hashed over st_buckets.

    The coupled passes take the fluid contact radius as 2 * pointRadius
    (HeterogeneousSimulator.cpp:179, RigidBodySimulator.cpp:517), so this sets
    common.pointRadius = fluid_radius / 2 on the fluid parameters passed in."""
    if sets:
        world = np.concatenate([np.concatenate([s, np.full((len(s), 1), b)], 1)
                                for b, s in enumerate(sets)]).astype(np.float32)
    else:
        world = np.zeros((0, 4), np.float32)
    offsets = np.concatenate([[0], np.cumsum([len(s) for s in sets])]).astype(np.uint32)
    static = None
    if static_sets:
        st = np.concatenate([np.asarray(s, np.float64) for s in static_sets])
        static = np.zeros((len(st), 4), np.float32)
        static[:, :3] = st
    sphere_r = 0.5 * h
    fluid_r = 0.25 * h
    common.pointRadius = float(np.float32(0.5 * fluid_r))
    fluid_r = 2.0 * common.pointRadius
    contact = fluid_r + sphere_r + 0.003
    rb = abi.CommonRigidBodyParams()
    rb.rb_to_fluid_relation = abi.i3(1, 1, 1)
    rb.st_bucket_size = abi.i3(*st_buckets) if static is not None else abi.i3(1, 1, 1)
    rb.dn_bucket_size = abi.i3(*buckets)
    rb.st_keyBits = key_bits(rb.st_bucket_size.x * rb.st_bucket_size.y * rb.st_bucket_size.z)
    rb.velocity_limit = common.velocity_limit
    rb.volume_min = abi.Float3(*common.volume_min)
    rb.volume_max = abi.Float3(*common.volume_max)
    g = common.global_external_force.y
    m0 = float(masses[0]) if len(masses) else 0.0
    rb.global_external_force.x, rb.global_external_force.y = 0.0, float(np.float32(g * m0))
    rb.global_external_force.z, rb.global_external_force.w = 0.0, 0.0
    rb.simulation_scale = common.simulation_scale
    rb.deltaTimeIntegration = common.deltaTimeIntegration
    # The body grid cell (getCellSize) spans the whole contact distance of
    # both the fluid contact and the body-body contact (2 r + 0.001), so the
    # 27-cell stencil finds every contact.
    rb.particleRadius = float(np.float32(max(contact, 2.0 * sphere_r + 0.001) * 1.001))
    return BodySet(world, offsets, np.asarray(masses, np.float64), float(np.float32(sphere_r)), rb, spring, damping,
                   float(np.float32(fluid_r)), names or [], static)


def coupled_config_scene(cfg_id=1, seed=None):
    """Config 1 (1M dam break + one rigid torus, R=0.15, r=0.05, density 500,
    0.3 m above the fluid surface) or config 2 (8M + four 0.1 m boxes,
    density 700, seeded poses above the fluid)."""
    s = config_scene(cfg_id)
    h = s.info["h"]
    nx, ny, nz = s.info["lattice"]
    top = 0.5 * h * ny
    cx, cz = 0.5 * h * (nx + 1) / 2.0, 0.5 * h * (nz + 1) / 2.0
    if cfg_id == 1:
        R, r = 0.15, 0.05
        pts = torus_spheres(R, r, 0.5 * h) + np.array([cx, top + 0.3 + r, cz])
        mass = 500.0 * 2.0 * np.pi ** 2 * R * r * r
        bodies = make_bodies([pts], [mass], h, s.common, names=["torus"])
    elif cfg_id == 2:
        rng = np.random.default_rng(SEED_BASE + 2)
        sets, masses = [], []
        for b in range(4):
            q = rng.normal(size=4)
            q /= np.linalg.norm(q)
            c = np.array([cx + (b % 2 - 0.5) * 0.25, top + 0.15 + 0.12 * b, cz + (b // 2 - 0.5) * 0.25])
            sets.append(quat_rotate(q, box_spheres(0.1, 0.5 * h)) + c)
            masses.append(700.0 * 0.1 ** 3)
        bodies = make_bodies(sets, masses, h, s.common, names=[f"box{b}" for b in range(4)])
    else:
        raise ValueError(cfg_id)
    s.name = f"coupled_cfg{cfg_id}"
    s.info["bodies"] = bodies
    return s


def config4_scene(seed=None, num_bodies=64):
    """Config 4: the 33,554,432-particle dam break (512x256x256 lattice at
    h/2, h = 0.003, box [0,2]x[0,1]x[0,1], dt = 8e-5) with a 0.5 m/s inflow
    through the x-min face above the block, and 64 rigid bodies, half tori
    and half boxes, overall size uniform in [0.03, 0.1] m and density uniform
    in [300, 900] kg/m^3, on three layers above the fluid (seeded poses)."""
    s = config_scene(4)
    h = s.info["h"]
    rng = np.random.default_rng(SEED_BASE + 4 if seed is None else seed)
    top = float(s.pos[:, 1].max())
    sets, masses, names = [], [], []
    nx, nz = 8, 4
    for b in range(num_bodies):
        size = rng.uniform(0.03, 0.1)
        rho = rng.uniform(300.0, 900.0)
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        ix, iz, iy = b % nx, (b // nx) % nz, b // (nx * nz)
        c = np.array([0.2 + 0.225 * ix, top + 0.1 + 0.15 * iy, 0.125 + 0.25 * iz])
        if b % 2 == 0:
            R = size / 2.6
            r = 0.3 * R
            pts = torus_spheres(R, r, 0.5 * h)
            masses.append(rho * 2.0 * np.pi ** 2 * R * r * r)
            names.append(f"torus{b}")
        else:
            pts = box_spheres(size, 0.5 * h)
            masses.append(rho * size ** 3)
            names.append(f"box{b}")
        sets.append(quat_rotate(q, pts) + c)
    bodies = make_bodies(sets, masses, h, s.common, names=names, buckets=(128, 128, 128))
    s = with_inflow(s, (top + 0.05, 0.35), (top + 0.2, 0.5), speed=0.5, layers=6)
    s.name = "inflow_cfg4"
    s.info["bodies"] = bodies
    return s


# ---------------------------------------------------------------------------
# Batched ensembles
# ---------------------------------------------------------------------------

@dataclass
class Ensemble:
    """Independent dam-break scenes advanced together (scene-major);
    bodies[s] is scene s's BodySet or None."""
    scenes: list
    seed: int
    bodies: list = None

    @property
    def n(self):
        return sum(s.n for s in self.scenes)

    @property
    def num_buckets(self):
        return sum(s.num_buckets for s in self.scenes)


def ensemble(num_scenes=256, seed=SEED_BASE + 3, n_range=(4096, 65536), edge_range=(0.5, 1.5), fill=0.3,
             bodies=False, max_bodies=4):
    """Scenes with log-uniform particle counts in n_range and box edges L
    uniform in edge_range; h is chosen so that the block of N particles at
    spacing h/2 fills `fill` of the box: N (h/2)^3 = fill L^3.  The block is
    nx x ny x nz with nx = nz ~ (2N)^(1/3) (a dam-break slab); dt = 0.015 h
    (the step/h ratio of the single-scene configurations); other constants as
    in the dam-break configurations."""
    u = uniform01(seed, 2 * num_scenes).reshape(num_scenes, 2)
    lo, hi = math.log(n_range[0]), math.log(n_range[1])
    scenes = []
    for s in range(num_scenes):
        n_target = math.exp(lo + (hi - lo) * u[s, 0])
        edge = edge_range[0] + (edge_range[1] - edge_range[0]) * u[s, 1]
        nx = max(2, int(round((2.0 * n_target) ** (1.0 / 3.0))))
        ny = max(1, int(round(n_target / (nx * nx))))
        while nx * ny * nx < n_range[0]:
            ny += 1
        while ny > 1 and nx * ny * nx > n_range[1]:
            ny -= 1
        n = nx * ny * nx
        h = 2.0 * (fill * edge ** 3 / n) ** (1.0 / 3.0)
        sc = dam_break((nx, ny, nx), h, (edge, edge, edge), 0.015 * h, steps=500, seed=seed * 1000 + s,
                       name=f"ensemble_{s}")
        sc.info["scene_index"] = s
        scenes.append(sc)
    return Ensemble(scenes, seed, [_ensemble_bodies(sc, seed * 1000 + k, max_bodies) for k, sc in enumerate(scenes)]
                    if bodies else None)


def _ensemble_bodies(sc, seed, max_bodies=4):
    """0..max_bodies rigid spheres or boxes for one ensemble scene: size
    (sphere diameter or box edge) uniform in [0.12, 0.2] of the box edge,
    density uniform in [300, 900] kg/m^3, floating half-submerged at the
    fluid surface in a row along x, each touching the next (so neighbours
    are in body-body contact), at a seeded depth-wise position, boxes at
    seeded orientations; boundary spheres of radius h/2 on a 16^3-bucket
    body grid.  None for a scene drawn with no body."""
    rng = np.random.default_rng(seed)
    k = int(rng.integers(0, max_bodies + 1))
    if k == 0:
        return None
    h = sc.info["h"]
    edge = sc.info["box"][0]
    top = float(sc.pos[:, 1].max())
    sets, masses, names = [], [], []
    cz = rng.uniform(0.3 * edge, 0.7 * edge)
    x = 0.1 * edge
    for b in range(k):
        size = edge * rng.uniform(0.12, 0.2)
        rho = rng.uniform(300.0, 900.0)
        cx = x + 0.5 * size
        x += size
        y = top
        if rng.random() < 0.5:
            pts = ball_spheres(0.5 * size, 0.5 * h)
            masses.append(rho * np.pi * size ** 3 / 6.0)
            names.append("sphere")
        else:
            q = rng.normal(size=4)
            q /= np.linalg.norm(q)
            pts = quat_rotate(q, box_spheres(size, 0.5 * h))
            masses.append(rho * size ** 3)
            names.append("box")
        sets.append(pts + np.array([cx, y, cz]))
    return make_bodies(sets, masses, h, sc.common, names=names, buckets=(16, 16, 16))
