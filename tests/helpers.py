"""Shared helpers for the GPU parity tests."""
import numpy as np

from oracle import binding as O


def scene_world(scene):
    pos, idx, inst, m4, mats = scene.arrays()
    return O.flatten(pos, idx, inst, m4)


def make_ctx(scene=None):
    import snkpt
    ctx = snkpt.Context(0)
    if scene is not None:
        ctx.load_scene(scene)
        ctx.build_accel()
    return ctx


def tri_soup_ctx(pos, idx):
    """Context holding one identity-transformed instance of (pos, idx)."""
    import snkpt
    ctx = snkpt.Context(0)
    ctx.upload_mesh(pos, idx)
    ctx.set_instances(np.array([[0, 0, 0, 0, idx.size]], np.uint32), np.eye(4, dtype=np.float32).reshape(1, 16))
    m = snkpt.Material()
    m.albedo_tex_idx = m.normal_tex_idx = m.roughness_tex_idx = m.metallic_tex_idx = m.ao_tex_idx = 0xFFFFFFFF
    m.albedo[:] = [0.5, 0.5, 0.5, 1.0]
    m.roughness = 1.0
    ctx.set_materials([m])
    ctx.set_lights(None, None)
    ctx.build_accel()
    return ctx


def gpu_trace(ctx, rays_np, any_hit=False, counted=True):
    import torch
    r = torch.from_numpy(np.ascontiguousarray(rays_np, np.float32).reshape(-1, 8)).cuda()
    org = r[:, :4].contiguous()
    dr = r[:, 4:].contiguous()
    out = ctx.trace(org, dr, any_hit=any_hit, counted=counted)
    torch.cuda.synchronize()
    return [x.cpu().numpy() if x is not None else None for x in out]


def as_bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def assert_bits_equal(a, b, what):
    a, b = as_bits(a), as_bits(b)
    bad = np.nonzero(a != b)[0]
    assert bad.size == 0, f"{what}: {bad.size} of {a.size} differ, first at {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"


class ArrayScene:
    """A hand-built scene: meshes with identity transforms, one instance each."""

    def __init__(self):
        self.pos, self.nrm, self.idx, self.inst, self.mats = [], [], [], [], []

    def add(self, pos, idx, nrm, albedo, flags, roughness=1.0, metallic=0.0):
        vo = sum(len(p) for p in self.pos)
        io = sum(len(i) for i in self.idx)
        self.pos.append(np.asarray(pos, np.float32).reshape(-1, 3))
        self.nrm.append(np.asarray(nrm, np.float32).reshape(-1, 3))
        self.idx.append(np.asarray(idx, np.uint32))
        m = np.zeros(16, np.uint32)
        m[0:5] = 0xFFFFFFFF
        f = np.array([roughness, metallic, 1.0] + list(albedo) + [1.0], np.float32)
        m[5:12] = f.view(np.uint32)
        m[12] = flags
        self.mats.append(m)
        self.inst.append([0, io, vo, len(self.mats) - 1, len(idx)])

    def quad(self, o, a, b, albedo, flags, **kw):
        o, a, b = (np.asarray(x, np.float64) for x in (o, a, b))
        n = np.cross(a, b)
        n /= np.linalg.norm(n)
        self.add([o, o + a, o + a + b, o + b], [0, 1, 2, 0, 2, 3], [n] * 4, albedo, flags, **kw)

    def arrays(self):
        pos = np.concatenate(self.pos)
        idx = np.concatenate(self.idx)
        inst = np.array(self.inst, np.uint32)
        m4 = np.eye(4, dtype=np.float32).reshape(1, 16)
        mats = np.array(self.mats, np.uint32)
        nrm = np.concatenate(self.nrm)
        em = np.array([i for i, m in enumerate(self.inst) if self.mats[m[3]][12] & 1], np.uint32)
        return pos, idx, inst, m4, mats, nrm, em

    def world(self):
        pos, idx, inst, m4, mats, nrm, em = self.arrays()
        return O.flatten(pos, idx, inst, m4)

    def load_into(self, ctx, lights=None):
        import snkpt
        pos, idx, inst, m4, mats, nrm, em = self.arrays()
        ctx.upload_mesh(pos, idx, nrm=nrm)
        ctx.set_instances(inst, m4)
        ms = []
        for m in mats:
            mm = snkpt.Material.from_buffer_copy(m.tobytes())
            ms.append(mm)
        ctx.set_materials(ms)
        ctx.set_lights(lights, em)
        ctx.build_accel()


def look_at_camera(eye, target, fov=90.0, w=1, h=1):
    """A snkpt.Camera (glm lookAt / perspective ZO with the Vulkan Y flip)."""
    import snkpt
    eye, target = np.asarray(eye, np.float64), np.asarray(target, np.float64)
    up = np.array([0, 1.0, 0]) if abs(np.dot(target - eye, [0, 1, 0])) < 0.999 * np.linalg.norm(target - eye) \
        else np.array([0, 0, -1.0])
    f = (target - eye) / np.linalg.norm(target - eye)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, u, -f
    view[0, 3], view[1, 3], view[2, 3] = -s @ eye, -u @ eye, f @ eye
    zn, zf = 0.1, 2500.0
    th = np.tan(np.radians(fov * 0.5) / 2)
    proj = np.zeros((4, 4))
    proj[0, 0] = 1 / ((w / h) * th)
    proj[1, 1] = -1 / th
    proj[2, 2] = zf / (zn - zf)
    proj[3, 2] = -1
    proj[2, 3] = -(zf * zn) / (zf - zn)
    cam = snkpt.Camera()
    cam.view[:] = view.T.reshape(-1).astype(np.float32).tolist()
    cam.proj[:] = proj.T.reshape(-1).astype(np.float32).tolist()
    cam.proj_view[:] = (proj @ view).T.reshape(-1).astype(np.float32).tolist()
    cam.cam_pos[:] = [float(eye[0]), float(eye[1]), float(eye[2]), 1.0]
    cam.cam_forward[:] = [float(f[0]), float(f[1]), float(f[2]), 1.0]
    return cam


def polygon_irradiance(x, n, verts, L):
    """Irradiance at x (normal n) from a Lambertian polygon of radiance L
    (Lambert's formula: E = L/2 * sum_i beta_i * n.g_i)."""
    x, n = np.asarray(x, np.float64), np.asarray(n, np.float64)
    r = [np.asarray(v, np.float64) - x for v in verts]
    e = 0.0
    for i in range(len(r)):
        a, b = r[i], r[(i + 1) % len(r)]
        beta = np.arccos(np.clip(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)), -1, 1))
        g = np.cross(a, b)
        g /= np.linalg.norm(g)
        e += beta * (n @ g)
    return abs(0.5 * L * e)


def floor_light_scene(L=10.0, albedo=0.5, side=1.0, height=2.0):
    """Lambertian floor (y = 0) lit by a square emitter facing down at `height`."""
    sc = ArrayScene()
    sc.quad([-10, 0, -10], [0, 0, 20], [20, 0, 0], [albedo] * 3, 2)
    h = side / 2
    sc.quad([-h, height, -h], [side, 0, 0], [0, 0, side], [L] * 3, 1)
    verts = [[-h, height, -h], [h, height, -h], [h, height, h], [-h, height, h]]
    return sc, verts


class ArraysScene:
    """A scene given as the path tracer's input tuple (pos, idx, inst5, mat4s,
    mats16, nrm, emissive_instances), loadable into a context."""

    def __init__(self, arrs):
        self.arrs = tuple(np.ascontiguousarray(a) for a in arrs)

    def arrays(self):
        return self.arrs

    def world(self):
        pos, idx, inst, m4 = self.arrs[:4]
        return O.flatten(pos, idx, inst, m4)

    def load_into(self, ctx, lights=None):
        import snkpt
        pos, idx, inst, m4, mats, nrm, em = self.arrs
        ctx.upload_mesh(pos, idx, nrm=nrm)
        ctx.set_instances(inst, m4)
        ctx.set_materials([snkpt.Material.from_buffer_copy(m.tobytes()) for m in mats])
        ctx.set_lights(lights, em)
        ctx.build_accel()


def ggx_room(seed=5):
    """The config-1 room with config 2's material distribution: every
    non-emissive material GGX (flags 0), roughness U[0.05, 1], metallic
    Bernoulli(0.3), albedo U[0.1, 0.9]^3; the first two are forced to a
    rough dielectric and a glossy metal so both lobes always appear."""
    import snkpt
    scene = snkpt.SceneData(1)
    pos, idx, inst, m4, mats, nrm, em = O.scene_arrays(scene)
    mats = mats.copy()
    rng = np.random.default_rng(seed)
    k = 0
    for m in mats:
        if m[12] & 1:
            continue
        f = m[5:12].view(np.float32)
        f[0] = rng.uniform(0.05, 1.0)
        f[1] = float(rng.random() < 0.3)
        f[3:6] = rng.uniform(0.1, 0.9, 3)
        if k == 0:
            f[0], f[1] = 0.6, 0.0
        elif k == 1:
            f[0], f[1] = 0.15, 1.0
        m[12] = 0
        k += 1
    cam_scene = scene
    return ArraysScene((pos.copy(), idx.copy(), inst.copy(), m4.copy(), mats, nrm, em)), cam_scene
