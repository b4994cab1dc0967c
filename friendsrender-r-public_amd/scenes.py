"""Seeded procedural scenes for the SPPM workloads.

The reference only loads OBJ/MTL files (Scene.cpp:13-218).  These generators
produce the same facet records that loader would (positions, per-vertex
normals, UVs, material id per facet) plus the attribute table a
SceneMaterialManager subclass would fill (main.cpp:14-114 pattern), and
RGBA32F textures exactly as TextureUtility::LoadTextureFromFile stores 8-bit
images (byte / 255.0f, alpha 1 for RGB images; TextureUtility.cpp:12-35).

Workloads:
  box_scene()      Cornell box, 5 walls x 2 triangles, diffuse + perfect
                   specular sphere (40x20 UV spheres), 0.25^2 ceiling light
                   with emission 15, untextured.
  full_scene()     the same box plus a 200,000-triangle geodesic sphere
                   displaced by 3-octave value noise (amplitude 0.05), GGX
                   conductor and GGX dielectric spheres, a light with a
                   256x256 RGB8 8x8 checker (5/20) emission texture and
                   512x512 albedo / roughness / normal noise textures.
"""
import numpy as np

from sppm_amd import (CAM_PINHOLE, CAM_THINLENS, MAT_BIMETAL, MAT_DIFFUSE, MAT_DIFFUSE_TRANSMIT, MAT_GGX_GLASS,
                      MAT_GLOSS_TRANSMIT, MAT_METAL, MAT_ROUGH_CONDUCTOR, SAMPLER_UNIFORM)


class Scene:
    ibl = None  # texture id of an environment map (SPHERE mapping), or None

    def __init__(self, pos, nrm, uv, attr, attrs, textures):
        self.pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 9)
        self.nrm = np.ascontiguousarray(nrm, np.float32).reshape(-1, 9)
        self.uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 9)
        self.attr = np.ascontiguousarray(attr, np.int32)
        self.attrs = attrs
        self.textures = [np.ascontiguousarray(t, np.float32) for t in textures]

    @property
    def nfacets(self):
        return self.pos.shape[0]

    def diagonal(self):
        p = self.pos.reshape(-1, 3)
        return float(np.linalg.norm(p.max(0) - p.min(0)))


class _Builder:
    def __init__(self):
        self.pos, self.nrm, self.uv, self.attr = [], [], [], []

    def tri(self, p, n, t, a):
        self.pos.append(np.asarray(p, np.float32).reshape(9))
        self.nrm.append(np.asarray(n, np.float32).reshape(9))
        self.uv.append(np.asarray(t, np.float32).reshape(9))
        self.attr.append(a)

    def quad(self, p0, p1, p2, p3, n, a, uvs=((0, 0), (1, 0), (1, 1), (0, 1))):
        """Two triangles (p0,p1,p2), (p0,p2,p3) with a constant normal."""
        t = [(u, v, 0.0) for u, v in uvs]
        self.tri([p0, p1, p2], [n, n, n], [t[0], t[1], t[2]], a)
        self.tri([p0, p2, p3], [n, n, n], [t[0], t[2], t[3]], a)

    def mesh(self, V, N, T, F, a):
        F = np.asarray(F)
        self.pos.extend(V[F].reshape(-1, 9))
        self.nrm.extend(N[F].reshape(-1, 9))
        self.uv.extend(T[F].reshape(-1, 9))
        self.attr.extend([a] * len(F))

    def scene(self, attrs, textures=()):
        return Scene(np.array(self.pos), np.array(self.nrm), np.array(self.uv), np.array(self.attr, np.int32),
                     attrs, list(textures))


def uv_sphere(center, r, nseg=40, nring=20):
    """UV sphere: nseg segments, nring rings, fans at the poles (no degenerate triangles)."""
    c = np.asarray(center, np.float64)
    V, N, T = [], [], []
    for j in range(nring + 1):
        th = np.pi * j / nring
        for i in range(nseg + 1):
            ph = 2 * np.pi * i / nseg
            n = np.array([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)])
            V.append(c + r * n)
            N.append(n)
            T.append((i / nseg, j / nring, 0.0))
    V, N, T = np.array(V, np.float32), np.array(N, np.float32), np.array(T, np.float32)
    F = []
    row = nseg + 1
    for j in range(nring):
        for i in range(nseg):
            a, b = j * row + i, j * row + i + 1
            c_, d = (j + 1) * row + i, (j + 1) * row + i + 1
            if j != 0:
                F.append((a, c_, b))
            if j != nring - 1:
                F.append((b, c_, d))
    return V, N, T, np.array(F)


def box_walls(b, white, red, green):
    """Box [0,1]^3 open at z=1; inward normals (Scene.cpp keeps OBJ normals)."""
    b.quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), white)      # floor y=0
    b.quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0), (0, -1, 0), white)     # ceiling y=1
    b.quad((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), (0, 0, 1), white)      # back z=0
    b.quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), (1, 0, 0), red)        # left x=0
    b.quad((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), (-1, 0, 0), green)     # right x=1


def light_quad(b, a, size=0.25, y=0.999, cx=0.5, cz=0.5):
    h = size / 2
    b.quad((cx - h, y, cz - h), (cx - h, y, cz + h), (cx + h, y, cz + h), (cx + h, y, cz - h), (0, -1, 0), a)


def box_scene(with_spheres=True):
    """Plain SPPM box workload (diffuse + specular spheres)."""
    attrs = [
        dict(name="white", kind=MAT_DIFFUSE, albedo=(0.73, 0.73, 0.73), is_sppm=1, if_opaque=1),       # 0
        dict(name="red", kind=MAT_DIFFUSE, albedo=(0.63, 0.06, 0.05), is_sppm=1, if_opaque=1),         # 1
        dict(name="green", kind=MAT_DIFFUSE, albedo=(0.14, 0.45, 0.09), is_sppm=1, if_opaque=1),       # 2
        dict(name="light", kind=MAT_DIFFUSE, albedo=(0.78, 0.78, 0.78), radiance=(15.0, 15.0, 15.0), is_sppm=1,
             if_opaque=1),                                                                            # 3 light
        dict(name="diffuse_sphere", kind=MAT_DIFFUSE, albedo=(0.75, 0.75, 0.75), is_sppm=1, if_opaque=1),  # 4
        dict(name="mirror_sphere", kind=MAT_METAL, albedo=(0.95, 0.95, 0.95), if_delta=1, is_sppm=0, if_opaque=1),  # 5
    ]
    b = _Builder()
    box_walls(b, 0, 1, 2)
    light_quad(b, 3)
    if with_spheres:
        b.mesh(*uv_sphere((0.3, 0.2, 0.4), 0.2), a=4)
        b.mesh(*uv_sphere((0.72, 0.2, 0.62), 0.2), a=5)
    return b.scene(attrs)


def pinhole_camera(width, height, sampler=SAMPLER_UNIFORM, sampler_id=0):
    return dict(type=CAM_PINHOLE, width=width, height=height, pos=(0.5, 0.5, 2.2), target=(0.5, 0.5, 0.5),
                fov_deg=45.0, F=1.0, dist_to_lens=1.0, aspect=width / height, sampler=sampler,
                sampler_id=sampler_id, sampler_nsample=1)


def thinlens_camera(width, height, aperture=0.02, focus=2.0, fov_deg=45.0, sampler_id=0):
    """ThinLensCamera: lens size == image-plane size (Camera.h:40-43), so the
    aperture fixes DistanceToLens; F from the focus distance B = A*F/(A-F)."""
    half = np.float32(np.pi) * np.float32(fov_deg) / np.float32(360.0)
    A = float(aperture / (2.0 * np.tan(float(half)) * (width / height)))
    F = A * focus / (A + focus)
    return dict(type=CAM_THINLENS, width=width, height=height, pos=(0.5, 0.5, 2.5), target=(0.5, 0.5, 0.5),
                fov_deg=fov_deg, F=F, dist_to_lens=A, aspect=width / height, sampler=SAMPLER_UNIFORM,
                sampler_id=sampler_id, sampler_nsample=1)


def aperture_image(width, height, kind="hexagon", seed=0):
    """A thin-lens diaphragm image, (height, width, 4) float32 with red 0 where
    the aperture is open and 1 where it is closed (ThinLensCamera weights each
    pixel by 1 - red): "hexagon", a hexagonal opening with a soft edge, or
    "noise", random red with a few fully open pixels.  Values are multiples
    of 1/255, so an 8-bit PPM of the image loads back to the same floats."""
    rng = np.random.default_rng(seed)
    if kind == "hexagon":
        y, x = np.mgrid[0:height, 0:width]
        px = (x + 0.5) / width * 2.0 - 1.0
        py = (y + 0.5) / height * 2.0 - 1.0
        r = np.max([np.abs(px * np.cos(a) + py * np.sin(a)) for a in (0.0, np.pi / 3, 2 * np.pi / 3)], axis=0)
        red = np.clip((r - 0.6) / 0.2, 0.0, 1.0)
    elif kind == "noise":
        red = rng.random((height, width))
        red[rng.random((height, width)) < 0.1] = 0.0
    else:
        raise ValueError(kind)
    img = np.zeros((height, width, 4), np.float32)
    img[..., 0] = np.round(red * 255.0).astype(np.float32) / np.float32(255.0)
    img[..., 3] = 1.0
    return img


def write_ppm(img, path):
    """8-bit binary PPM (P6) of the RGB channels of a (height, width, 4) image in [0, 1]."""
    h, w = img.shape[:2]
    with open(path, "wb") as f:
        f.write(b"P6\n%d %d\n255\n" % (w, h))
        f.write(np.round(np.clip(img[..., :3], 0.0, 1.0) * 255.0).astype(np.uint8).tobytes())


# ---- procedural textures ----------------------------------------------------
def _value_noise(n, seed, octaves=3, base=8):
    rng = np.random.default_rng(seed)
    out = np.zeros((n, n), np.float64)
    amp, tot = 1.0, 0.0
    for o in range(octaves):
        g = base * (2 ** o)
        lat = rng.random((g + 1, g + 1))
        x = np.linspace(0, g, n, endpoint=False)
        i0 = np.floor(x).astype(int)
        f = x - i0
        f = f * f * (3 - 2 * f)
        a = lat[i0][:, i0]
        b_ = lat[i0][:, i0 + 1]
        c = lat[i0 + 1][:, i0]
        d = lat[i0 + 1][:, i0 + 1]
        fx, fy = f[None, :], f[:, None]
        out += amp * ((a * (1 - fx) + b_ * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy)
        tot += amp
        amp *= 0.5
    return out / tot


def rgb8_to_rgba32f(img8):
    """Texture2D storage of an 8-bit RGB image: byte / 255.0f, alpha 1."""
    h, w = img8.shape[:2]
    out = np.ones((h, w, 4), np.float32)
    out[..., :3] = img8.astype(np.float32) / np.float32(255.0)
    return out


def checker_texture(n=256, cells=8, lo=5, hi=20):
    idx = (np.arange(n) * cells // n)
    chk = ((idx[:, None] + idx[None, :]) % 2).astype(np.uint8)
    v = np.where(chk == 1, hi, lo).astype(np.uint8)
    return rgb8_to_rgba32f(np.repeat(v[..., None], 3, axis=2))


def noise_textures(n=512, seed=7):
    alb = _value_noise(n, seed)
    albedo8 = np.clip(np.stack([0.35 + 0.5 * alb, 0.3 + 0.45 * alb, 0.25 + 0.4 * alb], -1) * 255, 0, 255).astype(np.uint8)
    rough8 = np.clip(_value_noise(n, seed + 1) * 255, 0, 255).astype(np.uint8)
    hgt = _value_noise(n, seed + 2, base=16)
    gy, gx = np.gradient(hgt * 8.0)
    nrm = np.stack([-gx, -gy, np.ones_like(gx)], -1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    normal8 = np.clip((nrm * 0.5 + 0.5) * 255, 0, 255).astype(np.uint8)
    return (rgb8_to_rgba32f(albedo8), rgb8_to_rgba32f(np.repeat(rough8[..., None], 3, axis=2)),
            rgb8_to_rgba32f(normal8))


def geodesic_sphere(center, r, freq, seed, amp=0.05):
    """Icosahedron with each face split into freq^2 triangles (20*freq^2
    triangles), projected to the sphere and displaced radially by 3-octave
    value noise on the sphere (amplitude amp)."""
    t = (1 + 5 ** 0.5) / 2
    iv = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                   (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], np.float64)
    iv /= np.linalg.norm(iv, axis=1, keepdims=True)
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    rng = np.random.default_rng(seed)
    octs = [(rng.normal(size=(3,)) * 0.0 + rng.random(3) * 10, 2.0 ** o) for o in range(3)]
    lat = rng.random((3, 16, 16, 16))

    def noise(p):
        val = np.zeros(len(p))
        amp_, tot = 1.0, 0.0
        for o, (off, fr) in enumerate(octs):
            q = (p * fr * 2.0 + off) % 16
            i0 = np.floor(q).astype(int)
            f = q - i0
            f = f * f * (3 - 2 * f)
            i1 = (i0 + 1) % 16
            L = lat[o]
            acc = 0
            for dx in (0, 1):
                for dy in (0, 1):
                    for dz in (0, 1):
                        ix = i1[:, 0] if dx else i0[:, 0]
                        iy = i1[:, 1] if dy else i0[:, 1]
                        iz = i1[:, 2] if dz else i0[:, 2]
                        w = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * \
                            (f[:, 2] if dz else 1 - f[:, 2])
                        acc = acc + w * L[ix, iy, iz]
            val += amp_ * acc
            tot += amp_
            amp_ *= 0.5
        return val / tot - 0.5

    tris = []
    for (a, b, c) in faces:
        A, B, Cc = iv[a], iv[b], iv[c]
        grid = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                p = A + (B - A) * (i / freq) + (Cc - A) * (j / freq)
                grid[(i, j)] = p / np.linalg.norm(p)
        for i in range(freq):
            for j in range(freq - i):
                tris.append((grid[(i, j)], grid[(i + 1, j)], grid[(i, j + 1)]))
                if j < freq - i - 1:
                    tris.append((grid[(i + 1, j)], grid[(i + 1, j + 1)], grid[(i, j + 1)]))
    P = np.array(tris, np.float64)  # (T, 3, 3) unit directions
    flat = P.reshape(-1, 3)
    disp = 1.0 + amp / r * noise(flat) * 2.0
    Vd = flat * (r * disp)[:, None]
    V = (Vd + np.asarray(center)).reshape(-1, 3, 3)
    # face normals used as vertex normals of each facet (Scene.cpp:124 rule)
    e1, e2 = V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]
    fn = np.cross(e1, e2)
    fn /= np.linalg.norm(fn, axis=1, keepdims=True) + 1e-30
    Nn = np.repeat(fn[:, None, :], 3, axis=1)
    uv = np.zeros_like(V)
    uv[..., 0] = 0.5 + np.arctan2(flat[:, 2], flat[:, 0]).reshape(-1, 3) / (2 * np.pi)
    uv[..., 1] = 0.5 - np.arcsin(np.clip(flat[:, 1], -1, 1)).reshape(-1, 3) / np.pi
    return V.astype(np.float32), Nn.astype(np.float32), uv.astype(np.float32)


def full_scene(freq=100, seed=1234, tex_size=512):
    """Full-material workload: 200k-triangle displaced mesh (freq=100)."""
    albedo, rough, normal = noise_textures(tex_size, seed)
    checker = checker_texture()
    textures = [albedo, rough, normal, checker]
    attrs = [
        dict(name="white", kind=MAT_DIFFUSE, albedo=(0.73, 0.73, 0.73), is_sppm=1, if_opaque=1),              # 0
        dict(name="red", kind=MAT_DIFFUSE, albedo=(0.63, 0.06, 0.05), is_sppm=1, if_opaque=1),                # 1
        dict(name="green", kind=MAT_DIFFUSE, albedo=(0.14, 0.45, 0.09), is_sppm=1, if_opaque=1),              # 2
        dict(name="light", kind=MAT_DIFFUSE, albedo=(1.0, 1.0, 1.0), radiance=(255.0, 255.0, 255.0), is_sppm=1,
             if_opaque=1, tex_albedo=3),                                                                      # 3 light
        # GGX conductor: albedo = F0, scalar complex IOR (eta=1, k=20) so F(0) ~ 0.99
        dict(name="conductor_sphere", kind=MAT_ROUGH_CONDUCTOR, albedo=(0.95, 0.64, 0.54), roughness=0.3, index=1.0, extinction=20.0,
             is_sppm=1, if_opaque=1),                                                                         # 4
        dict(name="glass_sphere", kind=MAT_GGX_GLASS, albedo=(0.98, 0.98, 0.98), roughness=0.1, index=1.5, is_sppm=0,
             if_opaque=0),                                                                                    # 5
        dict(name="floor", kind=MAT_DIFFUSE, albedo=(0.8, 0.8, 0.8), is_sppm=1, if_opaque=1, tex_albedo=0,
             tex_bump=2),                                                                             # 6
        dict(name="displaced_mesh", kind=MAT_ROUGH_CONDUCTOR, albedo=(0.9, 0.9, 0.9), roughness=0.4, index=1.0, extinction=20.0,
             is_sppm=1, if_opaque=1, tex_roughness=1),                                                        # 7 mesh
    ]
    b = _Builder()
    b.quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), 6)
    b.quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0), (0, -1, 0), 0)
    b.quad((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), (0, 0, 1), 0)
    b.quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), (1, 0, 0), 1)
    b.quad((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), (-1, 0, 0), 2)
    light_quad(b, 3)
    b.mesh(*uv_sphere((0.27, 0.2, 0.35), 0.2), a=4)
    b.mesh(*uv_sphere((0.73, 0.2, 0.65), 0.2), a=5)
    V, N, T = geodesic_sphere((0.5, 0.6, 0.3), 0.15, freq, seed)
    b.pos.extend(V.reshape(-1, 9))
    b.nrm.extend(N.reshape(-1, 9))
    b.uv.extend(T.reshape(-1, 9))
    b.attr.extend([7] * len(V))
    return b.scene(attrs, textures)


# ---- OBJ / MTL export ----------------------------------------------------------
def _f(x):
    return "%.9g" % float(np.float32(x))


def write_obj(scene, directory, stem="scene"):
    """Writes the scene as OBJ + MTL + binary PPM textures (the format the
    C++ facade's Scene::CreateScene reads).  Every facet is written with its
    own three v / vt / vn records, in facet order, so facet i of the file is
    facet i of the scene.  Materials are written in attribute order with
    Kd = albedo, Ke = radiance and map_Kd / map_d / map_bump; the surface
    kind and the remaining parameters come from the material name (the
    facade's example material manager maps the names used here).  Textures
    must be 8-bit (stored as byte / 255, alpha 1).  Returns the OBJ path."""
    import os
    os.makedirs(directory, exist_ok=True)
    tex_names = []
    for i, t in enumerate(scene.textures):
        b = np.rint(t[..., :3].astype(np.float64) * 255.0)
        if not (np.all(t[..., 3] == 1.0) and np.array_equal((b.astype(np.float32) / np.float32(255.0)), t[..., :3])):
            raise ValueError(f"texture {i} is not an 8-bit RGB image")
        name = f"{stem}_tex{i}.ppm"
        h, w = t.shape[:2]
        with open(os.path.join(directory, name), "wb") as f:
            f.write(b"P6\n%d %d\n255\n" % (w, h))
            f.write(b.astype(np.uint8).tobytes())
        tex_names.append(name)
    with open(os.path.join(directory, stem + ".mtl"), "w") as f:
        for i, a in enumerate(scene.attrs):
            f.write(f"newmtl {a.get('name', f'material{i}')}\n")
            f.write("Kd %s\n" % " ".join(_f(v) for v in a.get("albedo", (0.0, 0.0, 0.0))))
            f.write("Ke %s\n" % " ".join(_f(v) for v in a.get("radiance", (0.0, 0.0, 0.0))))
            for key, role in (("tex_albedo", "map_Kd"), ("tex_roughness", "map_d"), ("tex_bump", "map_bump")):
                if a.get(key, -1) >= 0:
                    f.write(f"{role} {tex_names[a[key]]}\n")
    P, N, T = scene.pos.reshape(-1, 3), scene.nrm.reshape(-1, 3), scene.uv.reshape(-1, 3)
    lines = [f"mtllib {stem}.mtl\n"]
    lines += ["v %s %s %s\n" % (_f(p[0]), _f(p[1]), _f(p[2])) for p in P]
    lines += ["vt %s %s\n" % (_f(t[0]), _f(t[1])) for t in T]
    lines += ["vn %s %s %s\n" % (_f(n[0]), _f(n[1]), _f(n[2])) for n in N]
    cur = None
    for i, a in enumerate(scene.attr):
        if a != cur:
            lines.append(f"usemtl {scene.attrs[a].get('name', f'material{a}')}\n")
            cur = a
        k = 3 * i + 1
        lines.append(f"f {k}/{k}/{k} {k + 1}/{k + 1}/{k + 1} {k + 2}/{k + 2}/{k + 2}\n")
    path = os.path.join(directory, stem + ".obj")
    with open(path, "w") as f:
        f.writelines(lines)
    return path


def occluded_scene(clutter_freq=61, seed=99):
    """Occluded-light workload: two 1x1x1 rooms (x in [0,1] and [1,2]) joined
    by a 0.1 x 0.6 door gap in the shared wall x = 1 (z in [0.45, 0.55],
    y in [0, 0.6]); a 0.3 x 0.3 ceiling light with emission 20 in room B
    (x in [1,2]); the camera (occluded_camera) looks from room A toward the
    door, so light reaches the image only through the gap.  Clutter: four
    geodesic spheres displaced by 3-octave value noise, 20 * freq^2 triangles
    each (~300k total at freq 61), two per room, diffuse."""
    attrs = [
        dict(name="white", kind=MAT_DIFFUSE, albedo=(0.73, 0.73, 0.73), is_sppm=1, if_opaque=1),        # 0
        dict(name="red", kind=MAT_DIFFUSE, albedo=(0.63, 0.06, 0.05), is_sppm=1, if_opaque=1),          # 1
        dict(name="green", kind=MAT_DIFFUSE, albedo=(0.14, 0.45, 0.09), is_sppm=1, if_opaque=1),        # 2
        dict(name="light", kind=MAT_DIFFUSE, albedo=(0.78, 0.78, 0.78), radiance=(20.0, 20.0, 20.0),
             is_sppm=1, if_opaque=1),                                                                 # 3
        dict(name="clutter", kind=MAT_DIFFUSE, albedo=(0.6, 0.55, 0.5), is_sppm=1, if_opaque=1),        # 4
    ]
    b = _Builder()
    up, down = (0, 1, 0), (0, -1, 0)
    b.quad((0, 0, 0), (2, 0, 0), (2, 0, 1), (0, 0, 1), up, 0)          # floor y = 0
    b.quad((0, 1, 0), (0, 1, 1), (2, 1, 1), (2, 1, 0), down, 0)        # ceiling y = 1
    b.quad((0, 0, 0), (0, 1, 0), (2, 1, 0), (2, 0, 0), (0, 0, 1), 0)   # z = 0
    b.quad((0, 0, 1), (2, 0, 1), (2, 1, 1), (0, 1, 1), (0, 0, -1), 0)  # z = 1
    b.quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), (1, 0, 0), 1)   # x = 0 (room A end wall)
    b.quad((2, 0, 0), (2, 1, 0), (2, 1, 1), (2, 0, 1), (-1, 0, 0), 2)  # x = 2 (room B end wall)
    # shared wall x = 1 around the door gap z in [0.45, 0.55], y in [0, 0.6]
    n = (-1, 0, 0)
    b.quad((1, 0, 0), (1, 1, 0), (1, 1, 0.45), (1, 0, 0.45), n, 0)
    b.quad((1, 0, 0.55), (1, 1, 0.55), (1, 1, 1), (1, 0, 1), n, 0)
    b.quad((1, 0.6, 0.45), (1, 1, 0.45), (1, 1, 0.55), (1, 0.6, 0.55), n, 0)
    light_quad(b, 3, size=0.3, y=0.999, cx=1.5, cz=0.5)
    centers = [((0.3, 0.16, 0.3), 0.15), ((0.62, 0.13, 0.75), 0.12), ((1.4, 0.15, 0.3), 0.14),
               ((1.7, 0.12, 0.7), 0.11)]
    for k, (c, r) in enumerate(centers):
        V, N, T = geodesic_sphere(c, r, clutter_freq, seed + k, amp=0.03)
        b.pos.extend(V.reshape(-1, 9))
        b.nrm.extend(N.reshape(-1, 9))
        b.uv.extend(T.reshape(-1, 9))
        b.attr.extend([4] * len(V))
    return b.scene(attrs)


def occluded_camera(width, height, sampler_id=0):
    """Pinhole camera in room A, next to the shared wall, looking back into the
    room: every visible surface is lit only through the door, which is out of
    view.  Rays start on the image plane DistanceToLens behind the camera
    position (Camera.h:73-82), so it is kept small to stay inside the room."""
    return dict(type=CAM_PINHOLE, width=width, height=height, pos=(0.95, 0.55, 0.1), target=(0.0, 0.3, 0.8),
                fov_deg=60.0, F=1.0, dist_to_lens=0.01, aspect=width / height, sampler=SAMPLER_UNIFORM,
                sampler_id=sampler_id, sampler_nsample=1)


# ---- batched workload: many uneven scenes ----------------------------------------
_ICO = None


def _icosahedron():
    global _ICO
    if _ICO is None:
        t = (1 + 5 ** 0.5) / 2
        v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                      (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], np.float64)
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                      (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                      (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)])
        _ICO = (v, f)
    return _ICO


def displaced_mesh(center, r, freq, seed, amp=0.05):
    """Vectorised geodesic sphere: 20 * freq^2 triangles, radially displaced by
    3-octave value noise on a 16^3 lattice (a function of the direction, so
    shared vertices stay shared).  Per-facet vertex arrays (V, N, UV), face
    normals as vertex normals.  The noise is evaluated once per lattice point
    (i, j), i + j <= freq, of each icosahedron face, and the facets gather
    their corners from those points."""
    v, f = _icosahedron()
    i, j = np.meshgrid(np.arange(freq), np.arange(freq), indexing="ij")
    up = j < freq - i
    dn = j < freq - i - 1
    iu, ju, idn, jdn = i[up], j[up], i[dn], j[dn]
    # integer corners (i, j) of the up and down triangles
    cu = np.stack([np.stack([iu, ju], -1), np.stack([iu + 1, ju], -1), np.stack([iu, ju + 1], -1)], 1)
    cd = np.stack([np.stack([idn + 1, jdn], -1), np.stack([idn + 1, jdn + 1], -1), np.stack([idn, jdn + 1], -1)], 1)
    corners = np.concatenate([cu, cd], 0)  # (T, 3, 2)
    gi, gj = np.meshgrid(np.arange(freq + 1), np.arange(freq + 1), indexing="ij")
    keep = gi + gj <= freq
    gi, gj = gi[keep], gj[keep]
    lut = np.full((freq + 1, freq + 1), -1, np.int64)
    lut[gi, gj] = np.arange(len(gi))
    tri = lut[corners[..., 0], corners[..., 1]]  # (T, 3) lattice point of each corner
    bary = np.stack([gi, gj], -1).astype(np.float64) / freq  # (K, 2)
    A, B, Cc = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]  # (20, 3)
    P = (A[:, None, :] + (B - A)[:, None, :] * bary[None, :, 0:1] + (Cc - A)[:, None, :] * bary[None, :, 1:2])
    P /= np.linalg.norm(P, axis=-1, keepdims=True)  # (20, K, 3)
    rng = np.random.default_rng(seed)
    offs = [rng.random(3) * 10 for _ in range(3)]
    lat = rng.random((3, 16, 16, 16))
    flat = P.reshape(-1, 3)
    val = np.zeros(len(flat))
    amp_, tot = 1.0, 0.0
    for o in range(3):
        q = (flat * (2.0 ** o) * 2.0 + offs[o]) % 16
        i0 = np.floor(q).astype(np.int64)
        fr = q - i0
        fr = fr * fr * (3 - 2 * fr)
        i1 = (i0 + 1) % 16
        acc = np.zeros(len(flat))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    ix = i1[:, 0] if dx else i0[:, 0]
                    iy = i1[:, 1] if dy else i0[:, 1]
                    iz = i1[:, 2] if dz else i0[:, 2]
                    w = (fr[:, 0] if dx else 1 - fr[:, 0]) * (fr[:, 1] if dy else 1 - fr[:, 1]) * \
                        (fr[:, 2] if dz else 1 - fr[:, 2])
                    acc += w * lat[o][ix, iy, iz]
        val += amp_ * acc
        tot += amp_
        amp_ *= 0.5
    disp = 1.0 + amp / r * (val / tot - 0.5) * 2.0
    Vu = (flat * (r * disp)[:, None] + np.asarray(center)).reshape(20, -1, 3)
    uvu = np.zeros_like(P)
    uvu[..., 0] = 0.5 + np.arctan2(P[..., 2], P[..., 0]) / (2 * np.pi)
    uvu[..., 1] = 0.5 - np.arcsin(np.clip(P[..., 1], -1, 1)) / np.pi
    V = Vu[:, tri].reshape(-1, 3, 3)
    uv = uvu[:, tri].reshape(-1, 3, 3)
    e1, e2 = V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]
    fn = np.cross(e1, e2)
    fn /= np.linalg.norm(fn, axis=1, keepdims=True) + 1e-30
    N = np.repeat(fn[:, None, :], 3, axis=1)
    return V.astype(np.float32), N.astype(np.float32), uv.astype(np.float32)


def random_material(rng, name):
    k = rng.integers(4)
    alb = tuple(float(x) for x in rng.uniform(0.3, 0.9, 3))
    if k == 0:
        return dict(name=name, kind=MAT_DIFFUSE, albedo=alb, is_sppm=1, if_opaque=1)
    if k == 1:
        return dict(name=name, kind=MAT_METAL, albedo=alb, if_delta=1, is_sppm=0, if_opaque=1)
    if k == 2:
        return dict(name=name, kind=MAT_ROUGH_CONDUCTOR, albedo=alb, roughness=float(rng.uniform(0.1, 0.5)),
                    index=1.0, extinction=20.0, is_sppm=1, if_opaque=1)
    return dict(name=name, kind=MAT_GGX_GLASS, albedo=(0.98, 0.98, 0.98), roughness=float(rng.uniform(0.05, 0.3)),
                index=1.5, is_sppm=0, if_opaque=0)


def batch_scene(seed, tri_range=(10_000, 500_000), max_meshes=8):
    """One scene of the batched workload: the box (walls + 0.25^2 light with
    emission 15) plus 1-8 displaced meshes whose triangle counts are drawn
    log-uniform in tri_range, each with a random material over diffuse /
    specular / GGX conductor / GGX dielectric."""
    rng = np.random.default_rng(seed)
    attrs = [
        dict(name="white", kind=MAT_DIFFUSE, albedo=(0.73, 0.73, 0.73), is_sppm=1, if_opaque=1),
        dict(name="red", kind=MAT_DIFFUSE, albedo=(0.63, 0.06, 0.05), is_sppm=1, if_opaque=1),
        dict(name="green", kind=MAT_DIFFUSE, albedo=(0.14, 0.45, 0.09), is_sppm=1, if_opaque=1),
        dict(name="light", kind=MAT_DIFFUSE, albedo=(0.78, 0.78, 0.78), radiance=(15.0, 15.0, 15.0), is_sppm=1,
             if_opaque=1),
    ]
    b = _Builder()
    box_walls(b, 0, 1, 2)
    light_quad(b, 3)
    parts = [(np.array(b.pos), np.array(b.nrm), np.array(b.uv), np.array(b.attr, np.int32))]
    nmesh = int(rng.integers(1, max_meshes + 1))
    for k in range(nmesh):
        ntri = float(np.exp(rng.uniform(np.log(tri_range[0]), np.log(tri_range[1]))))
        freq = max(2, int(round(np.sqrt(ntri / 20.0))))
        r = float(rng.uniform(0.06, 0.15))
        c = (float(rng.uniform(0.2, 0.8)), float(rng.uniform(r, 0.6)), float(rng.uniform(0.2, 0.8)))
        attrs.append(random_material(rng, f"mesh{k}"))
        V, N, T = displaced_mesh(c, r, freq, seed * 31 + k, amp=0.2 * r)
        parts.append((V.reshape(-1, 9), N.reshape(-1, 9), T.reshape(-1, 9), np.full(len(V), len(attrs) - 1, np.int32)))
    return Scene(*(np.concatenate([p[k] for p in parts]) for k in range(4)), attrs, [])


def batch_workload(n=64, seed=2024, tri_range=(10_000, 500_000), max_meshes=8, resolutions=(256, 512, 1024),
                   progress=None):
    """The batched workload: n scenes of uneven size, square resolutions drawn
    from `resolutions`, photons per pass = U(0.5, 2) x pixels."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n):
        if progress:
            progress(s)
        res = int(rng.choice(resolutions))
        ppp = float(rng.uniform(0.5, 2.0))
        sc = batch_scene(seed * 1000 + s, tri_range, max_meshes)
        out.append(dict(scene=sc, cam=pinhole_camera(res, res), photons=int(ppp * res * res), alpha=0.7,
                        r0=0.02 * sc.diagonal()))
    return out


# ---- caustic workload: glass objects on a textured floor, small textured light --
def caustic_scene(seed=5, tris=2_000_000, tex_size=2048, nobj=4):
    """Caustic-heavy workload: `nobj` displaced geodesic meshes of GGX glass
    (IOR 1.5, roughness 0.05), ~tris triangles in total, on a diffuse floor
    with tex_size^2 procedural albedo / roughness / normal textures, lit by a
    0.05 x 0.05 textured light (emission up to 200) 1.0 above the floor."""
    rng = np.random.default_rng(seed)
    albedo, rough, normal = noise_textures(tex_size, seed)
    light_tex = checker_texture(n=256, cells=8, lo=128, hi=255)
    attrs = [
        dict(name="floor", kind=MAT_DIFFUSE, albedo=(0.8, 0.8, 0.8), is_sppm=1, if_opaque=1, tex_albedo=0,
             tex_roughness=1, tex_bump=2),                                                                   # 0
        dict(name="light", kind=MAT_DIFFUSE, albedo=(1.0, 1.0, 1.0), radiance=(200.0, 200.0, 200.0), is_sppm=1,
             if_opaque=1, tex_albedo=3),                                                                     # 1
        dict(name="glass", kind=MAT_GGX_GLASS, albedo=(0.98, 0.98, 0.98), roughness=0.05, index=1.5, is_sppm=0,
             if_opaque=0),                                                                                   # 2
    ]
    b = _Builder()
    b.quad((-1.0, 0.0, -1.0), (-1.0, 0.0, 2.0), (2.0, 0.0, 2.0), (2.0, 0.0, -1.0), (0, 1, 0), 0)
    h = 0.025
    b.quad((0.5 - h, 1.0, 0.5 - h), (0.5 - h, 1.0, 0.5 + h), (0.5 + h, 1.0, 0.5 + h), (0.5 + h, 1.0, 0.5 - h),
           (0, -1, 0), 1)
    base = b.scene(attrs)
    freq = max(2, int(round(np.sqrt(tris / nobj / 20.0))))
    parts = [base]
    for k in range(nobj):
        ang = 2.0 * np.pi * (k + rng.uniform(-0.15, 0.15)) / nobj
        r = float(rng.uniform(0.14, 0.2))
        c = (0.5 + 0.32 * np.cos(ang), r * 1.05, 0.5 + 0.32 * np.sin(ang))
        V, N, T = displaced_mesh(c, r, freq, seed * 101 + k, amp=0.15 * r)
        parts.append(Scene(V.reshape(-1, 9), N.reshape(-1, 9), T.reshape(-1, 9), np.full(len(V), 2, np.int32),
                           attrs, []))
    return Scene(np.concatenate([p.pos for p in parts]), np.concatenate([p.nrm for p in parts]),
                 np.concatenate([p.uv for p in parts]), np.concatenate([p.attr for p in parts]), attrs,
                 [albedo, rough, normal, light_tex])


def caustic_cameras(n=8, width=2048, height=2048, aperture=0.02):
    """`n` thin-lens variants on a circle around the objects, each focused on
    the scene centre (0.5, 0.15, 0.5)."""
    cams = []
    tgt = np.array([0.5, 0.15, 0.5])
    for k in range(n):
        ang = 2.0 * np.pi * (k + 0.5) / n
        pos = tgt + np.array([1.6 * np.cos(ang), 0.75, 1.6 * np.sin(ang)])
        focus = float(np.linalg.norm(pos - tgt))
        c = thinlens_camera(width, height, aperture=aperture, focus=focus, fov_deg=40.0, sampler_id=k)
        c["pos"] = tuple(float(x) for x in pos)
        c["target"] = tuple(float(x) for x in tgt)
        cams.append(c)
    return cams


# ---- environment light --------------------------------------------------------
def sky_texture(w=256, h=128):
    """Equirectangular sky for the SPHERE mapping (u = azimuth, v = polar
    angle from +y): a vertical gradient with a warm sun patch, RGB8."""
    v = (np.arange(h) + 0.5) / h
    u = (np.arange(w) + 0.5) / w
    base = np.stack([0.35 + 0.5 * (1 - v), 0.5 + 0.4 * (1 - v), 0.9 - 0.3 * v], -1)[:, None, :] * np.ones((1, w, 1))
    sun = np.exp(-(((u[None, :] - 0.3) / 0.04) ** 2 + ((v[:, None] - 0.25) / 0.06) ** 2))
    img = np.clip(base + sun[..., None] * np.array([1.0, 0.8, 0.5]), 0, 1)
    return rgb8_to_rgba32f((img * 255).astype(np.uint8))


def box_scene_ibl():
    """The box scene lit also by an environment map through its open side
    (z = 1): camera rays that leave the box and eye paths that bounce out of it
    see the sky."""
    sc = box_scene()
    out = Scene(sc.pos, sc.nrm, sc.uv, sc.attr, sc.attrs, list(sc.textures) + [sky_texture()])
    out.ibl = len(out.textures) - 1
    return out


def alpha_texture(alpha):
    """RGBA32F texture, grey, with the given alpha plane (H, W)."""
    a = np.asarray(alpha, np.float32)
    out = np.full(a.shape + (4,), 0.7, np.float32)
    out[..., 3] = a
    return out


def screens_scene(seed=11, stack=8):
    """The box with screens facing the open side (z = 1), for shadow queries
    through alpha-textured surfaces:
      A  z = 0.5,  x, y in [0.1, 0.9]: alpha 0.2 for u < 0.5 (1 - alpha = 0.8),
         0.9 for u >= 0.5 (1 - alpha = 0.1), 64x64 texels;
      B  z = 0.3,  x in [0.1, 0.3]: untextured (always occludes);
      C  z = 0.15, x in [0.5, 0.9]: value-noise alpha in [0, 1], 128x128;
      D  `stack` fully clear layers (alpha 0) at z = 0.9, 0.85, ... over
         x in [0.91, 0.99]."""
    half = np.where(np.arange(64)[None, :] < 32, 0.2, 0.9) * np.ones((64, 1))
    nz = _value_noise(128, seed)
    nz = (nz - nz.min()) / (nz.max() - nz.min())
    textures = [alpha_texture(half), alpha_texture(nz), alpha_texture(np.zeros((4, 4)))]
    attrs = [
        dict(name="white", kind=MAT_DIFFUSE, albedo=(0.73, 0.73, 0.73), is_sppm=1, if_opaque=1),   # 0
        dict(name="red", kind=MAT_DIFFUSE, albedo=(0.63, 0.06, 0.05), is_sppm=1, if_opaque=1),     # 1
        dict(name="green", kind=MAT_DIFFUSE, albedo=(0.14, 0.45, 0.09), is_sppm=1, if_opaque=1),   # 2
        dict(name="light", kind=MAT_DIFFUSE, albedo=(0.78, 0.78, 0.78), radiance=(15.0, 15.0, 15.0), is_sppm=1,
             if_opaque=1),                                                                        # 3
        dict(name="screen_halves", kind=MAT_DIFFUSE, albedo=(1, 1, 1), is_sppm=1, if_opaque=1, tex_albedo=0),  # 4
        dict(name="screen_noise", kind=MAT_DIFFUSE, albedo=(1, 1, 1), is_sppm=1, if_opaque=1, tex_albedo=1),   # 5
        dict(name="screen_clear", kind=MAT_DIFFUSE, albedo=(1, 1, 1), is_sppm=1, if_opaque=1, tex_albedo=2),   # 6
    ]
    b = _Builder()
    box_walls(b, 0, 1, 2)
    light_quad(b, 3)
    n = (0, 0, 1)
    b.quad((0.1, 0.1, 0.5), (0.9, 0.1, 0.5), (0.9, 0.9, 0.5), (0.1, 0.9, 0.5), n, 4)
    b.quad((0.1, 0.1, 0.3), (0.3, 0.1, 0.3), (0.3, 0.9, 0.3), (0.1, 0.9, 0.3), n, 0)
    b.quad((0.5, 0.1, 0.15), (0.9, 0.1, 0.15), (0.9, 0.9, 0.15), (0.5, 0.9, 0.15), n, 5)
    for k in range(stack):
        z = 0.9 - 0.05 * k
        b.quad((0.91, 0.1, z), (0.99, 0.1, z), (0.99, 0.9, z), (0.91, 0.9, z), n, 6)
    return b.scene(attrs, textures)


def transmit_scene():
    """The box with the reference main scene's extra surfaces
    (main.cpp:14-114 sets them up the same way):
      leaf    DiffuseTransmit quad (Reflectance 0.88, not opaque, hit points),
      bell    DiffuseTransmit sphere (Reflectance 0.7, not opaque, hit points),
      star    GlossTransmit sphere (roughness 0.25, Reflectance 0.1, not
              opaque, hit points),
      panel   BiMetal quad (delta, opaque, no hit points) leaning on the
              right wall, facing away from the light on one side."""
    attrs = [
        dict(name="white", kind=MAT_DIFFUSE, albedo=(0.73, 0.73, 0.73), is_sppm=1, if_opaque=1),     # 0
        dict(name="red", kind=MAT_DIFFUSE, albedo=(0.63, 0.06, 0.05), is_sppm=1, if_opaque=1),       # 1
        dict(name="green", kind=MAT_DIFFUSE, albedo=(0.14, 0.45, 0.09), is_sppm=1, if_opaque=1),     # 2
        dict(name="light", kind=MAT_DIFFUSE, albedo=(0.78, 0.78, 0.78), radiance=(15.0, 15.0, 15.0), is_sppm=1,
             if_opaque=1),                                                                          # 3
        dict(name="tree_leaf", kind=MAT_DIFFUSE_TRANSMIT, albedo=(0.3, 0.7, 0.2), reflectance=0.88, is_sppm=1,
             if_opaque=0),                                                                          # 4
        dict(name="ornament_bell_top", kind=MAT_DIFFUSE_TRANSMIT, albedo=(0.9, 0.8, 0.3), reflectance=0.7,
             is_sppm=1, if_opaque=0),                                                               # 5
        dict(name="star_outside", kind=MAT_GLOSS_TRANSMIT, albedo=(0.97, 0.98, 0.98), roughness=0.25,
             reflectance=0.1, is_sppm=1, if_opaque=0),                                              # 6
        dict(name="spotlight_cone", kind=MAT_BIMETAL, albedo=(0.9, 0.9, 0.9), if_delta=1, is_sppm=0,
             if_opaque=1),                                                                          # 7
    ]
    b = _Builder()
    box_walls(b, 0, 1, 2)
    light_quad(b, 3)
    b.quad((0.12, 0.35, 0.6), (0.45, 0.35, 0.45), (0.45, 0.7, 0.45), (0.12, 0.7, 0.6), tuple(
        np.cross(np.subtract((0.45, 0.35, 0.45), (0.12, 0.35, 0.6)), (0, 1, 0)) /
        np.linalg.norm(np.cross(np.subtract((0.45, 0.35, 0.45), (0.12, 0.35, 0.6)), (0, 1, 0)))), 4)
    b.mesh(*uv_sphere((0.3, 0.15, 0.3), 0.15, nseg=24, nring=12), a=5)
    b.mesh(*uv_sphere((0.7, 0.22, 0.55), 0.2, nseg=32, nring=16), a=6)
    b.quad((0.97, 0.05, 0.1), (0.75, 0.05, 0.1), (0.75, 0.6, 0.35), (0.97, 0.6, 0.35), (0, -0.41, 0.91), 7)
    return b.scene(attrs)


def nee_scene():
    """The box with surfaces that carry no hit points but are not delta, where
    eye-path next-event estimation applies: a GGX conductor sphere (roughness
    0.3) and a GGX glass sphere, beside a mirror sphere (delta).  The light is
    seen directly and in the spheres."""
    attrs = [
        dict(name="white", kind=MAT_DIFFUSE, albedo=(0.73, 0.73, 0.73), is_sppm=1, if_opaque=1),       # 0
        dict(name="red", kind=MAT_DIFFUSE, albedo=(0.63, 0.06, 0.05), is_sppm=1, if_opaque=1),         # 1
        dict(name="green", kind=MAT_DIFFUSE, albedo=(0.14, 0.45, 0.09), is_sppm=1, if_opaque=1),       # 2
        dict(name="light", kind=MAT_DIFFUSE, albedo=(0.78, 0.78, 0.78), radiance=(15.0, 15.0, 15.0), is_sppm=1,
             if_opaque=1),                                                                            # 3
        dict(name="lamp_body", kind=MAT_ROUGH_CONDUCTOR, albedo=(0.95, 0.64, 0.54), roughness=0.3, index=1.1978,
             extinction=7.0488, is_sppm=0, if_opaque=1),                                               # 4
        dict(name="lamp_glass", kind=MAT_GGX_GLASS, albedo=(0.97, 0.98, 0.98), roughness=0.1, index=1.5,
             is_sppm=0, if_opaque=0),                                                                 # 5
        dict(name="mirror", kind=MAT_METAL, albedo=(0.95, 0.95, 0.95), if_delta=1, is_sppm=0, if_opaque=1),  # 6
    ]
    b = _Builder()
    box_walls(b, 0, 1, 2)
    light_quad(b, 3)
    b.mesh(*uv_sphere((0.3, 0.25, 0.45), 0.22, nseg=32, nring=16), a=4)
    b.mesh(*uv_sphere((0.72, 0.2, 0.6), 0.18, nseg=32, nring=16), a=5)
    b.mesh(*uv_sphere((0.75, 0.62, 0.3), 0.12, nseg=24, nring=12), a=6)
    return b.scene(attrs)
