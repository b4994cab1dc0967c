"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the
reference-built checker (oracle/_ref/libsppm_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product path.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libsppm_ref.so")

_PKG = os.path.join(os.path.dirname(HERE), "friendsrender-r-public_amd")
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)
from sppm_amd import (CameraDesc, Params, SceneDesc, camera_desc, diaphragm_rgba, params_struct,  # noqa: E402
                      scene_desc)

F = C.POINTER(C.c_float)
I = C.POINTER(C.c_int32)
D = C.POINTER(C.c_double)


def fp(a):
    return a.ctypes.data_as(F)


def ip(a):
    return a.ctypes.data_as(I)


class FParam(C.Structure):
    _fields_ = [("in_", C.c_float * 3), ("out", C.c_float * 3), ("n", C.c_float * 3), ("from_index", C.c_float),
                ("to_index", C.c_float), ("roughness", C.c_float), ("absorption", C.c_float),
                ("reflectance", C.c_float)]


def _load_oracle():
    if not os.path.exists(ORACLE_PATH):
        raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
    L = C.CDLL(ORACLE_PATH)
    L.orc_mt_real1_seq.argtypes = [C.c_uint32, C.c_int, D]
    L.orc_mt_raw_seq.argtypes = [C.c_uint32, C.c_int, C.POINTER(C.c_uint32)]
    L.orc_uniform_sampler_seq.argtypes = [C.c_int, C.c_int, F]
    L.orc_halton.argtypes = [C.c_int, C.c_int]
    L.orc_halton.restype = C.c_float
    L.orc_hammersley.argtypes = [C.c_int, C.c_int, F]
    L.orc_table_build.argtypes = [F, C.c_int, F, I, I, F]
    L.orc_table_sample_seq.argtypes = [F, I, I, C.c_int, C.c_uint32, C.c_int, I]
    L.orc_bilinear.argtypes = [C.c_void_p, C.c_float, C.c_float, F]
    L.orc_matrix3_inv.argtypes = [F, F]
    L.orc_bsdf_brdf.argtypes = [C.c_int, C.POINTER(FParam)]
    L.orc_bsdf_brdf.restype = C.c_float
    L.orc_bsdf_pdf.argtypes = [C.c_int, C.POINTER(FParam)]
    L.orc_bsdf_pdf.restype = C.c_float
    L.orc_bsdf_sample.argtypes = [C.c_int, C.POINTER(FParam), C.c_uint32, F, F, C.POINTER(C.c_int)]
    L.orc_bsdf_eval_batch.argtypes = [C.c_int, C.POINTER(FParam), C.c_int, F, F, F]
    L.orc_bsdf_sample_batch.argtypes = [C.c_int, C.POINTER(FParam), C.c_uint32, C.c_int, F, F]
    L.orc_scene_create.argtypes = [C.POINTER(SceneDesc)]
    L.orc_scene_create.restype = C.c_void_p
    L.orc_scene_destroy.argtypes = [C.c_void_p]
    L.orc_scene_num_lights.argtypes = [C.c_void_p]
    L.orc_scene_light_info.argtypes = [C.c_void_p, I, F, I, F, I, I, F]
    L.orc_closest_hit_batch.argtypes = [C.c_void_p, C.c_int, C.c_int, F, F, C.c_float, C.c_float, I, F, F, F]
    L.orc_occluded_batch.argtypes = [C.c_void_p, C.c_int, F, F, C.c_float, F, C.c_float, I]
    L.orc_camera_ray.argtypes = [C.POINTER(CameraDesc), C.c_float, C.c_float, F, F, F, F]
    L.orc_camera_ray_diaphragm.argtypes = [C.POINTER(CameraDesc), C.c_int, C.c_int, F, C.c_float, C.c_float, F, F,
                                           F, F]
    L.orc_render_diaphragm.argtypes = [C.c_void_p, C.c_int, C.c_int, F]
    L.orc_render_create.argtypes = [C.c_void_p, C.POINTER(CameraDesc), C.POINTER(Params), C.c_int]
    L.orc_render_create.restype = C.c_void_p
    L.orc_render_destroy.argtypes = [C.c_void_p]
    L.orc_photon_pass.argtypes = [C.c_void_p, C.c_int]
    L.orc_flatten.argtypes = [C.c_void_p]
    L.orc_read_photons.argtypes = [C.c_void_p, F, F, F, F, F]
    L.orc_eye_pass.argtypes = [C.c_void_p, C.c_int]
    L.orc_read_hitpoints.argtypes = [C.c_void_p, I, I, F, F, F, F]
    L.orc_gather.argtypes = [C.c_void_p, C.c_int]
    L.orc_read_counts.argtypes = [C.c_void_p, I]
    L.orc_neighbour_total.argtypes = [C.c_void_p]
    L.orc_neighbour_total.restype = C.c_int64
    L.orc_read_neighbours.argtypes = [C.c_void_p, C.POINTER(C.c_int64), I]
    L.orc_read_shared.argtypes = [C.c_void_p, F, F, I, I]
    L.orc_read_direct.argtypes = [C.c_void_p, F, F, I]
    L.orc_ibl_color.argtypes = [C.c_void_p, F, F]
    L.orc_ibl_uv.argtypes = [F, F]
    L.orc_write_shared.argtypes = [C.c_void_p, F, F, I, I]
    L.orc_reconstruct.argtypes = [C.c_void_p, C.c_int32, C.c_int32, F]
    L.orc_full_pass.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.orc_uniform_sampler_at.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int64), F]
    L.orc_eye_pixel.argtypes = [C.c_void_p, C.c_int, C.c_int, F, I, I, F, F, F, F]
    L.orc_photon_path.argtypes = [C.c_void_p, C.c_int, C.c_int, F, F, F, F, F, I]
    L.orc_read_path_counts.argtypes = [C.c_void_p, I]
    L.orc_script_set.argtypes = [D, C.c_int]
    L.orc_script_used.restype = C.c_int
    L.orc_guide_create.argtypes = [C.c_void_p, C.c_float]
    L.orc_guide_create.restype = C.c_void_p
    L.orc_guide_destroy.argtypes = [C.c_void_p]
    L.orc_guide_set_tables.argtypes = [C.c_void_p, F, F]
    L.orc_render_set_guide.argtypes = [C.c_void_p, C.c_void_p]
    L.orc_guide_frame.argtypes = [C.c_void_p, F]
    L.orc_guide_point_cells.argtypes = [C.c_void_p, F, I, I]
    L.orc_guide_density.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, F, F, I]
    L.orc_guide_density.restype = C.c_float
    L.orc_guide_emit.argtypes = [C.c_void_p, C.c_int, C.c_int, F, F, F, I]
    L.orc_guide_density_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_int, I, I, F, F, F]
    L.orc_guide_emit_batch.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, F, F, F, I]
    return L


class script:
    """Context manager: RandomMT draws on this thread come from `draws`
    (0.5 past the end).  `used` afterwards = draws taken."""

    def __init__(self, draws):
        self.draws = np.ascontiguousarray(draws, np.float64)
        self.used = None

    def __enter__(self):
        orc().orc_script_set(self.draws.ctypes.data_as(D), len(self.draws))
        return self

    def __exit__(self, *exc):
        self.used = orc().orc_script_used()
        orc().orc_script_set(None, -1)
        return False


GUIDE_CELLS = 262144


class OracleGuide:
    """Guided emission over an OracleScene's single rectangular light."""

    def __init__(self, oscene, mix_plain):
        self.oscene = oscene
        self.h = orc().orc_guide_create(oscene.h, mix_plain)
        if not self.h:
            raise ValueError("guided emission needs exactly one rectangular light of two facets")
        self.textured = bool(oscene.lights()["textured"][0])

    def __del__(self):
        try:
            orc().orc_guide_destroy(self.h)
        except Exception:
            pass

    def set_tables(self, prob, cdf=None):
        prob = np.ascontiguousarray(prob, np.float32)
        if cdf is None:
            cdf = np.cumsum(prob, dtype=np.float64).astype(np.float32)
        cdf = np.ascontiguousarray(cdf, np.float32)
        orc().orc_guide_set_tables(self.h, fp(prob), fp(cdf))

    def frame(self):
        out = np.zeros(18, np.float32)
        orc().orc_guide_frame(self.h, fp(out))
        return dict(P0=out[0:3], e1=out[3:6], e2=out[6:9], n=out[9:12], t1=out[12:15], t2=out[15:18])

    def point_cells(self):
        mass = np.zeros(1024, np.float32)
        start = np.zeros(1025, np.int32)
        cell = np.zeros(2000, np.int32)
        orc().orc_guide_point_cells(self.h, fp(mass), ip(start), ip(cell))
        return mass, start, cell

    def density_batch(self, pos, d, light=None, point=None):
        pos = np.ascontiguousarray(pos, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        n = pos.shape[0]
        q = np.zeros(n, np.float32)
        la = np.ascontiguousarray(light, np.int32) if light is not None else None
        pa = np.ascontiguousarray(point, np.int32) if point is not None else None
        orc().orc_guide_density_batch(self.oscene.h, self.h, n, ip(la) if la is not None else None,
                                      ip(pa) if pa is not None else None, fp(pos), fp(d), fp(q))
        return q

    def density(self, pos, d, light=0, point=0):
        pos = np.ascontiguousarray(pos, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        c = np.zeros(1, np.int32)
        q = orc().orc_guide_density(self.oscene.h, self.h, light, point, fp(pos), fp(d), ip(c))
        return q, int(c[0])


_ORC = None


def orc():
    global _ORC
    if _ORC is None:
        _ORC = _load_oracle()
    return _ORC


def ref_available():
    return os.path.exists(REF_PATH)


_REF = None


def ref():
    global _REF
    if _REF is None:
        L = C.CDLL(REF_PATH)
        L.ref_mt_real1_seq.argtypes = [C.c_int, C.c_int, D]
        L.ref_uniform_sampler_seq.argtypes = [C.c_int, C.c_int, F]
        L.ref_halton.argtypes = [C.c_int, C.c_int]
        L.ref_halton.restype = C.c_float
        L.ref_hammersley.argtypes = [C.c_int, C.c_int, F]
        L.ref_table_build.argtypes = [F, C.c_int, F, I, I, F]
        L.ref_table_pdf.argtypes = [F, C.c_int, F]
        L.ref_table_sample_seq.argtypes = [F, C.c_int, C.c_int, C.c_int, I]
        L.ref_bilinear.argtypes = [C.c_int, C.c_int, F, C.c_float, C.c_float, F]
        L.ref_bilinear_batch.argtypes = [C.c_int, C.c_int, F, C.c_int, F, F]
        L.ref_normal_from_map.argtypes = [C.c_int, C.c_int, F, C.c_float, C.c_float, F]
        L.ref_matrix3_inv.argtypes = [F, F]
        L.ref_bump_rotate.argtypes = [F, F, F, F, F]
        _REF = L
    return _REF


def fparam(indir, outdir, n, fi=1.0, ti=1.0, rough=1.0, absorb=1.0, refl=1.0):
    p = FParam()
    p.in_[:] = [float(x) for x in indir]
    p.out[:] = [float(x) for x in outdir]
    p.n[:] = [float(x) for x in n]
    p.from_index, p.to_index, p.roughness, p.absorption, p.reflectance = fi, ti, rough, absorb, refl
    return p


def bsdf_eval(kind, p, outdirs):
    """PDF() and BRDF() of surface `kind` at each outgoing direction (n x 3)."""
    o = np.ascontiguousarray(outdirs, np.float32)
    n = o.shape[0]
    pdf, brdf = np.zeros(n, np.float32), np.zeros(n, np.float32)
    orc().orc_bsdf_eval_batch(kind, C.byref(p), n, fp(o), fp(pdf), fp(brdf))
    return pdf, brdf


def bsdf_sample(kind, p, n, seed0=1):
    """n sampled directions and their sampling pdfs (RandomMT(seed0 + j))."""
    d = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    orc().orc_bsdf_sample_batch(kind, C.byref(p), seed0, n, fp(d), fp(pdf))
    return d, pdf


class OracleScene:
    def __init__(self, scene):
        self.d, self._keep = scene_desc(scene)
        self.h = orc().orc_scene_create(C.byref(self.d))
        self.scene = scene

    def __del__(self):
        try:
            orc().orc_scene_destroy(self.h)
        except Exception:
            pass

    @staticmethod
    def ibl_uv(d):
        d = np.ascontiguousarray(d, np.float32)
        out = np.zeros(2, np.float32)
        orc().orc_ibl_uv(fp(d), fp(out))
        return out

    def ibl_color(self, d):
        d = np.ascontiguousarray(d, np.float32)
        out = np.zeros(3, np.float32)
        orc().orc_ibl_color(self.h, fp(d), fp(out))
        return out

    def closest_hit(self, org, dirs, mode=1, tmin=1e-4, tmax=3.4028234663852886e38):
        org = np.ascontiguousarray(org, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = org.shape[0]
        f = np.zeros(n, np.int32)
        t, u, v = (np.zeros(n, np.float32) for _ in range(3))
        orc().orc_closest_hit_batch(self.h, mode, n, fp(org), fp(dirs), tmin, tmax, ip(f), fp(t), fp(u), fp(v))
        return f, t, u, v

    def occluded(self, org, dirs, tmin, tmax, translucency=0.5):
        org = np.ascontiguousarray(org, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = org.shape[0]
        tmax = np.ascontiguousarray(np.broadcast_to(np.float32(tmax) if np.isscalar(tmax) else tmax, (n,)), np.float32)
        out = np.zeros(n, np.int32)
        orc().orc_occluded_batch(self.h, n, fp(org), fp(dirs), tmin, fp(tmax), translucency, ip(out))
        return out

    def lights(self):
        n = orc().orc_scene_num_lights(self.h)
        facet, tex = np.zeros(n, np.int32), np.zeros(n, np.int32)
        flux = np.zeros((n, 3), np.float32)
        prob, left, right = np.zeros(n, np.float32), np.zeros(n, np.int32), np.zeros(n, np.int32)
        avg = np.zeros(1, np.float32)
        orc().orc_scene_light_info(self.h, ip(facet), fp(flux), ip(tex), fp(prob), ip(left), ip(right), fp(avg))
        return dict(facet=facet, flux=flux, textured=tex, prob=prob, left=left, right=right, avg=avg[0])


class OracleRender:
    """CPU restatement of one SPPM render (same API shape as sppm_amd.Renderer)."""

    def __init__(self, oscene, cam, alpha, first_radius, photons, nthreads=None, nee=0):
        self.oscene = oscene
        self.cam = camera_desc(cam)
        self.p = params_struct(alpha, first_radius, photons, nee=nee)
        nt = nthreads or min(16, os.cpu_count() or 1)
        self.h = orc().orc_render_create(oscene.h, C.byref(self.cam), C.byref(self.p), nt)
        if cam.get("diaphragm") is not None:
            self._dia = diaphragm_rgba(cam["diaphragm"])
            if orc().orc_render_diaphragm(self.h, self._dia.shape[1], self._dia.shape[0], fp(self._dia)):
                raise ValueError("diaphragm image has no open pixel")
        self.W, self.H, self.N = cam["width"], cam["height"], photons
        self.passes = 0
        self.total_photons = 0

    def __del__(self):
        try:
            orc().orc_render_destroy(self.h)
        except Exception:
            pass

    def set_guide(self, guide):
        """Photon passes emit from `guide` (an OracleGuide; None: plain)."""
        self._guide = guide
        orc().orc_render_set_guide(self.h, guide.h if guide is not None else None)

    def guide_emit_batch(self, step, first, n):
        pos, d = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
        pdf, c = np.zeros(n, np.float32), np.zeros(n, np.int32)
        orc().orc_guide_emit_batch(self.h, step, first, n, fp(pos), fp(d), fp(pdf), ip(c))
        return pos, d, pdf, c

    def guide_emit(self, step, path):
        pos, d = np.zeros(3, np.float32), np.zeros(3, np.float32)
        pdf, c = np.zeros(1, np.float32), np.zeros(1, np.int32)
        orc().orc_guide_emit(self.h, step, int(path), fp(pos), fp(d), fp(pdf), ip(c))
        return pos, d, float(pdf[0]), int(c[0])

    def full_pass(self, record=False):
        e = orc().orc_full_pass(self.h, self.passes, int(record))
        self.passes += 1
        self.total_photons += e
        return e

    def photons(self, step):
        return orc().orc_photon_pass(self.h, step)

    def flatten(self):
        return orc().orc_flatten(self.h)

    def eye(self, step):
        orc().orc_eye_pass(self.h, step)

    def gather(self, record=False):
        orc().orc_gather(self.h, int(record))

    def read_photons(self, n):
        pos, flux, indir = (np.zeros((n, 3), np.float32) for _ in range(3))
        index, pdf = np.zeros(n, np.float32), np.zeros(n, np.float32)
        orc().orc_read_photons(self.h, fp(pos), fp(flux), fp(indir), fp(index), fp(pdf))
        return dict(pos=pos, flux=flux, indir=indir, index=index, pdf=pdf)

    def read_hitpoints(self):
        n = self.W * self.H
        t, f = np.zeros(n, np.int32), np.zeros(n, np.int32)
        pos, w, o, nn = (np.zeros((n, 3), np.float32) for _ in range(4))
        orc().orc_read_hitpoints(self.h, ip(t), ip(f), fp(pos), fp(w), fp(o), fp(nn))
        return dict(type=t, facet=f, pos=pos, weight=w, outdir=o, normal=nn)

    def read_counts(self):
        m = np.zeros(self.W * self.H, np.int32)
        orc().orc_read_counts(self.h, ip(m))
        return m

    def read_neighbours(self):
        tot = orc().orc_neighbour_total(self.h)
        off = np.zeros(self.W * self.H + 1, np.int64)
        idx = np.zeros(max(1, tot), np.int32)
        orc().orc_read_neighbours(self.h, off.ctypes.data_as(C.POINTER(C.c_int64)), ip(idx))
        return off, idx[:tot]

    def read_shared(self):
        n = self.W * self.H
        rad, r = np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
        k, g = np.zeros(n, np.int32), np.zeros(n, np.int32)
        orc().orc_read_shared(self.h, fp(rad), fp(r), ip(k), ip(g))
        return dict(radiance=rad, radius=r, num_photon=k, num_gather=g)

    def read_direct(self):
        n = self.W * self.H
        hd, sd, k = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros(n, np.int32)
        orc().orc_read_direct(self.h, fp(hd), fp(sd), ip(k))
        return dict(hit=hd, shared=sd, num=k)

    def write_shared(self, radiance, radius, num_photon, num_gather):
        radiance = np.ascontiguousarray(radiance, np.float32)
        radius = np.ascontiguousarray(radius, np.float32)
        num_photon = np.ascontiguousarray(num_photon, np.int32)
        num_gather = np.ascontiguousarray(num_gather, np.int32)
        orc().orc_write_shared(self.h, fp(radiance), fp(radius), ip(num_photon), ip(num_gather))

    def reconstruct(self, total_photons=None, passes=None):
        total = self.total_photons if total_photons is None else total_photons
        p = self.passes if passes is None else passes
        img = np.zeros((self.H, self.W, 4), np.float32)
        orc().orc_reconstruct(self.h, total, p, fp(img))
        return img

    def eye_pixels(self, step, pixels, samples):
        """createHitpoints for selected pixels of pass `step` (samples: n x 3)."""
        n = len(pixels)
        out = dict(type=np.zeros(n, np.int32), facet=np.zeros(n, np.int32),
                   pos=np.zeros((n, 3), np.float32), weight=np.zeros((n, 3), np.float32),
                   normal=np.zeros((n, 3), np.float32), outdir=np.zeros((n, 3), np.float32))
        smp = np.ascontiguousarray(samples, np.float32)
        t, f = np.zeros(1, np.int32), np.zeros(1, np.int32)
        for j, pix in enumerate(pixels):
            orc().orc_eye_pixel(self.h, step, int(pix), fp(smp[j]), ip(t), ip(f), fp(out["pos"][j]),
                                fp(out["weight"][j]), fp(out["normal"][j]), fp(out["outdir"][j]))
            out["type"][j], out["facet"][j] = t[0], f[0]
        return out

    def photon_path(self, step, path):
        """createPhotonMap for one path: its new photons (emission photon if none stored)."""
        pos, flux, indir = (np.zeros((32, 3), np.float32) for _ in range(3))
        index, pdf = np.zeros(32, np.float32), np.zeros(32, np.float32)
        nl = np.zeros(1, np.int32)
        k = orc().orc_photon_path(self.h, step, int(path), fp(pos), fp(flux), fp(indir), fp(index), fp(pdf), ip(nl))
        n = int(nl[0])
        return dict(k=k, pos=pos[:n], flux=flux[:n], indir=indir[:n], index=index[:n], pdf=pdf[:n])

    def read_path_counts(self):
        c = np.zeros(self.N, np.int32)
        orc().orc_read_path_counts(self.h, ip(c))
        return c


def uniform_samples_at(sampler_id, calls):
    """UniformSample(sampler_id) draws number calls[j] (any order)."""
    calls = np.asarray(calls, np.int64)
    order = np.argsort(calls, kind="stable")
    sc = np.ascontiguousarray(calls[order])
    out = np.zeros((len(sc), 3), np.float32)
    orc().orc_uniform_sampler_at(sampler_id, len(sc), sc.ctypes.data_as(C.POINTER(C.c_int64)), fp(out))
    res = np.zeros_like(out)
    res[order] = out
    return res


def camera_ray(cam, u, v, sample3):
    c = camera_desc(cam)
    s = np.ascontiguousarray(sample3, np.float32)
    org, d, pdf = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(1, np.float32)
    if cam.get("diaphragm") is not None:
        img = diaphragm_rgba(cam["diaphragm"])
        if orc().orc_camera_ray_diaphragm(C.byref(c), img.shape[1], img.shape[0], fp(img), u, v, fp(s), fp(org),
                                          fp(d), fp(pdf)):
            raise ValueError("diaphragm image has no open pixel")
    else:
        orc().orc_camera_ray(C.byref(c), u, v, fp(s), fp(org), fp(d), fp(pdf))
    return org, d, pdf[0]
