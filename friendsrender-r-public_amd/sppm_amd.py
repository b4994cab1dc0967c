"""ctypes binding of libsppm_amd.so (include/sppm_amd.h).

This is plumbing for tests and the benchmark: every call goes straight to the
C ABI, whose passes run as hand-written HIP kernels.  There is no CPU path:
if the shared library is missing or fails to load, import raises.

Mirrors the reference's per-pass structure (SPPM/SPPM/SPPM.cpp:770-854):
``Renderer.pass_()`` is one loop body, ``reconstruct`` is
``reconstructImageFromSharedPoints``.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; SPPM_AMD_LIB names another build of the same library
LIB_PATH = os.environ.get("SPPM_AMD_LIB", os.path.join(HERE, "libsppm_amd.so"))

MAT_DIFFUSE, MAT_METAL, MAT_ROUGH_CONDUCTOR, MAT_GGX_GLASS, MAT_GLASS = range(5)
MAT_BIMETAL, MAT_DIFFUSE_TRANSMIT, MAT_GLOSS_TRANSMIT = 5, 6, 7
CAM_PINHOLE, CAM_THINLENS = 0, 1
NEE_OFF, NEE_AS_WRITTEN, NEE_KEEP = 0, 1, 2  # sppm_params.eye_nee
SAMPLER_UNIFORM, SAMPLER_HALTON, SAMPLER_HAMMERSLEY = 0, 1, 2

EXPORTED = [
    "sppm_abi_version", "sppm_last_error", "sppm_create", "sppm_destroy", "sppm_scene_upload",
    "sppm_camera_set", "sppm_params_set", "sppm_begin", "sppm_pass", "sppm_pass_photons",
    "sppm_pass_build", "sppm_pass_eye", "sppm_pass_gather", "sppm_sync", "sppm_reconstruct",
    "sppm_read_photons", "sppm_read_hitpoints", "sppm_read_shared", "sppm_read_counts",
    "sppm_read_neighbours", "sppm_trace_closest", "sppm_kernel_ms", "sppm_counters",
    "sppm_read_path_counts", "sppm_scene_stats", "sppm_guide_set", "sppm_guide_weights", "sppm_guide_gradient",
    "sppm_guide_train", "sppm_guide_read", "sppm_graph_capture", "sppm_batch_create", "sppm_batch_add",
    "sppm_batch_capture", "sppm_batch_pass", "sppm_batch_sync", "sppm_batch_destroy", "sppm_autotune",
    "sppm_mt19937_jump", "sppm_checkpoint_size", "sppm_checkpoint_save", "sppm_checkpoint_load",
    "sppm_read_direct", "sppm_trace_occluded", "sppm_ray_counts", "sppm_guide_tables", "sppm_trace_closest_static",
    "sppm_tie_counts", "sppm_read_grid", "sppm_guide_path_cells", "sppm_camera_diaphragm", "sppm_batch_mode",
    "sppm_batch_get_mode",
]
TRACE_STATIC_MAX_RAYS = 65536  # SPPM_TRACE_STATIC_MAX_RAYS


def mt19937_jump(state, steps, lib=None):
    """Window of 624 raw mt19937 words `steps` words further on (host only)."""
    lib = lib or load_library()
    st = np.ascontiguousarray(state, dtype=np.uint32)
    assert st.shape == (624,)
    out = np.zeros(624, np.uint32)
    _check(lib.sppm_mt19937_jump(st.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_uint64(steps),
                                 out.ctypes.data_as(C.POINTER(C.c_uint32))))
    return out


class TuneResult(C.Structure):
    _fields_ = [("bvh_width", C.c_int32), ("grid_blocks", C.c_int32), ("ms_per_pass", C.c_float),
                ("trial_ms", C.c_float * 6), ("bvh_builder", C.c_int32), ("photon_kernel", C.c_int32),
                ("grid_ms", C.c_float * 3), ("kernel_ms", C.c_float * 2), ("photons_per_s", C.c_float * 3),
                ("photons_suggested", C.c_int32)]
GUIDE_CELLS = 262144
BVH_LBVH = 0x100  # bvh_width flag: radix tree over Morton codes instead of binned SAH
GUIDE_NPARAMS = 16705


class GuideParams(C.Structure):
    _fields_ = [("enable", C.c_int32), ("mix_plain", C.c_float), ("learning_rate", C.c_float),
                ("train_steps", C.c_int32), ("seed", C.c_uint32), ("history_decay", C.c_float),
                ("hidden", C.c_int32), ("layers", C.c_int32), ("fp16", C.c_int32), ("train_interval", C.c_int32),
                ("target", C.c_int32)]


GUIDE_TARGET_MEAN, GUIDE_TARGET_RMS = 0, 1


def guide_nparams(hidden=64, layers=4):
    """Parameter count of the proposal network (64 features -> hidden x layers -> 1)."""
    return 64 * hidden + hidden + (layers - 1) * (hidden * hidden + hidden) + hidden + 1


class Attr(C.Structure):
    _fields_ = [("kind", C.c_int32), ("index", C.c_float), ("if_delta", C.c_int32),
                ("if_opaque", C.c_int32), ("albedo", C.c_float * 3), ("roughness", C.c_float),
                ("radiance", C.c_float * 3), ("reflectance", C.c_float), ("extinction", C.c_float),
                ("is_sppm", C.c_int32), ("tex_albedo", C.c_int32), ("tex_roughness", C.c_int32),
                ("tex_bump", C.c_int32)]


class SceneInfo(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("nfacets", "nlights", "bvh_width", "bvh_nodes", "bvh_depth", "stack_lds",
                                         "stack_global")]


class Texture(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("rgba", C.POINTER(C.c_float))]


class SceneDesc(C.Structure):
    _fields_ = [("nfacets", C.c_int32), ("pos", C.POINTER(C.c_float)), ("nrm", C.POINTER(C.c_float)),
                ("uv", C.POINTER(C.c_float)), ("attr", C.POINTER(C.c_int32)), ("nattrs", C.c_int32),
                ("attrs", C.POINTER(Attr)), ("ntextures", C.c_int32), ("textures", C.POINTER(Texture)),
                ("has_ibl", C.c_int32), ("ibl_texture", C.c_int32)]


class CameraDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("pos", C.c_float * 3), ("target", C.c_float * 3), ("fov_deg", C.c_float),
                ("F", C.c_float), ("dist_to_lens", C.c_float), ("aspect", C.c_float),
                ("sampler", C.c_int32), ("sampler_id", C.c_int32), ("sampler_nsample", C.c_int32)]


class Params(C.Structure):
    _fields_ = [("alpha", C.c_float), ("first_radius", C.c_float), ("photons_per_pass", C.c_int32),
                ("bvh_width", C.c_int32), ("block_size", C.c_int32), ("check_numeric", C.c_int32),
                ("record_neighbours", C.c_int32), ("max_neighbours", C.c_int32), ("eye_nee", C.c_int32),
                ("step_offset", C.c_int32), ("photon_kernel", C.c_int32), ("wave_bounces", C.c_int32),
                ("wave_lanes", C.c_int32)]


class PassStats(C.Structure):
    _fields_ = [("photons_emitted", C.c_int32), ("photons_stored", C.c_int32), ("hit_points", C.c_int32),
                ("neighbours", C.c_int64), ("mean_radius", C.c_float), ("flags", C.c_int32),
                ("ms_photons", C.c_float), ("ms_build", C.c_float), ("ms_eye", C.c_float),
                ("ms_gather", C.c_float)]


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    lib.sppm_last_error.restype = C.c_char_p
    lib.sppm_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_void_p]
    lib.sppm_destroy.argtypes = [C.c_void_p]
    lib.sppm_scene_upload.argtypes = [C.c_void_p, C.POINTER(SceneDesc), C.c_int]
    lib.sppm_camera_set.argtypes = [C.c_void_p, C.POINTER(CameraDesc)]
    lib.sppm_camera_diaphragm.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_float)]
    lib.sppm_params_set.argtypes = [C.c_void_p, C.POINTER(Params)]
    lib.sppm_begin.argtypes = [C.c_void_p]
    lib.sppm_pass.argtypes = [C.c_void_p, C.POINTER(PassStats)]
    lib.sppm_pass_photons.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int32)]
    lib.sppm_pass_build.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    lib.sppm_pass_eye.argtypes = [C.c_void_p, C.c_int]
    lib.sppm_pass_gather.argtypes = [C.c_void_p]
    lib.sppm_sync.argtypes = [C.c_void_p]
    lib.sppm_reconstruct.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_float)]
    lib.sppm_read_photons.argtypes = [C.c_void_p, C.c_int32] + [C.POINTER(C.c_float)] * 5 + [C.POINTER(C.c_int32)]
    lib.sppm_read_hitpoints.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)] + [C.POINTER(C.c_float)] * 4
    lib.sppm_read_shared.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32)]
    lib.sppm_read_counts.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    lib.sppm_read_neighbours.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int64)]
    lib.sppm_trace_closest.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float,
                                       C.c_float, C.POINTER(C.c_int32)] + [C.POINTER(C.c_float)] * 3
    lib.sppm_trace_closest_static.argtypes = lib.sppm_trace_closest.argtypes
    lib.sppm_trace_occluded.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                        C.c_float, C.POINTER(C.c_float), C.c_float, C.POINTER(C.c_int32)]
    lib.sppm_kernel_ms.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_float)]
    lib.sppm_scene_stats.argtypes = [C.c_void_p, C.POINTER(SceneInfo)]
    lib.sppm_guide_set.argtypes = [C.c_void_p, C.POINTER(GuideParams)]
    lib.sppm_graph_capture.argtypes = [C.c_void_p]
    lib.sppm_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
    lib.sppm_batch_add.argtypes = [C.c_void_p, C.c_void_p]
    lib.sppm_batch_mode.argtypes = [C.c_void_p, C.c_int]
    lib.sppm_batch_get_mode.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    for n in ("sppm_batch_capture", "sppm_batch_pass", "sppm_batch_sync", "sppm_batch_destroy"):
        getattr(lib, n).argtypes = [C.c_void_p]
    lib.sppm_autotune.argtypes = [C.c_void_p, C.POINTER(SceneDesc), C.c_int, C.POINTER(TuneResult)]
    lib.sppm_mt19937_jump.argtypes = [C.POINTER(C.c_uint32), C.c_uint64, C.POINTER(C.c_uint32)]
    lib.sppm_checkpoint_size.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    lib.sppm_checkpoint_save.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    lib.sppm_checkpoint_load.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    lib.sppm_guide_weights.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.sppm_guide_gradient.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.sppm_guide_train.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_int]
    lib.sppm_guide_read.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.sppm_counters.argtypes = [C.c_void_p] + [C.POINTER(C.c_int64)] * 5
    lib.sppm_ray_counts.argtypes = [C.c_void_p] + [C.POINTER(C.c_int64)] * 2
    lib.sppm_tie_counts.argtypes = [C.c_void_p] + [C.POINTER(C.c_int64)] * 2
    lib.sppm_read_path_counts.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    lib.sppm_guide_path_cells.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    lib.sppm_read_grid.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.c_int64,
                                   C.POINTER(C.c_int32), C.c_int64, C.POINTER(C.c_int32)]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load_library()
    return _LIB


class SppmError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise SppmError(f"sppm error {rc}: {lib().sppm_last_error().decode()}")


def attr_records(attr_dicts):
    """Attribute dicts (SURFACE_PARAM fields; defaults as material.h:77-89
    except is_sppm, which a scene states) -> Attr array."""
    attrs = (Attr * len(attr_dicts))()
    for i, a in enumerate(attr_dicts):
        attrs[i].kind = a["kind"]
        attrs[i].index = a.get("index", 1.0)
        attrs[i].if_delta = int(a.get("if_delta", 0))
        attrs[i].if_opaque = int(a.get("if_opaque", 1))
        attrs[i].albedo[:] = list(a.get("albedo", (0.0, 0.0, 0.0)))
        attrs[i].roughness = a.get("roughness", 1.0)
        attrs[i].radiance[:] = list(a.get("radiance", (0.0, 0.0, 0.0)))
        attrs[i].reflectance = a.get("reflectance", 1.0)
        attrs[i].extinction = a.get("extinction", 1.0)
        attrs[i].is_sppm = int(a.get("is_sppm", 0))
        attrs[i].tex_albedo = a.get("tex_albedo", -1)
        attrs[i].tex_roughness = a.get("tex_roughness", -1)
        attrs[i].tex_bump = a.get("tex_bump", -1)
    return attrs


def scene_desc(scene):
    """Builds a SceneDesc from a scenes.Scene; returns (desc, keepalive)."""
    attrs = attr_records(scene.attrs)
    texs = (Texture * max(1, len(scene.textures)))()
    for i, t in enumerate(scene.textures):
        texs[i].height, texs[i].width = t.shape[0], t.shape[1]
        texs[i].rgba = _fp(t)
    d = SceneDesc()
    d.nfacets = scene.pos.shape[0]
    d.pos, d.nrm, d.uv, d.attr = _fp(scene.pos), _fp(scene.nrm), _fp(scene.uv), _ip(scene.attr)
    d.nattrs = len(scene.attrs)
    d.attrs = C.cast(attrs, C.POINTER(Attr))
    d.ntextures = len(scene.textures)
    d.textures = C.cast(texs, C.POINTER(Texture))
    ibl = getattr(scene, "ibl", None)
    d.has_ibl = 0 if ibl is None else 1
    d.ibl_texture = -1 if ibl is None else int(ibl)
    return d, (attrs, texs, scene)


def camera_desc(cam):
    c = CameraDesc()
    c.type = cam["type"]
    c.width, c.height = cam["width"], cam["height"]
    c.pos[:] = list(cam["pos"])
    c.target[:] = list(cam["target"])
    c.fov_deg, c.F, c.dist_to_lens, c.aspect = cam["fov_deg"], cam["F"], cam["dist_to_lens"], cam["aspect"]
    c.sampler = cam.get("sampler", SAMPLER_UNIFORM)
    c.sampler_id = cam.get("sampler_id", 0)
    c.sampler_nsample = cam.get("sampler_nsample", 1)
    return c


def diaphragm_rgba(img):
    """A thin-lens diaphragm image as (height, width, 4) float32 RGBA, pixel
    (x, y) at [y, x] (storage order x + y * width); only red is read."""
    a = np.ascontiguousarray(img, np.float32)
    if a.ndim != 3 or a.shape[2] != 4:
        raise ValueError("diaphragm image must be (height, width, 4) RGBA")
    return a


RECORD_OFF, RECORD_LISTS, RECORD_PAIRS = 0, 1, 2  # sppm_params.record_neighbours


PHOTON_MEGAKERNEL, PHOTON_WAVEFRONT = 0, 1  # sppm_params.photon_kernel


def params_struct(alpha, first_radius, photons, bvh_width=8, record=False, check_numeric=False, nee=0,
                  max_pairs=0, step_offset=0, photon_kernel=PHOTON_MEGAKERNEL, wave_bounces=0, wave_lanes=0):
    """record: False/0 none, True/1 lists from the recording gather kernel, 2
    pairs appended by the production gather (at most max_pairs per pass).
    step_offset: pass k runs the reference's step k + step_offset.
    photon_kernel / wave_bounces: how the photon pass runs (same photons)."""
    p = Params()
    p.photon_kernel, p.wave_bounces, p.wave_lanes = int(photon_kernel), int(wave_bounces), int(wave_lanes)
    p.alpha, p.first_radius, p.photons_per_pass = alpha, first_radius, photons
    p.bvh_width, p.block_size, p.check_numeric = bvh_width, 0, int(check_numeric)
    p.record_neighbours, p.max_neighbours = int(record), int(max_pairs)
    p.eye_nee = nee
    p.step_offset = int(step_offset)
    return p


class Renderer:
    """One SPPM render on one GPU: scene + camera + params, then passes."""

    def __init__(self, scene, cam, alpha, first_radius, photons, bvh_width=8, record=False, device=0, nee=NEE_OFF,
                 max_pairs=0, step_offset=0, photon_kernel=PHOTON_MEGAKERNEL, wave_bounces=0, wave_lanes=0):
        self.lib = lib()
        h = C.c_void_p()
        _check(self.lib.sppm_create(C.byref(h), device, None))
        self.h = h
        self.W, self.H = cam["width"], cam["height"]
        self.N = photons
        d, self._keep = scene_desc(scene)
        _check(self.lib.sppm_scene_upload(self.h, C.byref(d), bvh_width))
        self._cam = camera_desc(cam)
        _check(self.lib.sppm_camera_set(self.h, C.byref(self._cam)))
        if cam.get("diaphragm") is not None:
            self._dia = diaphragm_rgba(cam["diaphragm"])
            _check(self.lib.sppm_camera_diaphragm(self.h, self._dia.shape[1], self._dia.shape[0], _fp(self._dia)))
        self._params = params_struct(alpha, first_radius, photons, bvh_width & 0xff, record, nee=nee, max_pairs=max_pairs,
                                     step_offset=step_offset, photon_kernel=photon_kernel, wave_bounces=wave_bounces,
                                     wave_lanes=wave_lanes)
        _check(self.lib.sppm_params_set(self.h, C.byref(self._params)))
        _check(self.lib.sppm_begin(self.h))
        self.passes = 0
        self.total_photons = 0

    def upload_scene(self, scene, bvh_width=8):
        """Scene::CreateScene again: the render must be begun again (begin())."""
        d, keep = scene_desc(scene)
        _check(self.lib.sppm_scene_upload(self.h, C.byref(d), bvh_width))
        self._keep = keep

    def begin(self):
        """initializeSharedPoints: a new render with the current setup."""
        _check(self.lib.sppm_begin(self.h))
        self.passes = 0
        self.total_photons = 0

    def close(self):
        if self.h:
            self.lib.sppm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pass_(self, stats=True):
        st = PassStats() if stats else None
        _check(self.lib.sppm_pass(self.h, C.byref(st) if stats else None))
        self.passes += 1
        self.total_photons += self.N
        return st

    def photons(self, step):
        e = C.c_int32()
        _check(self.lib.sppm_pass_photons(self.h, step, C.byref(e)))
        return e.value

    def build(self):
        s = C.c_int32()
        _check(self.lib.sppm_pass_build(self.h, C.byref(s)))
        return s.value

    def eye(self, step):
        _check(self.lib.sppm_pass_eye(self.h, step))

    def gather(self):
        _check(self.lib.sppm_pass_gather(self.h))

    def sync(self):
        _check(self.lib.sppm_sync(self.h))

    def kernel_ms(self, name):
        v = C.c_float()
        _check(self.lib.sppm_kernel_ms(self.h, name.encode(), C.byref(v)))
        return v.value

    # ---- graph replay, autotuner --------------------------------------------
    def graph_capture(self):
        _check(self.lib.sppm_graph_capture(self.h))

    def autotune(self, scene, passes=2):
        d, keep = scene_desc(scene)
        res = TuneResult()
        _check(self.lib.sppm_autotune(self.h, C.byref(d), passes, C.byref(res)))
        self._keep = keep
        self.passes = 0
        self.total_photons = 0
        return dict(bvh_width=res.bvh_width, grid_blocks=res.grid_blocks, ms_per_pass=res.ms_per_pass,
                    trial_ms=list(res.trial_ms), bvh_builder=res.bvh_builder, photon_kernel=res.photon_kernel,
                    grid_ms=list(res.grid_ms), kernel_ms=list(res.kernel_ms),
                    photons_per_s=list(res.photons_per_s), photons_suggested=res.photons_suggested)

    # ---- learned emission proposal --------------------------------------
    def guide_set(self, enable=True, mix_plain=0.5, learning_rate=1e-3, train_steps=4, seed=1, history_decay=0.9,
                  hidden=64, layers=4, fp16=False, train_interval=1, target=GUIDE_TARGET_MEAN):
        g = GuideParams(int(enable), mix_plain, learning_rate, train_steps, seed, history_decay, hidden, layers,
                        int(fp16), int(train_interval), int(target))
        _check(self.lib.sppm_guide_set(self.h, C.byref(g)))
        self.guide_np = guide_nparams(hidden, layers)

    def guide_weights(self, new=None):
        out = np.zeros(self.guide_np, np.float32)
        src = None
        if new is not None:
            new = np.ascontiguousarray(new, np.float32)
            assert new.shape == (self.guide_np,)
            src = _fp(new)
        _check(self.lib.sppm_guide_weights(self.h, _fp(out), src))
        return out

    def guide_gradient(self, hist=None):
        grad = np.zeros(self.guide_np, np.float32)
        loss = C.c_float()
        hp = None
        if hist is not None:
            hist = np.ascontiguousarray(hist, np.float32)
            hp = _fp(hist)
        _check(self.lib.sppm_guide_gradient(self.h, hp, C.byref(loss), _fp(grad)))
        return loss.value, grad

    def guide_train(self, hist=None, steps=1):
        hp = None
        if hist is not None:
            hist = np.ascontiguousarray(hist, np.float32)
            hp = _fp(hist)
        _check(self.lib.sppm_guide_train(self.h, hp, steps))

    def guide_read(self):
        prob = np.zeros(GUIDE_CELLS, np.float32)
        hist = np.zeros(GUIDE_CELLS, np.float32)
        _check(self.lib.sppm_guide_read(self.h, _fp(prob), _fp(hist)))
        return prob, hist

    def guide_tables(self):
        """(prob, cdf) the next guided photon pass samples."""
        prob = np.zeros(GUIDE_CELLS, np.float32)
        cdf = np.zeros(GUIDE_CELLS, np.float32)
        _check(self.lib.sppm_guide_tables(self.h, _fp(prob), _fp(cdf)))
        return prob, cdf

    def guide_path_cells(self):
        """Emission cell of every photon path of the last guided pass."""
        cells = np.zeros(self.N, np.int32)
        _check(self.lib.sppm_guide_path_cells(self.h, _ip(cells)))
        return cells

    def checkpoint(self):
        """Bytes of a checkpoint of the render state (see sppm_checkpoint_save)."""
        n = C.c_int64()
        _check(self.lib.sppm_checkpoint_size(self.h, C.byref(n)))
        buf = np.empty(n.value, np.uint8)
        _check(self.lib.sppm_checkpoint_save(self.h, buf.ctypes.data_as(C.c_void_p), n))
        return buf, self.passes, self.total_photons

    def restore(self, ckpt):
        buf, passes, total = ckpt
        buf = np.ascontiguousarray(buf, np.uint8)
        _check(self.lib.sppm_checkpoint_load(self.h, buf.ctypes.data_as(C.c_void_p), C.c_int64(buf.size)))
        self.passes, self.total_photons = passes, total

    def scene_stats(self):
        info = SceneInfo()
        _check(self.lib.sppm_scene_stats(self.h, C.byref(info)))
        return {n: getattr(info, n) for n, _ in SceneInfo._fields_}

    def counters(self):
        v = [C.c_int64() for _ in range(5)]
        _check(self.lib.sppm_counters(self.h, *[C.byref(x) for x in v]))
        rays = [C.c_int64() for _ in range(4)]
        _check(self.lib.sppm_ray_counts(self.h, C.byref(rays[0]), C.byref(rays[1])))
        _check(self.lib.sppm_tie_counts(self.h, C.byref(rays[2]), C.byref(rays[3])))
        return dict(hit_points=v[0].value, neighbours=v[1].value, stored=v[2].value, candidates=v[3].value,
                    passes=v[4].value, photon_rays=rays[0].value, eye_rays=rays[1].value,
                    photon_tie_rays=rays[2].value, eye_tie_rays=rays[3].value)

    def read_photons(self):
        n = C.c_int32()
        _check(self.lib.sppm_read_photons(self.h, 0, None, None, None, None, None, C.byref(n)))
        n = n.value
        pos, flux, indir = (np.zeros((n, 3), np.float32) for _ in range(3))
        index, pdf = np.zeros(n, np.float32), np.zeros(n, np.float32)
        _check(self.lib.sppm_read_photons(self.h, n, _fp(pos), _fp(flux), _fp(indir), _fp(index), _fp(pdf),
                                          C.byref(C.c_int32())))
        return dict(pos=pos, flux=flux, indir=indir, index=index, pdf=pdf)

    def read_hitpoints(self):
        np_ = self.W * self.H
        t, f = np.zeros(np_, np.int32), np.zeros(np_, np.int32)
        pos, w, o, n = (np.zeros((np_, 3), np.float32) for _ in range(4))
        _check(self.lib.sppm_read_hitpoints(self.h, _ip(t), _ip(f), _fp(pos), _fp(w), _fp(o), _fp(n)))
        return dict(type=t, facet=f, pos=pos, weight=w, outdir=o, normal=n)

    def read_shared(self):
        np_ = self.W * self.H
        rad, r = np.zeros((np_, 3), np.float32), np.zeros(np_, np.float32)
        n, g = np.zeros(np_, np.int32), np.zeros(np_, np.int32)
        _check(self.lib.sppm_read_shared(self.h, _fp(rad), _fp(r), _ip(n), _ip(g)))
        return dict(radiance=rad, radius=r, num_photon=n, num_gather=g)

    def read_direct(self):
        np_ = self.W * self.H
        hd, sd = np.zeros((np_, 3), np.float32), np.zeros((np_, 3), np.float32)
        n = np.zeros(np_, np.int32)
        _check(self.lib.sppm_read_direct(self.h, _fp(hd), _fp(sd), _ip(n)))
        return dict(hit=hd, shared=sd, num=n)

    def read_path_counts(self):
        c = np.zeros(self.N, np.int32)
        _check(self.lib.sppm_read_path_counts(self.h, _ip(c)))
        return c

    def read_grid(self):
        """The photon grid of the last build: dims (nx, ny, nz, ncells,
        nphotons), geometry (bmin, cell edge), cell starts and the flattened
        photon index at each grid position."""
        dims, geom = np.zeros(5, np.int32), np.zeros(4, np.float32)
        _check(self.lib.sppm_read_grid(self.h, _ip(dims), _fp(geom), 0, None, 0, None))
        start = np.zeros(int(dims[3]) + 1, np.int32)
        order = np.zeros(max(int(dims[4]), 1), np.int32)
        _check(self.lib.sppm_read_grid(self.h, None, None, start.size, _ip(start), order.size, _ip(order)))
        return dict(nx=int(dims[0]), ny=int(dims[1]), nz=int(dims[2]), ncells=int(dims[3]),
                    nphotons=int(dims[4]), bmin=geom[:3].copy(), cell=np.float32(geom[3]), start=start,
                    order=order[:int(dims[4])])

    def read_counts(self):
        m = np.zeros(self.W * self.H, np.int32)
        _check(self.lib.sppm_read_counts(self.h, _ip(m)))
        return m

    def read_neighbours(self):
        tot = C.c_int64()
        _check(self.lib.sppm_read_neighbours(self.h, 0, None, None, C.byref(tot)))
        off = np.zeros(self.W * self.H + 1, np.int64)
        idx = np.zeros(max(1, tot.value), np.int32)
        _check(self.lib.sppm_read_neighbours(self.h, tot.value, off.ctypes.data_as(C.POINTER(C.c_int64)), _ip(idx),
                                             C.byref(tot)))
        return off, idx[: tot.value]

    def reconstruct(self, total_photons=None, passes=None):
        total = self.total_photons if total_photons is None else total_photons
        p = self.passes if passes is None else passes
        img = np.zeros((self.H, self.W, 4), np.float32)
        _check(self.lib.sppm_reconstruct(self.h, total, p, _fp(img)))
        return img

    def trace(self, org, dirs, tmin=1e-4, tmax=3.4028234663852886e38):
        org = np.ascontiguousarray(org, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = org.shape[0]
        f = np.zeros(n, np.int32)
        t, u, v = (np.zeros(n, np.float32) for _ in range(3))
        _check(self.lib.sppm_trace_closest(self.h, n, _fp(org), _fp(dirs), tmin, tmax, _ip(f), _fp(t), _fp(u), _fp(v)))
        return f, t, u, v

    def trace_static(self, org, dirs, tmin=1e-4, tmax=3.4028234663852886e38):
        """sppm_trace_closest_static: the same query through the fixed staging
        array (at most TRACE_STATIC_MAX_RAYS rays, no allocation)."""
        org = np.ascontiguousarray(org, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = org.shape[0]
        f = np.zeros(n, np.int32)
        t, u, v = (np.zeros(n, np.float32) for _ in range(3))
        _check(self.lib.sppm_trace_closest_static(self.h, n, _fp(org), _fp(dirs), tmin, tmax, _ip(f), _fp(t), _fp(u),
                                                  _fp(v)))
        return f, t, u, v

    def occluded(self, org, dirs, tmin, tmax, translucency=0.5):
        """Scene::IsOccluded per ray (Scene.cpp:246): 1 where a surface in
        [tmin, tmax) blocks it (tmax scalar or per ray), surfaces with
        1 - alpha > translucency passed."""
        org = np.ascontiguousarray(org, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = org.shape[0]
        tmax = np.ascontiguousarray(np.broadcast_to(np.float32(tmax) if np.isscalar(tmax) else tmax, (n,)), np.float32)
        out = np.zeros(n, np.int32)
        _check(self.lib.sppm_trace_occluded(self.h, n, _fp(org), _fp(dirs), tmin, _fp(tmax), translucency, _ip(out)))
        return out


BATCH_MULTI, BATCH_BRANCHES, BATCH_AUTO = 0, 1, 2  # sppm_batch_mode


class Batch:
    """One graph launch per pass over several Renderers (each begun, not yet
    stepped).  mode BATCH_MULTI: the photon, eye and gather-chunk kernels of
    all scenes run as one launch each; BATCH_BRANCHES: each scene's pass is
    its own graph branch; BATCH_AUTO picks one by the mean photons per pass."""

    def __init__(self, renderers, device=0, mode=BATCH_AUTO):
        self.lib = lib()
        h = C.c_void_p()
        _check(self.lib.sppm_batch_create(C.byref(h), device))
        self.h = h
        self.renderers = list(renderers)
        for r in self.renderers:
            _check(self.lib.sppm_batch_add(self.h, r.h))
        _check(self.lib.sppm_batch_mode(self.h, mode))
        _check(self.lib.sppm_batch_capture(self.h))
        m = C.c_int32()
        _check(self.lib.sppm_batch_get_mode(self.h, C.byref(m)))
        self.mode = m.value  # the mode in use (what BATCH_AUTO chose)

    def pass_(self):
        _check(self.lib.sppm_batch_pass(self.h))
        for r in self.renderers:
            r.passes += 1
            r.total_photons += r.N

    def sync(self):
        _check(self.lib.sppm_batch_sync(self.h))

    def close(self):
        if self.h:
            self.lib.sppm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
