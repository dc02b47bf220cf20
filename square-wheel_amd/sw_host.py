"""ctypes binding of libsw_host.so (the CPU renderer front-end, include/sw_host.h) and the seeded
workload recipes of the benchmark configurations."""
import ctypes as C
import math
import os

import numpy as np

import sw_abi as A

_lib = None


class MapConfig(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("grid", C.c_uint32 * 3), ("solid_probability", C.c_float),
                ("num_lights", C.c_uint32), ("light_intensity_min", C.c_float), ("light_intensity_max", C.c_float),
                ("sun", C.c_uint32), ("sky_fraction", C.c_float), ("texture_size", C.c_uint32),
                ("num_materials", C.c_uint32), ("normal_maps", C.c_uint32), ("specular_fraction", C.c_float),
                ("alpha_blend_fraction", C.c_float), ("directional_lightmaps", C.c_uint32),
                ("num_decal_textures", C.c_uint32), ("num_model_textures", C.c_uint32), ("num_submodels", C.c_uint32),
                ("num_view_portals", C.c_uint32)]


class MapInfo(C.Structure):
    _fields_ = [("num_polygons", C.c_uint32), ("num_leafs", C.c_uint32), ("num_nodes", C.c_uint32),
                ("num_portals", C.c_uint32), ("num_empty_cells", C.c_uint32), ("num_texels", C.c_uint64),
                ("num_lightmap_elements", C.c_uint64)]


class FrameConfig(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("color_bits", C.c_uint32), ("fov", C.c_float),
                ("textures_mip_bias", C.c_float), ("use_directional_lightmaps", C.c_uint32),
                ("clear_background", C.c_uint32), ("force_clear", C.c_uint32), ("num_dynamic_lights", C.c_uint32),
                ("num_models", C.c_uint32), ("num_decals", C.c_uint32), ("seed", C.c_uint32), ("time", C.c_float),
                ("virtual_threads", C.c_uint32), ("view_model", C.c_uint32),
                ("portals_depth", C.c_uint32), ("device_models", C.c_uint32)]


class ModelData(C.Structure):
    _fields_ = [("instances", C.c_void_p), ("num_instances", C.c_uint32), ("runs", C.c_void_p), ("num_runs", C.c_uint32),
                ("run_positions", C.c_void_p), ("run_num_draws", C.c_void_p), ("prepared_vertices", C.c_void_p),
                ("num_prepared_vertices", C.c_uint64), ("prepared_triangles", C.c_void_p),
                ("num_prepared_triangles", C.c_uint64), ("prepared_counts", C.c_void_p), ("num_prepared_counts", C.c_uint32),
                ("vertex_capacity", C.c_uint64), ("triangle_capacity", C.c_uint64)]


class MeshDataC(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32), ("num_frames", C.c_uint32),
                ("tc", C.c_void_p), ("frame_vertices", C.c_void_p), ("triangles", C.c_void_p)]


class ViewC(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("cam_pos", C.c_float * 3),
                ("clear_background", C.c_uint32), ("background_color", C.c_uint64), ("first_surface", C.c_uint32),
                ("num_surfaces", C.c_uint32), ("first_draw", C.c_uint32), ("num_draws", C.c_uint32),
                ("fb_offset", C.c_uint64), ("surface_arena_offset", C.c_uint64)]


class FrameData(C.Structure):
    _fields_ = [("view", ViewC), ("surfaces", C.c_void_p), ("num_surfaces", C.c_uint32), ("lights", C.c_void_p),
                ("num_lights", C.c_uint32), ("light_indices", C.c_void_p), ("num_light_indices", C.c_uint32),
                ("draws", C.c_void_p), ("num_draws", C.c_uint32), ("vertex_data", C.c_void_p),
                ("num_vertex_ints", C.c_uint64), ("surface_arena_texels", C.c_uint64),
                ("num_visible_leafs", C.c_uint32), ("num_triangles", C.c_uint32), ("first_submodel_surface", C.c_uint32),
                ("num_portal_texels", C.c_uint64), ("num_portal_views", C.c_uint32), ("num_visible_portals", C.c_uint32)]


assert C.sizeof(ViewC) == A.VIEW_DTYPE.itemsize

PORTAL_MIRROR, PORTAL_CAMERA, PORTAL_PARALLAX = 0, 1, 2
VIEW_PORTAL_DTYPE = np.dtype([("kind", "<u4"), ("blending_mode", "<u4"), ("plane", "<f4", 4),
                              ("tex_coord_equation", "<f4", (2, 4)), ("vertices", "<f4", (4, 3)),
                              ("camera_position", "<f4", 3), ("camera_yaw", "<f4"), ("camera_pitch", "<f4"),
                              ("camera_fov", "<f4"), ("transform", "<f4", 16), ("texture_level", "<u4"),
                              ("texture_blending_mode", "<u4"), ("texture_light_scale", "<f4"),
                              ("texture_light_add", "<f4", 3)])
PORTAL_INFO_DTYPE = np.dtype([("visible", "<u4"), ("resolution", "<u4", 2), ("tc_min", "<i4", 2), ("mip", "<u4"),
                              ("texture_offset", "<u8"), ("mix_level", "<u4"), ("mix_blend", "<u4"),
                              ("mix_light", "<f4", 3), ("reserved", "<u4")], align=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(A.HOST_LIB_PATH):
            raise RuntimeError(f"{A.HOST_LIB_PATH} is not built: run square-wheel_amd/build.py")
        lib = C.CDLL(A.HOST_LIB_PATH)
        lib.swh_map_create.argtypes = [C.POINTER(MapConfig), C.POINTER(C.c_void_p)]
        lib.swh_map_create.restype = C.c_int
        lib.swh_map_destroy.argtypes = [C.c_void_p]
        lib.swh_map_destroy.restype = None
        lib.swh_map_desc.argtypes = [C.c_void_p, C.POINTER(A.MapDesc)]
        lib.swh_map_desc.restype = C.c_int
        lib.swh_map_info_get.argtypes = [C.c_void_p, C.POINTER(MapInfo)]
        lib.swh_map_info_get.restype = C.c_int
        lib.swh_camera_path.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
        lib.swh_camera_path.restype = C.c_int
        lib.swh_frame_create.argtypes = [C.POINTER(C.c_void_p)]
        lib.swh_frame_create.restype = C.c_int
        lib.swh_frame_destroy.argtypes = [C.c_void_p]
        lib.swh_frame_destroy.restype = None
        lib.swh_frame_build.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(FrameConfig), C.POINTER(C.c_float), C.c_float,
                                        C.c_float]
        lib.swh_frame_build.restype = C.c_int
        lib.swh_frame_get.argtypes = [C.c_void_p, C.POINTER(FrameData)]
        lib.swh_frame_get.restype = C.c_int
        lib.swh_frame_portal_view.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p)]
        lib.swh_frame_portal_view.restype = C.c_int
        lib.swh_frame_portal_info.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
        lib.swh_frame_portal_info.restype = C.c_int
        lib.swh_map_view_portals.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
        lib.swh_map_view_portals.restype = C.c_int
        lib.swh_frame_model_data.argtypes = [C.c_void_p, C.POINTER(ModelData)]
        lib.swh_frame_model_data.restype = C.c_int
        lib.swh_map_model_mesh.argtypes = [C.c_void_p, C.POINTER(MeshDataC)]
        lib.swh_map_model_mesh.restype = C.c_int
        lib.swh_assemble_model_draws.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                                 C.c_uint64, C.c_void_p, C.c_uint64]
        lib.swh_assemble_model_draws.restype = C.c_int64
        lib.swh_shadow_create.argtypes = [C.POINTER(C.c_void_p)]
        lib.swh_shadow_create.restype = C.c_int
        lib.swh_shadow_destroy.argtypes = [C.c_void_p]
        lib.swh_shadow_destroy.restype = None
        lib.swh_shadow_clear.argtypes = [C.c_void_p]
        lib.swh_shadow_clear.restype = None
        lib.swh_shadow_add_cube.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_float), C.c_uint32]
        lib.swh_shadow_add_cube.restype = C.c_int
        lib.swh_shadow_add_projector.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float,
                                                 C.c_uint32, C.POINTER(C.c_float)]
        lib.swh_shadow_add_projector.restype = C.c_int
        lib.swh_shadow_get.argtypes = [C.c_void_p, C.POINTER(ShadowData)]
        lib.swh_shadow_get.restype = C.c_int
        lib.swh_pack_normal_roughness_from_rgb.argtypes = [C.c_uint32, C.c_float]
        lib.swh_pack_normal_roughness_from_rgb.restype = C.c_int32
        lib.swh_build_texture_mips.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
        lib.swh_build_texture_mips.restype = C.c_uint64
        lib.swh_map_save_sqwm.argtypes = [C.c_void_p, C.c_char_p]
        lib.swh_map_save_sqwm.restype = C.c_int
        lib.swh_map_load_sqwm.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        lib.swh_map_load_sqwm.restype = C.c_int
        lib.swh_map_save_quake_map.argtypes = [C.c_void_p, C.c_char_p]
        lib.swh_map_save_quake_map.restype = C.c_int
        lib.swh_last_error.argtypes = []
        lib.swh_last_error.restype = C.c_char_p
        lib.swh_compress_color.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_uint8)]
        lib.swh_compress_color.restype = None
        lib.swh_decompress_color.argtypes = [C.POINTER(C.c_uint8), C.POINTER(C.c_float)]
        lib.swh_decompress_color.restype = None
        lib.swh_compress_vector.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_int8), C.POINTER(C.c_uint8)]
        lib.swh_compress_vector.restype = None
        lib.swh_turb_shifts.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.c_void_p, C.c_uint32]
        lib.swh_turb_shifts.restype = C.c_uint32
        lib.swh_evaluate_function.argtypes = [C.c_uint32, C.POINTER(C.c_float), C.c_float]
        lib.swh_evaluate_function.restype = C.c_float
        lib.swh_layer_shift.argtypes = [C.c_float, C.c_uint32]
        lib.swh_layer_shift.restype = C.c_int32
        _lib = lib
    return _lib


def _copy(ptr, n, dtype):
    dtype = np.dtype(dtype)
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    buf = (C.c_char * (n * dtype.itemsize)).from_address(ptr)
    return np.frombuffer(bytes(buf), dtype=dtype).copy()


class Map:
    """A generated map (host arrays) and its upload description."""

    def __init__(self, **kw):
        lib = load()
        cfg = MapConfig()
        defaults = dict(seed=0, grid=(8, 8, 2), solid_probability=0.3, num_lights=4, light_intensity_min=0.5,
                        light_intensity_max=2.0, sun=0, sky_fraction=0.1, texture_size=64, num_materials=8,
                        normal_maps=0, specular_fraction=0.0, alpha_blend_fraction=0.0, directional_lightmaps=0,
                        num_decal_textures=0, num_model_textures=0, num_submodels=0, num_view_portals=0)
        defaults.update(kw)
        for k, v in defaults.items():
            if k == "grid":
                cfg.grid[:] = list(v)
            else:
                setattr(cfg, k, v)
        self.config = defaults
        h = C.c_void_p()
        if lib.swh_map_create(C.byref(cfg), C.byref(h)) != 0:
            raise RuntimeError("swh_map_create failed")
        self._attach(h)

    @classmethod
    def load_sqwm(cls, path):
        """A map loaded from a .sqwm file written by save_sqwm (swh_map_load_sqwm)."""
        lib = load()
        h = C.c_void_p()
        if lib.swh_map_load_sqwm(os.fsencode(path), C.byref(h)) != 0:
            raise RuntimeError("swh_map_load_sqwm: " + lib.swh_last_error().decode())
        m = cls.__new__(cls)
        m.config = None
        m._attach(h)
        return m

    def save_sqwm(self, path):
        lib = load()
        if lib.swh_map_save_sqwm(self.h, os.fsencode(path)) != 0:
            raise RuntimeError("swh_map_save_sqwm: " + lib.swh_last_error().decode())

    def save_quake_map(self, path):
        lib = load()
        if lib.swh_map_save_quake_map(self.h, os.fsencode(path)) != 0:
            raise RuntimeError("swh_map_save_quake_map: " + lib.swh_last_error().decode())

    def _attach(self, h):
        lib = load()
        self.h = h
        d = A.MapDesc()
        lib.swh_map_desc(self.h, C.byref(d))
        self.desc_ = d
        info = MapInfo()
        lib.swh_map_info_get(self.h, C.byref(info))
        self.info = {n: int(getattr(info, n)) for n, _ in MapInfo._fields_}
        # numpy copies of the map data (for the oracle and the GPU upload)
        self.data = A.MapData(
            _copy(d.texels, d.num_texels, A.TEXEL_DTYPE), _copy(d.lite_texels, d.num_lite_texels, np.uint32),
            _copy(d.lightmaps_simple, d.num_lightmap_elements * 3, np.float32),
            _copy(d.lightmaps_directional, d.num_directional_lightmap_elements, A.DIRLM_DTYPE),
            _copy(d.levels, d.num_levels, A.LEVEL_DTYPE), _copy(d.skybox_texels32, d.num_skybox_texels32, np.uint32),
            _copy(d.skybox_texels64, d.num_skybox_texels64, np.uint64))

    def view_portals(self):
        """The map's view portals (VIEW_PORTAL_DTYPE; transform column-major)."""
        lib = load()
        n = lib.swh_map_view_portals(self.h, None, 0)
        out = np.zeros(max(n, 0), dtype=VIEW_PORTAL_DTYPE)
        if n > 0:
            assert lib.swh_map_view_portals(self.h, C.c_void_p(out.ctypes.data), n) == n
        return out

    def model_mesh(self):
        """The map's animated model mesh (swh_map_model_mesh) as sw_model_mesh's VERTEX_ANIMATED data."""
        md = MeshDataC()
        if load().swh_map_model_mesh(self.h, C.byref(md)) != 0:
            raise RuntimeError("swh_map_model_mesh failed")
        nv, nt, nf = int(md.num_vertices), int(md.num_triangles), int(md.num_frames)
        tc = _copy(md.tc, nv * 2, np.float32).reshape(nv, 2)
        fv = _copy(md.frame_vertices, nf * nv * 6, np.float32).reshape(nf, nv, 6)
        tris = _copy(md.triangles, nt * 3, np.uint16).reshape(nt, 3)
        return A.ModelMeshData(A.MODEL_VERTEX_ANIMATED, tc, tris, frame_vertices=fv)

    def camera_path(self, seed, n):
        out = np.zeros((n, 5), dtype=np.float32)
        if load().swh_camera_path(self.h, seed, n, C.c_void_p(out.ctypes.data)) != 0:
            raise RuntimeError("swh_camera_path failed")
        return out

    def close(self):
        if self.h:
            load().swh_map_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _frame_dict(h, color_bits):
    """The command buffers of a built frame (copies), with the views seen through its portals."""
    lib = load()
    fd = FrameData()
    lib.swh_frame_get(h, C.byref(fd))
    portals = []
    i = 0
    info = np.zeros(1, dtype=PORTAL_INFO_DTYPE)
    while lib.swh_frame_portal_info(h, i, C.c_void_p(info.ctypes.data)) == 0:
        portals.append(info[0].copy())
        i += 1
    views = []
    for i in range(fd.num_portal_views):
        ch = C.c_void_p()
        if lib.swh_frame_portal_view(h, i, C.byref(ch)) != 0:
            raise RuntimeError("swh_frame_portal_view failed")
        views.append(_frame_dict(ch, color_bits))
    md = ModelData()
    if lib.swh_frame_model_data(h, C.byref(md)) != 0:
        raise RuntimeError("swh_frame_model_data failed")
    models = {
        "instances": _copy(md.instances, md.num_instances, A.MODEL_INSTANCE_DTYPE),
        "runs": _copy(md.runs, md.num_runs, A.MESH_RUN_DTYPE),
        "run_positions": _copy(md.run_positions, md.num_runs, np.uint32),
        "run_num_draws": _copy(md.run_num_draws, md.num_runs, np.uint32),
        "prepared_vertices": _copy(md.prepared_vertices, md.num_prepared_vertices, A.MODEL_VERTEX_DTYPE),
        "prepared_triangles": _copy(md.prepared_triangles, md.num_prepared_triangles * 3, np.uint16).reshape(-1, 3),
        "prepared_counts": _copy(md.prepared_counts, md.num_prepared_counts, np.uint32),
        "vertex_capacity": int(md.vertex_capacity),
        "triangle_capacity": int(md.triangle_capacity),
    }
    return {
        "models": models,
        "view": np.frombuffer(bytes(fd.view), dtype=A.VIEW_DTYPE).copy(),
        "surfaces": _copy(fd.surfaces, fd.num_surfaces, A.SURFACE_DTYPE),
        "lights": _copy(fd.lights, fd.num_lights, A.LIGHT_DTYPE),
        "light_indices": _copy(fd.light_indices, fd.num_light_indices, np.uint32),
        "draws": _copy(fd.draws, fd.num_draws, A.DRAW_DTYPE),
        "vertex_data": _copy(fd.vertex_data, fd.num_vertex_ints, np.int32),
        "arena_texels": int(fd.surface_arena_texels),
        "num_visible_leafs": int(fd.num_visible_leafs),
        "num_triangles": int(fd.num_triangles),
        "first_submodel_surface": int(fd.first_submodel_surface),
        "color_bits": color_bits,
        "portal_texels": int(fd.num_portal_texels),
        "num_visible_portals": int(fd.num_visible_portals),
        "portal_views": views,
        "portals": np.array(portals, dtype=PORTAL_INFO_DTYPE),
    }


class FrameBuilder:
    def __init__(self):
        h = C.c_void_p()
        if load().swh_frame_create(C.byref(h)) != 0:
            raise RuntimeError("swh_frame_create failed")
        self.h = h

    def build(self, m: Map, pose, **kw):
        cfg = FrameConfig()
        defaults = dict(width=320, height=200, color_bits=32, fov=math.pi / 2, textures_mip_bias=0.0,
                        use_directional_lightmaps=0, clear_background=1, force_clear=1, num_dynamic_lights=0,
                        num_models=0, num_decals=0, seed=0, time=0.0, virtual_threads=1, view_model=0,
                        portals_depth=2, device_models=0)
        defaults.update(kw)
        for k, v in defaults.items():
            setattr(cfg, k, v)
        cam = (C.c_float * 3)(float(pose[0]), float(pose[1]), float(pose[2]))
        lib = load()
        if lib.swh_frame_build(self.h, m.h, C.byref(cfg), cam, float(pose[3]), float(pose[4])) != 0:
            raise RuntimeError("swh_frame_build failed")
        return _frame_dict(self.h, int(cfg.color_bits))

    def close(self):
        if self.h:
            load().swh_frame_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ShadowData(C.Structure):
    _fields_ = [("targets", C.c_void_p), ("num_targets", C.c_uint32), ("draws", C.c_void_p), ("num_draws", C.c_uint32),
                ("vertex_data", C.c_void_p), ("num_vertex_ints", C.c_uint64), ("num_depth_texels", C.c_uint64)]


class ShadowBuilder:
    """Draw lists of cube shadow maps for sw_render_depth (swh_shadow_add_cube; the reference's
    PartialRenderer::build_shadow_maps cube branch with DepthRenderer::draw_map,
    partial_renderer.rs:541-566, depth_renderer.rs:47-287): six size x size sides per light, at
    consecutive texel offsets, in the order +x, -x, +y, -y, +z, -z."""

    def __init__(self):
        h = C.c_void_p()
        if load().swh_shadow_create(C.byref(h)) != 0:
            raise RuntimeError("swh_shadow_create failed")
        self.h = h

    def clear(self):
        load().swh_shadow_clear(self.h)

    def add_cube(self, m: Map, position, size):
        pos = (C.c_float * 3)(*(float(x) for x in position))
        if load().swh_shadow_add_cube(self.h, m.h, pos, int(size)) != 0:
            raise ValueError(f"swh_shadow_add_cube failed (size {size})")

    def add_projector(self, m: Map, position, rotation, fov, size):
        """One projector map (SW_DEPTH_CIRCLE target); rotation is the quaternion (s, x, y, z). Returns
        the map's basis (float32 [3, 3]: basis_x, basis_y, basis_z of LIGHT_SHADOW_DTYPE)."""
        pos = (C.c_float * 3)(*(float(x) for x in position))
        rot = (C.c_float * 4)(*(float(x) for x in rotation))
        basis = (C.c_float * 9)()
        if load().swh_shadow_add_projector(self.h, m.h, pos, rot, float(fov), int(size), basis) != 0:
            raise ValueError(f"swh_shadow_add_projector failed (size {size}, fov {fov})")
        return np.array(basis, np.float32).reshape(3, 3)

    def arrays(self):
        """(targets, draws, vertex_data, num_depth_texels): copies of the lists built so far."""
        sd = ShadowData()
        load().swh_shadow_get(self.h, C.byref(sd))
        return (_copy(sd.targets, sd.num_targets, A.DEPTH_TARGET_DTYPE), _copy(sd.draws, sd.num_draws, A.DEPTH_DRAW_DTYPE),
                _copy(sd.vertex_data, sd.num_vertex_ints, np.int32), int(sd.num_depth_texels))

    def close(self):
        if self.h:
            load().swh_shadow_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cube_shadows_for_lights(m: Map, lights, size):
    """Cube shadow maps for every light of a frame's light list: the draw lists and the matching
    LIGHT_SHADOW_DTYPE descriptors (light i's six sides at texel offset 6 * size * size * i)."""
    sb = ShadowBuilder()
    try:
        for lt in lights:
            sb.add_cube(m, lt["position"], size)
        targets, draws, vdata, n = sb.arrays()
    finally:
        sb.close()
    sh = np.zeros(len(lights), dtype=A.LIGHT_SHADOW_DTYPE)
    sh["kind"], sh["size"] = A.SHADOW_CUBE, size
    sh["offset"] = np.arange(len(lights), dtype=np.uint64) * 6 * size * size
    return targets, draws, vdata, n, sh


def batch_from_frames(frames):
    """Concatenates built frames into one BatchData (views get consecutive ranges)."""
    views = np.concatenate([f["view"] for f in frames]).copy()
    surf, lights, lidx, draws, vdata = [], [], [], [], []
    s_off = d_off = l_off = li_off = v_off = 0
    arena_off = fb_off = portal_off = 0
    for k, f in enumerate(frames):
        v = views[k]
        v["first_surface"], v["num_surfaces"] = s_off, len(f["surfaces"])
        v["first_draw"], v["num_draws"] = d_off, len(f["draws"])
        v["fb_offset"], v["surface_arena_offset"] = fb_off, arena_off
        s = f["surfaces"].copy()
        s["first_light"] += li_off
        li = f["light_indices"] + l_off
        d = f["draws"].copy()
        d["first_vertex_int"] += v_off
        d["texture_offset"][d["texture_source"] == A.TEXSRC_PORTAL] += np.uint64(portal_off)
        surf.append(s)
        lights.append(f["lights"])
        lidx.append(li.astype(np.uint32))
        draws.append(d)
        vdata.append(f["vertex_data"])
        s_off += len(s)
        d_off += len(d)
        l_off += len(f["lights"])
        li_off += len(li)
        v_off += len(f["vertex_data"])
        arena_off += f["arena_texels"]
        fb_off += int(v["width"]) * int(v["height"])
        portal_off += f.get("portal_texels", 0)
    cat = lambda xs, dt: np.concatenate(xs) if xs else np.zeros(0, dtype=dt)  # noqa: E731
    return A.BatchData(views, cat(surf, A.SURFACE_DTYPE), cat(lights, A.LIGHT_DTYPE), cat(lidx, np.uint32),
                       cat(draws, A.DRAW_DTYPE), cat(vdata, np.int32), frames[0]["color_bits"], arena_texels=arena_off,
                       portal_texels=portal_off)


def portal_levels(frame):
    """The frame and the views seen through its portals, level by level: level k + 1 lists the portal
    views of level k's frames in order, so that level k + 1's framebuffers laid end to end (one batch)
    are level k's portal textures laid end to end. The last level's portal draws sample zeros."""
    levels = [[frame]]
    while True:
        nxt = [c for f in levels[-1] for c in f.get("portal_views", [])]
        if not nxt:
            return levels
        tex = sum(f["portal_texels"] for f in levels[-1])
        px = sum(int(c["view"]["width"][0]) * int(c["view"]["height"][0]) for c in nxt)
        assert tex == px, (tex, px)
        levels.append(nxt)


def assemble_model_draws(frame, fans, int_base):
    """The draw list of a frame built with device_models (swh_assemble_model_draws): the frame's draws
    with, at each mesh run's position, the triangles of its fans (sw_expand_mesh_runs' fan sizes
    `fans`), triangle t's fan in order, t in order; their vertex data sits at vertex int int_base +
    the run's layout."""
    md = frame["models"]
    runs = np.ascontiguousarray(md["runs"], dtype=A.MESH_RUN_DTYPE)
    pos = np.ascontiguousarray(md["run_positions"], dtype=np.uint32)
    draws = np.ascontiguousarray(frame["draws"], dtype=A.DRAW_DTYPE)
    fans = np.ascontiguousarray(fans, dtype=np.uint8)
    cap = len(draws) + int(fans.astype(np.int64).sum())
    out = np.zeros(max(cap, 1), dtype=A.DRAW_DTYPE)
    n = load().swh_assemble_model_draws(A._ptr(draws), len(draws), A._ptr(runs) if runs.size else None,
                                        A._ptr(pos) if pos.size else None, len(runs), A._ptr(fans) if fans.size else None,
                                        int(int_base), A._ptr(out), len(out))
    if n < 0:
        raise ValueError("swh_assemble_model_draws: run positions or fan sizes out of range")
    return out[:n]


class DeviceModels:
    """Frames built with device_models drawn with their models prepared on the device: the map's
    mesh stays in HBM; per frame sw_prepare_models (animation, lighting, back faces, sort), then
    sw_expand_mesh_runs (draw_mesh's clipping and fans) into the frame's vertex data in HBM, the
    runs' triangles put into the draw list at their places, and one sw_render_batch. The device
    buffers grow as needed and are kept between frames."""

    def __init__(self, ctx, m: Map):
        self.ctx = ctx
        self.mesh = m.model_mesh()
        self.mesh_desc, self.mesh_bufs = ctx.upload_meshes([self.mesh])
        self.bufs = {}

    def _buf(self, name, nbytes):
        b = self.bufs.get(name)
        if b is None or b.nbytes < nbytes:
            if b is not None:
                b.free()
            b = self.bufs[name] = self.ctx.alloc(max(nbytes, 256))
        return b

    def _upload(self, name, a):
        b = self._buf(name, a.nbytes)
        if a.nbytes:
            self.ctx.check(self.ctx.lib.sw_copy_to_device(self.ctx.h, b.ptr, A._ptr(a), a.nbytes), "sw_copy_to_device")
        return b

    def render(self, map_h, frame, out=None):
        """Returns (framebuffer, fans); `out`: a device framebuffer to draw into (kept, nothing read
        back), else the framebuffer is read back."""
        ctx = self.ctx
        md = frame["models"]
        inst = np.ascontiguousarray(md["instances"])
        vcap, tcap = max(md["vertex_capacity"], 1), max(md["triangle_capacity"], 1)
        host_ints = len(frame["vertex_data"])
        _, run_ints, _ = A.mesh_run_layout(md["runs"])
        vbuf = self._buf("vertices", vcap * A.MODEL_VERTEX_DTYPE.itemsize)
        tbuf = self._buf("triangles", tcap * 6)
        cbuf = self._buf("counts", max(len(inst), 1) * 4)
        ints = self._buf("vertex_data", (host_ints + run_ints) * 4)
        if len(inst):
            ctx.check(ctx.lib.sw_prepare_models(ctx.h, self.mesh_desc, 1, A._ptr(inst), len(inst), vbuf.ptr, vcap, tbuf.ptr,
                                                tcap, cbuf.ptr), "sw_prepare_models")
        vd = np.ascontiguousarray(frame["vertex_data"], dtype=np.int32)
        if host_ints:
            # queued: sw_expand_mesh_runs below waits for it
            ctx.check(ctx.lib.sw_copy_to_device_async(ctx.h, ints.ptr, A._ptr(vd), host_ints * 4), "sw_copy_to_device_async")
        fans = ctx.expand_mesh_runs(md["runs"], vbuf, vcap, tbuf, tcap, cbuf, len(inst),
                                    (ints.ptr.value or 0) + host_ints * 4, run_ints)
        draws = assemble_model_draws(frame, fans, host_ints)
        # the frame's command buffers, queued behind the expansion (kept alive until the render
        # call below has waited for them)
        keep = {"surfaces": np.ascontiguousarray(frame["surfaces"]), "lights": np.ascontiguousarray(frame["lights"]),
                "light_indices": np.ascontiguousarray(frame["light_indices"], dtype=np.uint32), "draws": draws}
        ptrs = {}
        for k, a in keep.items():
            buf = self._buf(k, a.nbytes)
            if a.nbytes:
                ctx.check(ctx.lib.sw_copy_to_device_async(ctx.h, buf.ptr, A._ptr(a), a.nbytes), "sw_copy_to_device_async")
            ptrs[k] = buf.ptr
        view = np.array(frame["view"], dtype=A.VIEW_DTYPE).reshape(1)
        view["first_surface"] = 0
        view["num_surfaces"] = len(keep["surfaces"])
        view["first_draw"] = 0
        view["num_draws"] = len(draws)
        view["fb_offset"] = 0
        view["surface_arena_offset"] = 0
        bits = int(frame["color_bits"])
        fb_texels = int(view["width"][0]) * int(view["height"][0])
        cb = A.Batch()
        cb.views, cb.num_views, cb.color_bits = A._ptr(view), 1, bits
        cb.surfaces, cb.num_surfaces = ptrs["surfaces"], len(keep["surfaces"])
        cb.lights, cb.num_lights = ptrs["lights"], len(keep["lights"])
        cb.light_indices, cb.num_light_indices = ptrs["light_indices"], len(keep["light_indices"])
        cb.draws, cb.num_draws = ptrs["draws"], len(draws)
        cb.vertex_data, cb.num_vertex_ints = ints.ptr, host_ints + run_ints
        cb.surface_arena_texels = int(frame["arena_texels"])
        cb.framebuffer_texels = fb_texels
        cb.inputs_on_device = 1
        if out is not None:
            ctx.check(ctx.lib.sw_render_batch(ctx.h, map_h, C.byref(cb), out.ptr), "sw_render_batch")
            return None, fans
        dtype = np.uint64 if bits == 64 else np.uint32
        fb = self._buf("framebuffer", max(fb_texels, 1) * np.dtype(dtype).itemsize)
        ctx.check(ctx.lib.sw_render_batch(ctx.h, map_h, C.byref(cb), fb.ptr), "sw_render_batch")
        return fb.read(dtype, fb_texels), fans

    def close(self):
        for b in list(self.mesh_bufs) + list(self.bufs.values()):
            b.free()
        self.mesh_bufs = []
        self.bufs = {}


def portal_level_batches(frame):
    """One BatchData per portal level (portal_levels), level 0 (the frame itself) first."""
    return [batch_from_frames(level) for level in portal_levels(frame)]


def portal_level_mixes(frame):
    """The portal textures' overlays per portal level (sw_mix_portal_textures' input): entry k (k >= 1)
    lists, as PORTAL_MIX_DTYPE, the overlays of the portals of level k - 1's frames, with their rects
    placed in the portal textures level k - 1 samples (level k's framebuffers laid end to end; for
    the last level, whose portals have no views, the black textures below it); they are mixed after
    level k is drawn and before level k - 1 is. Entry 0 is empty; there are len(levels) + 1 entries."""
    levels = portal_levels(frame)
    out = [np.zeros(0, dtype=A.PORTAL_MIX_DTYPE)]
    for k in range(1, len(levels) + 1):
        rows = []
        base = 0
        for f in levels[k - 1]:
            for p in f["portals"]:
                if p["visible"] and int(p["mix_level"]) != A.NO_TEXTURE:
                    m = np.zeros(1, dtype=A.PORTAL_MIX_DTYPE)[0]
                    m["dst_offset"] = base + int(p["texture_offset"])
                    m["size"] = p["resolution"]
                    m["tc_start"] = p["tc_min"]
                    m["level"] = p["mix_level"]
                    m["blend"] = p["mix_blend"]
                    m["light"] = p["mix_light"]
                    rows.append(m)
            base += int(f["portal_texels"])
        out.append(np.array(rows, dtype=A.PORTAL_MIX_DTYPE))
    return out


# ---------------------------------------------------------------------------------------------
# Workload recipes of the benchmark configurations
# ---------------------------------------------------------------------------------------------
def small_map(seed=0):
    """The 320x200 CPU-scale scene: 8x8 grid (4 layers, so that two interior layers remain inside the
    always-solid perimeter), p=0.3, 4 white lights, 64x64 textures, simple lightmaps."""
    return Map(seed=seed, grid=(8, 8, 4), solid_probability=0.3, num_lights=4, texture_size=64, num_materials=6,
               normal_maps=0, specular_fraction=0.0, directional_lightmaps=0)


def large_map(seed=0, grid=(32, 32, 4), **kw):
    """The 1080p scene: 32x32x4 grid, p=0.35, 32 lights + sun through sky faces, 256x256 textures with
    normal maps, a third of the materials specular (half of those metal), directional lightmaps."""
    args = dict(seed=seed, grid=grid, solid_probability=0.35, num_lights=32, sun=1, sky_fraction=0.1,
                texture_size=256, num_materials=12, normal_maps=1, specular_fraction=1.0 / 3.0,
                directional_lightmaps=1, num_decal_textures=4, num_model_textures=4)
    args.update(kw)
    return Map(**args)


def large_map_alpha(seed=0, **kw):
    """The 1080p scene with 10% alpha-blended materials (the HDR + models scene)."""
    return large_map(seed, alpha_blend_fraction=0.1, **kw)


def plain_1080p_frames(m, seed=0, n=64, width=1920, height=1080, fbld=None):
    """The plain drop-in workload: an n-pose seeded camera path through map m (camera_path(seed + 1)),
    each pose a Color32 frame with directional lightmaps that clears its background."""
    fbld = fbld or FrameBuilder()
    return [fbld.build(m, p, width=width, height=height, color_bits=32, use_directional_lightmaps=1, clear_background=1,
                       force_clear=1) for p in m.camera_path(seed + 1, n)]


def hdr_models_frames(m, seed=0, n=16, width=2560, height=1440, fbld=None):
    """The HDR + models workload on the alpha map: Color64 frames with 16 dynamic lights, 64 animated
    models and 256 decals along an n-pose path (camera_path(seed + 3)), model time 0.1 s per pose."""
    fbld = fbld or FrameBuilder()
    return [fbld.build(m, p, width=width, height=height, color_bits=64, use_directional_lightmaps=1, num_dynamic_lights=16,
                       num_models=64, num_decals=256, seed=seed + 5, time=0.1 * i, clear_background=1, force_clear=1)
            for i, p in enumerate(m.camera_path(seed + 3, n))]


def batch_view_frames(m, seed=0, n_views=256, fbld=None, color_bits=32):
    """The batched workload: n_views views of map m with independent seeded poses (camera_path(seed + 2)),
    widths log-uniform in [160, 1920], each view 16:9 or 4:3 by a seeded coin, capped at 1920x1080."""
    rng = np.random.default_rng(seed + 17)
    poses = m.camera_path(seed + 2, n_views)
    fbld = fbld or FrameBuilder()
    frames = []
    for i in range(n_views):
        wid = int(round(math.exp(rng.uniform(math.log(160), math.log(1920)))))
        hei = max(1, int(round(wid * (9 / 16 if rng.uniform() < 0.5 else 3 / 4))))
        wid, hei = min(wid, 1920), min(hei, 1080)
        frames.append(fbld.build(m, poses[i], width=wid, height=hei, color_bits=color_bits, use_directional_lightmaps=1,
                                 clear_background=1, force_clear=1))
    return frames


SERVING_GRIDS = [(16, 16, 4), (24, 24, 4), (32, 32, 4), (48, 48, 4)]


def serving_maps(seed=0):
    """The serving workload's four resident maps (the 1080p recipe with seeds seed..seed+3)."""
    return [large_map(seed + i, grid=g) for i, g in enumerate(SERVING_GRIDS)]


def serving_request_pool(maps, seed=0, per_map=12, fbld=None):
    """Requests of the serving workload: per map a seeded path of poses (camera_path(seed + 40 + i)),
    each a frame of log-uniform width in [320, 2560] at 16:10, Color32 or Color64 by a seeded coin,
    4 dynamic lights. Returns [(map index, frame)]."""
    rng = np.random.default_rng(seed + 77)
    fbld = fbld or FrameBuilder()
    pool = []
    for i, m in enumerate(maps):
        for p in m.camera_path(seed + 40 + i, per_map):
            w = int(round(math.exp(rng.uniform(math.log(320), math.log(2560)))))
            h = max(1, int(round(w * 10 / 16)))
            bits = 64 if rng.uniform() < 0.5 else 32
            pool.append((i, fbld.build(m, p, width=w, height=h, color_bits=bits, use_directional_lightmaps=1,
                                       num_dynamic_lights=4, clear_background=1, force_clear=1)))
    return pool


class TurbShifts:
    """make_turb_distortion's shift buffer for one effect instance (swh_turb_shifts), kept between
    calls like the reference's temp buffer."""

    def __init__(self):
        self.buf = np.zeros(1 << 16, dtype=np.int32)
        self.len = 0

    def compute(self, amplitude, wave_length, frequency, time_s, width, height, mip):
        self.len = int(load().swh_turb_shifts(amplitude, wave_length, frequency, time_s, width, height, mip,
                                              C.c_void_p(self.buf.ctypes.data), self.len))
        return self.buf[:self.len].copy()


def evaluate_function(kind, params, x):
    """SingleArgumentFunction::evaluate (swh_evaluate_function)."""
    p = (C.c_float * 4)(*(list(params) + [0.0] * (4 - len(params))))
    return float(load().swh_evaluate_function(kind, p, x))


def layer_shift(value, mip):
    return int(load().swh_layer_shift(value, mip))
