"""ctypes bindings of the two native libraries (include/fooska.h,
include/fooska_mesh.h).  Loading fails loudly: there is no Python fallback
for any part of the finite-volume step."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_int64_p = ctypes.POINTER(ctypes.c_int64)

FK_OK = 0
EINVAL = -1  # FK_EINVAL (include/fooska_mesh.h)
MODEL_SINGLE = 0
MODEL_TWO_PHASE = 1
FP64 = 0
MIXED = 1
RELAX_PLAIN = 0
RELAX_MLP = 1
RELAX_COLD = 2


class FkParams(ctypes.Structure):
    """Mirror of fk_params (include/fooska.h): the reference INPUT struct."""
    _fields_ = [(n, ctypes.c_int32) for n in (
        "model", "precision", "rkSteps", "adaptDt", "adaptDtItr", "viscFlag", "turbulence",
        "heatFlux", "woodsFlag", "neglectX", "neglectY", "neglectZ", "chemFlag", "thermItrMax",
        "chemItrMax", "hllcEnergy", "relaxFlag", "relaxMode")] + [(n, ctypes.c_double) for n in (
        "dt", "cfl", "maxDt", "avfMin", "thermUrf", "chemUrf", "chemLimit", "satTempUrf",
        "polishTol")] + [(n, ctypes.c_double * 2) for n in (
        "gamma", "pInf", "Cp", "eta", "satConst", "dynVisc", "Pr")] + [
        ("thermFntRlxFlag", ctypes.c_int32), ("thermRlxRate", ctypes.c_double),
        ("inletP", ctypes.c_double), ("inletT", ctypes.c_double), ("pRef", ctypes.c_double)]

    @classmethod
    def from_dict(cls, d):
        p = cls()
        for k, v in d.items():
            if isinstance(getattr(p, k), ctypes.Array):
                arr = getattr(p, k)
                for i, x in enumerate(v):
                    arr[i] = x
            else:
                setattr(p, k, v)
        return p

    def to_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if isinstance(v, ctypes.Array) else v
        return out


class FkRelaxStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("n_therm", "n_chem", "newton_iters", "tsat_iters", "n_fallback",
                                               "n_warm")]


class FkmPhaseEos(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double * 2) for n in ("gamma", "pinf", "cp", "eta")]


class FkmRegion(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("c", ctypes.c_double * 3), ("r", ctypes.c_double),
                ("alphaL", ctypes.c_double), ("p", ctypes.c_double), ("T", ctypes.c_double),
                ("u", ctypes.c_double * 3)]


def _load(name):
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run python -m fooskagpu_amd.build")
    return ctypes.CDLL(path)


_mesh = None
_hip = None

# (name, restype, argtypes) of the C ABI
MESH_API = [
    ("fkm_last_error", ctypes.c_char_p, []),
    ("fkm_box", ctypes.c_int, [ctypes.c_int] * 3 + [c_double_p, c_double_p, ctypes.c_double, ctypes.c_uint64,
                               ctypes.c_int, c_int32_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("fkm_box_kinds", ctypes.c_int, [ctypes.c_int] * 3 + [c_double_p, c_double_p, ctypes.c_double, ctypes.c_uint64,
                                                         c_int32_p, ctypes.c_int, c_int32_p,
                                                         ctypes.POINTER(ctypes.c_void_p)]),
    ("fkm_read_neu", ctypes.c_int, [ctypes.c_char_p, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("fkm_write_neu", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    ("fkm_free", None, [ctypes.c_void_p]),
    ("fkm_sizes", ctypes.c_int, [ctypes.c_void_p, c_int64_p, c_int64_p, c_int64_p, c_int64_p]),
    ("fkm_array", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    ("fkm_init_sod", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, c_double_p,
                                    c_double_p, c_double_p]),
    ("fkm_init_tgv", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_double] * 4 + [c_double_p]),
    ("fkm_init_regions", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(FkmPhaseEos), ctypes.POINTER(FkmRegion),
                                        ctypes.c_int, c_double_p]),
    ("fkm_place_bubbles", ctypes.c_int, [ctypes.c_int, c_double_p, c_double_p, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_uint64, c_double_p, c_double_p]),
]

_ctx = ctypes.c_void_p
HIP_API = [
    ("fk_last_error", ctypes.c_char_p, []),
    ("fk_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(FkParams), ctypes.POINTER(ctypes.c_void_p)]),
    ("fk_destroy", ctypes.c_int, [_ctx]),
    ("fk_set_params", ctypes.c_int, [_ctx, ctypes.POINTER(FkParams)]),
    ("fk_upload_mesh", ctypes.c_int, [_ctx, ctypes.c_int64, ctypes.c_int64, c_int32_p, c_int32_p, c_double_p,
                                      c_double_p, c_double_p, c_double_p, c_double_p, c_double_p]),
    ("fk_upload_state", ctypes.c_int, [_ctx, c_double_p]),
    ("fk_download_state", ctypes.c_int, [_ctx, c_double_p]),
    ("fk_init_rk", ctypes.c_int, [_ctx]),
    ("fk_compute_face_values", ctypes.c_int, [_ctx]),
    ("fk_compute_rhs", ctypes.c_int, [_ctx]),
    ("fk_compute_time_step", ctypes.c_int, [_ctx, c_double_p]),
    ("fk_update", ctypes.c_int, [_ctx, ctypes.c_int, ctypes.c_double]),
    ("fk_compute_rhs_update", ctypes.c_int, [_ctx, ctypes.c_int, ctypes.c_double]),
    ("fk_relaxation", ctypes.c_int, [_ctx, ctypes.c_int, ctypes.POINTER(FkRelaxStats)]),
    ("fk_relax_totals", ctypes.c_int, [_ctx, ctypes.POINTER(FkRelaxStats), ctypes.c_int]),
    ("fk_cons_calc", ctypes.c_int, [_ctx, c_double_p]),
    ("fk_cons_calc_noalloc", ctypes.c_int, [_ctx, c_double_p]),
    ("fk_run_steps", ctypes.c_int, [_ctx, ctypes.c_int, c_double_p, c_double_p]),
    ("fk_health", ctypes.c_int, [_ctx, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64), ctypes.c_int]),
    ("fk_sync", ctypes.c_int, [_ctx]),
    ("fk_get_clock", ctypes.c_int, [_ctx, c_double_p, c_double_p, ctypes.POINTER(ctypes.c_int64)]),
    ("fk_set_clock", ctypes.c_int, [_ctx, ctypes.c_double, ctypes.c_double, ctypes.c_int64]),
    ("fk_download", ctypes.c_int, [_ctx, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int64]),
    ("fk_set_debug", ctypes.c_int, [_ctx, ctypes.c_int]),
    ("fk_set_tuning", ctypes.c_int, [_ctx, ctypes.c_char_p, ctypes.c_int64]),
    ("fk_stream_copy", ctypes.c_int, [_ctx, ctypes.c_int64, c_double_p]),
    ("fk_mlp_set_weights", ctypes.c_int, [_ctx, ctypes.c_void_p, ctypes.c_void_p]),
    ("fk_mlp_get_weights", ctypes.c_int, [_ctx, ctypes.c_void_p, ctypes.c_void_p]),
    ("fk_mlp_set_norm", ctypes.c_int, [_ctx, ctypes.c_void_p]),
    ("fk_mlp_label", ctypes.c_int, [_ctx, c_double_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
    ("fk_mlp_data", ctypes.c_int, [_ctx, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int]),
    ("fk_mlp_emit", ctypes.c_int, [_ctx, ctypes.c_int64, ctypes.c_int64]),
    ("fk_mlp_download_data", ctypes.c_int, [_ctx, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, c_int64_p]),
    ("fk_mlp_grad", ctypes.c_int, [_ctx, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                   ctypes.POINTER(ctypes.c_float)]),
    ("fk_mlp_train", ctypes.c_int, [_ctx, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64,
                                    ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    ("fk_mlp_infer", ctypes.c_int, [_ctx, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
    ("fk_set_timing", ctypes.c_int, [_ctx, ctypes.c_int]),
    ("fk_kernel_times", ctypes.c_int, [_ctx, ctypes.c_int, ctypes.c_void_p, c_int64_p, c_double_p]),
]


def _bind(lib, api):
    for name, res, args in api:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def mesh_lib():
    global _mesh
    if _mesh is None:
        _mesh = _bind(_load("libfooska_mesh.so"), MESH_API)
    return _mesh


def hip_lib():
    global _hip
    if _hip is None:
        _hip = _bind(_load("libfooska_hip.so"), HIP_API)
    return _hip


class FkError(RuntimeError):
    pass


def check_mesh(rc):
    if rc != FK_OK:
        raise FkError(f"fooska mesh error {rc}: {mesh_lib().fkm_last_error().decode()}")
    return rc


def check_hip(rc):
    if rc < 0:
        raise FkError(f"fooska error {rc}: {hip_lib().fk_last_error().decode()}")
    return rc


def dptr(a):
    return a.ctypes.data_as(c_double_p)


def iptr(a):
    return a.ctypes.data_as(c_int32_p)
