"""ctypes binding of the C ABI in include/oslmi.h (liboslmi.so).

The HIP library is the only compute path: if it has not been built, or no
gfx950 GPU is present, calls raise instead of falling back to anything.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "_lib", "liboslmi.so")

_lib = None


class OSLMIError(RuntimeError):
    pass


def _declare(L):
    P, I, I64, C = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_char_p
    sigs = {
        "oslmi_version": ([], C),
        "oslmi_last_error": ([], C),
        "oslmi_init": ([], I),
        "oslmi_noise": ([C, I, I, I, P, P, I64, P], I),
        "oslmi_pnoise": ([C, I, I, I, P, P, P, I64, P], I),
        "oslmi_hash": ([I, P, P, I64, P], I),
        "oslmi_shade_fbm_cell": ([I64, P, P, P], I),
        "oslmi_shade_layered3": ([I64, P, P, P, P, P, P, P, P, P], I),
        "oslmi_shadingsys_create": ([C, C], P),
        "oslmi_shadingsys_destroy": ([P], None),
        "oslmi_shadingsys_attribute": ([P, C, C], I),
        "oslmi_renderer_name_transform": ([P, C, P, P], I),
        "oslmi_renderer_set_attribute": ([P, C, C, P], I),
        "oslmi_load_memory_compiled_shader": ([P, C, C], I),
        "oslmi_shader_group_begin": ([P, C], P),
        "oslmi_parameter": ([P, P, C, C, P, I], I),
        "oslmi_shader": ([P, P, C, C, C], I),
        "oslmi_connect_shaders": ([P, P, C, C, C, C], I),
        "oslmi_shader_group_end": ([P, P], I),
        "oslmi_group_release": ([P], None),
        "oslmi_group_outputs": ([P, P, I, P], I),
        "oslmi_jit_group": ([P, P], I),
        "oslmi_compile_group": ([P, P], I),
        "oslmi_group_emit": ([P, P, P, I64], I),
        "oslmi_group_output_planes": ([P, P], I),
        "oslmi_group_jit_seconds": ([P], ctypes.c_double),
        "oslmi_execute_batch": ([P, P, P, P, P, P, P, P], I),
        "oslmi_execute_batch_sample": ([P, P, P, P, P, P, P, P], I),
        "oslmi_batch_create": ([P, I, P, I, I], P),
        "oslmi_batch_destroy": ([P], None),
        "oslmi_batch_jit_seconds": ([P], ctypes.c_double),
        "oslmi_batch_kernels": ([P], I),
        "oslmi_batch_bind": ([P, I, P], I),
        "oslmi_batch_set_config": ([P, P], I),
        "oslmi_batch_get_config": ([P, P], I),
        "oslmi_batch_launch": ([P, P], I),
        "oslmi_batch_launch_per_group": ([P, P], I),
        "oslmi_batch_capture": ([P, P], I),
        "oslmi_batch_autotune": ([P, P, I, I, P, P], I),
        "oslmi_batch_launch_noalloc": ([P, I, P, P], I),
        "oslmi_execute_batch_closures": ([P, P, P, P, P, I, P, P], I),
        "oslmi_layered_closure": ([P, I64, P, P, P, P, P, P, P, P], I),
        "oslmi_layered_eval_local": ([P, I64, P, P, P], I),
        "oslmi_bsdl_tables": ([P, I64], I),
        "oslmi_surrogate_create": ([I64], P),
        "oslmi_surrogate_destroy": ([P], None),
        "oslmi_surrogate_param_count": ([], I),
        "oslmi_surrogate_set_normalization": ([P, P, P, P, P, P, P], I),
        "oslmi_surrogate_set_params": ([P, P, I64], I),
        "oslmi_surrogate_get_params": ([P, P, I64], I),
        "oslmi_surrogate_grads": ([P, I64, P, P, P, P], I),
        "oslmi_surrogate_train_step": ([P, I64, P, P, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_float, I, P], I),
        "oslmi_surrogate_set_routing": ([P, P, I], I),
        "oslmi_surrogate_set_spot_check": ([P, ctypes.c_float], I),
        "oslmi_surrogate_infer": ([P, P, I64, P, P, P, P, P], I),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res


def lib():
    """Load liboslmi.so (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSLMIError(
                "liboslmi.so is not built (%s); run python -m openshadinglanguage_amd.build"
                % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


class LayeredMaterial(ctypes.Structure):
    """oslmi_layered_material (include/oslmi.h)."""
    _fields_ = [("coat_refl_tint", ctypes.c_float * 3),
                ("coat_refr_tint", ctypes.c_float * 3),
                ("base_refl_tint", ctypes.c_float * 3),
                ("base_refr_tint", ctypes.c_float * 3),
                ("base_exponent", ctypes.c_float),
                ("base_roughness", ctypes.c_float),
                ("sheen_albedo", ctypes.c_float * 3)]

    def as_list(self):
        """The 17 floats in declaration order."""
        out = []
        for name, _ in self._fields_:
            v = getattr(self, name)
            out.extend(list(v) if hasattr(v, "__len__") else [v])
        return out


def check(rc):
    if rc != 0:
        raise OSLMIError("oslmi error %d: %s" % (rc, lib().oslmi_last_error().decode()))
    return rc


def exported_symbols():
    return [n for n in dir(lib()) if n.startswith("oslmi_")]
