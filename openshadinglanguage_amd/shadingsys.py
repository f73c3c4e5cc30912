"""ShadingSystem / ShaderGroup / BatchedExecutor over the C ABI.

Mirrors the reference's host API (src/include/OSL/oslexec.h):
LoadMemoryCompiledShader, ShaderGroupBegin/End, Parameter, Shader,
ConnectShaders, attribute, and the batched executor's jit_group / execute.
Inputs and outputs are torch CUDA tensors of SoA planes.
"""
import ctypes
import os

from ._capi import check, lib

PKG = os.path.dirname(os.path.abspath(__file__))
SHADER_DIR = os.path.join(PKG, "shaders")
CSRC = os.path.join(PKG, "csrc")
JIT_CACHE = os.environ.get("OSLMI_JIT_CACHE", os.path.join(PKG, "_lib", "jitcache"))

# global slot order of oslmi_batched_globals (include/oslmi.h)
GLOBALS = ["P", "dPdx", "dPdy", "N", "Ng", "I", "u", "v", "dudx", "dudy", "dvdx", "dvdy",
           "time", "dPdu", "dPdv", "Ps"]
TRIPLES = {"P", "dPdx", "dPdy", "N", "Ng", "I", "dPdu", "dPdv", "Ps"}


class BatchedGlobals(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64),
                ("planes", ctypes.c_void_p * 16),
                ("uniform", (ctypes.c_float * 3) * 16)]


class BatchItem(ctypes.Structure):
    _fields_ = [("globals", BatchedGlobals),
                ("d_outputs", ctypes.c_void_p),
                ("d_wo", ctypes.c_void_p),
                ("d_wi", ctypes.c_void_p),
                ("d_eval", ctypes.c_void_p)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int), ("ppt", ctypes.c_int), ("blocks_per_cu", ctypes.c_int),
                ("bb_lds", ctypes.c_int), ("order", ctypes.c_int), ("streams", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def batched_globals(n, globals_):
    """oslmi_batched_globals for n points. globals_: dict name -> CUDA
    tensor of planes (varying) or a float / tuple of floats (uniform)."""
    bg = BatchedGlobals()
    bg.n = n
    for k, name in enumerate(GLOBALS):
        v = globals_.get(name)
        if v is None:
            continue
        if hasattr(v, "data_ptr"):
            bg.planes[k] = v.data_ptr()
        else:
            vals = list(v) if hasattr(v, "__len__") else [v]
            for c in range(3):
                bg.uniform[k][c] = float(vals[min(c, len(vals) - 1)])
    return bg


class ShaderGroup:
    def __init__(self, ss, handle, name):
        self.ss, self.h, self.name = ss, handle, name

    def __del__(self):
        try:
            lib().oslmi_group_release(self.h)
        except Exception:
            pass


class ShadingSystem:
    def __init__(self, jit_cache=JIT_CACHE):
        L = lib()
        self.h = L.oslmi_shadingsys_create(CSRC.encode(),
                                           jit_cache.encode() if jit_cache else None)
        if not self.h:
            raise RuntimeError("oslmi_shadingsys_create failed")

    def __del__(self):
        try:
            lib().oslmi_shadingsys_destroy(self.h)
        except Exception:
            pass

    def attribute(self, name, value):
        check(lib().oslmi_shadingsys_attribute(self.h, name.encode(), str(value).encode()))

    def name_transform(self, name, matrix, inverse=None):
        """A named coordinate system of the renderer services: matrix is
        the space-to-common 4x4 in Imath's row-vector layout (translation in
        the last row); inverse, if given, is what get_inverse_matrix returns
        (SimpleRenderer::name_transform, src/testshade/simplerend.cpp:464).
        "shader" and "object" name the batch's shader2common and
        object2common matrices."""
        import numpy as np
        m = np.ascontiguousarray(np.asarray(matrix, np.float32).reshape(16))
        inv = None if inverse is None else np.ascontiguousarray(
            np.asarray(inverse, np.float32).reshape(16))
        check(lib().oslmi_renderer_name_transform(
            self.h, name.encode(), m.ctypes.data_as(ctypes.c_void_p),
            None if inv is None else inv.ctypes.data_as(ctypes.c_void_p)))

    def set_attribute(self, name, typ, value):
        """A renderer attribute for getattribute(): typ "int", "float", a
        triple type or "matrix"."""
        import numpy as np
        dt = np.int32 if typ == "int" else np.float32
        v = np.ascontiguousarray(np.asarray(value, dt).reshape(-1))
        check(lib().oslmi_renderer_set_attribute(self.h, name.encode(), typ.encode(),
                                                 v.ctypes.data_as(ctypes.c_void_p)))

    def LoadMemoryCompiledShader(self, shadername, oso_text):
        check(lib().oslmi_load_memory_compiled_shader(self.h, shadername.encode(),
                                                      oso_text.encode()))

    def load_shader_file(self, shadername, path=None):
        path = path or os.path.join(SHADER_DIR, shadername + ".oso")
        with open(path) as f:
            self.LoadMemoryCompiledShader(shadername, f.read())

    def ShaderGroupBegin(self, name="group"):
        h = lib().oslmi_shader_group_begin(self.h, name.encode())
        return ShaderGroup(self, h, name)

    def Parameter(self, group, name, typename, value):
        if typename == "string":
            vals = [value] if isinstance(value, str) else list(value)
            arr = (ctypes.c_char_p * len(vals))(*[v.encode() for v in vals])
            check(lib().oslmi_parameter(self.h, group.h, name.encode(), b"string",
                                        ctypes.cast(arr, ctypes.c_void_p), len(vals)))
            return
        if typename == "int":
            vals = [value] if isinstance(value, int) else list(value)
            arr = (ctypes.c_int * len(vals))(*vals)
            count = len(vals)
        else:
            vals = [float(value)] if isinstance(value, (int, float)) else [float(v) for v in value]
            arr = (ctypes.c_float * len(vals))(*vals)
            per = {"float": 1, "matrix": 16}.get(typename, 3)
            count = max(1, len(vals) // per)
        check(lib().oslmi_parameter(self.h, group.h, name.encode(), typename.encode(),
                                    ctypes.cast(arr, ctypes.c_void_p), count))

    def Shader(self, group, usage, shadername, layername):
        check(lib().oslmi_shader(self.h, group.h, usage.encode(), shadername.encode(),
                                 layername.encode()))

    def ConnectShaders(self, group, srclayer, srcparam, dstlayer, dstparam):
        check(lib().oslmi_connect_shaders(self.h, group.h, srclayer.encode(), srcparam.encode(),
                                          dstlayer.encode(), dstparam.encode()))

    def ShaderGroupEnd(self, group):
        check(lib().oslmi_shader_group_end(self.h, group.h))

    def set_outputs(self, group, names):
        arr = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
        check(lib().oslmi_group_outputs(self.h, group.h, len(names), arr))

    def emit(self, group):
        """HIP source of the group's fused kernel (no device needed)."""
        n = lib().oslmi_group_emit(self.h, group.h, None, 0)
        check(n if n < 0 else 0)
        buf = ctypes.create_string_buffer(n + 1)
        lib().oslmi_group_emit(self.h, group.h, buf, n + 1)
        return buf.value.decode()

    def compile_group(self, group):
        """Emit + hiprtc compile into the code-object cache (no device)."""
        check(lib().oslmi_compile_group(self.h, group.h))
        return lib().oslmi_group_jit_seconds(group.h)

    def jit_group(self, group):
        check(lib().oslmi_jit_group(self.h, group.h))
        return lib().oslmi_group_jit_seconds(group.h)

    def output_planes(self, group):
        n = lib().oslmi_group_output_planes(self.h, group.h)
        check(n if n < 0 else 0)
        return n

    def execute(self, group, n, globals_, outputs=None, wo=None, wi=None, eval_out=None,
                stream=None):
        """Batched execution over n points. globals_: dict name -> CUDA
        tensor of planes (varying) or a tuple of floats (uniform)."""
        self.bind(group, n, globals_, outputs, wo, wi, eval_out, stream)()

    def bind(self, group, n, globals_, outputs=None, wo=None, wi=None, eval_out=None,
             stream=None):
        """Prepares the argument block once; returns a callable that
        launches the batched execution (repeated launches reuse it)."""
        bg = batched_globals(n, globals_)
        import torch
        sp = ctypes.c_void_p(stream.cuda_stream if stream is not None
                             else torch.cuda.current_stream().cuda_stream)
        fn = lib().oslmi_execute_batch
        args = (self.h, group.h, ctypes.byref(bg), _ptr(outputs), _ptr(wo), _ptr(wi),
                _ptr(eval_out), sp)
        keep = (bg, globals_, outputs, wo, wi, eval_out)

        def launch():
            check(fn(*args))
            return keep
        return launch


    def sample(self, group, n, globals_, wo, rnd, samples, outputs=None, stream=None):
        """Batched execution with the surface closure sampled at wo
        (CompositeBSDF::sample): rnd 3 planes (rx, ry, rz), samples 7
        planes (wi xyz, weight rgb, pdf). Needs the group compiled after
        attribute("closure_sample", 1)."""
        bg = batched_globals(n, globals_)
        import torch
        sp = ctypes.c_void_p(stream.cuda_stream if stream is not None
                             else torch.cuda.current_stream().cuda_stream)
        check(lib().oslmi_execute_batch_sample(self.h, group.h, ctypes.byref(bg),
                                               _ptr(outputs), _ptr(wo), _ptr(rnd),
                                               _ptr(samples), sp))


    def closures(self, group, n, globals_, closures, max_lobes=8, wo=None, outputs=None,
                 stream=None):
        """Batched execution with the surface closure handed out as
        testrender's list of (closure id, weight rgb, parameters)
        (oslmi_execute_batch_closures): closures 1 + max_lobes *
        CLOSURE_RECORD planes. Needs the group compiled after
        attribute("closure_export", 1)."""
        bg = batched_globals(n, globals_)
        import torch
        sp = ctypes.c_void_p(stream.cuda_stream if stream is not None
                             else torch.cuda.current_stream().cuda_stream)
        check(lib().oslmi_execute_batch_closures(self.h, group.h, ctypes.byref(bg),
                                                 _ptr(outputs), _ptr(wo), max_lobes,
                                                 _ptr(closures), sp))


# planes per exported closure component: id, weight rgb, 21 parameters
CLOSURE_RECORD = 25


def build_fbm_cell(ss=None, outputs=("Cout",)):
    """The single-layer fbm_cell group."""
    ss = ss or ShadingSystem()
    ss.load_shader_file("fbm_cell")
    g = ss.ShaderGroupBegin("fbm_cell")
    ss.Shader(g, "surface", "fbm_cell", "fbm")
    ss.ShaderGroupEnd(g)
    ss.set_outputs(g, list(outputs))
    return ss, g


def build_layered3(ss=None, outputs=()):
    """The 3-layer layered3 group: pattern -> color -> surface."""
    ss = ss or ShadingSystem()
    for name, fn in (("c2_pattern", "layered3_pattern"), ("c2_color", "layered3_color"),
                     ("c2_surface", "layered3_surface")):
        ss.load_shader_file(name, os.path.join(SHADER_DIR, fn + ".oso"))
    g = ss.ShaderGroupBegin("layered3")
    ss.Shader(g, "shader", "c2_pattern", "pattern")
    ss.Shader(g, "shader", "c2_color", "color")
    ss.Shader(g, "surface", "c2_surface", "surface")
    for p in ("fbm", "vn", "gab"):
        ss.ConnectShaders(g, "pattern", p, "color", p)
    ss.ConnectShaders(g, "color", "Cd", "surface", "Cd")
    ss.ShaderGroupEnd(g)
    ss.set_outputs(g, list(outputs))
    return ss, g


def build_mx_layered(ss=None, params=None):
    """The closure workload's layered MaterialX material as a one-layer
    group (shaders/mx_layered.osl): per-point parameters in the globals P
    (coat ior, coat roughness, sheen roughness), dPdu (F0), dPdv (F90);
    params: instance values of coat_tint, base_tint, sheen_albedo,
    base_exponent."""
    ss = ss or ShadingSystem()
    ss.load_shader_file("mx_layered", os.path.join(SHADER_DIR, "mx_layered.oso"))
    g = ss.ShaderGroupBegin("mx_layered")
    for name, (typ, value) in (params or {}).items():
        ss.Parameter(g, name, typ, value)
    ss.Shader(g, "surface", "mx_layered", "surface")
    ss.ShaderGroupEnd(g)
    return ss, g


def build_generated_group(ss, spec, name, outputs=("fout", "cout")):
    """A group from groupgen's (layers, connections, ...) description:
    layers [(shadername, layername, oso_text)], the last one the surface."""
    layers, conns = spec[0], spec[1]
    for sn, _, txt in layers:
        ss.LoadMemoryCompiledShader(sn, txt)
    g = ss.ShaderGroupBegin(name)
    for k, (sn, ln, _) in enumerate(layers):
        ss.Shader(g, "surface" if k == len(layers) - 1 else "shader", sn, ln)
    for c in conns:
        ss.ConnectShaders(g, *c)
    ss.ShaderGroupEnd(g)
    ss.set_outputs(g, list(outputs))
    return g


class GroupBatch:
    """Many groups shaded by one persistent launch (oslmi_batch_*).

    bind() takes one item per group: (n, globals dict, outputs, wo, wi,
    eval_out) with the meaning ShadingSystem.execute gives them."""

    def __init__(self, ss, groups, groups_per_kernel=0, load=True):
        L = lib()
        self.ss, self.groups = ss, list(groups)
        arr = (ctypes.c_void_p * len(self.groups))(*[g.h for g in self.groups])
        self.h = L.oslmi_batch_create(ss.h, len(self.groups), arr, groups_per_kernel,
                                      1 if load else 0)
        if not self.h:
            raise RuntimeError("oslmi_batch_create: %s" % L.oslmi_last_error().decode())
        self.jit_seconds = L.oslmi_batch_jit_seconds(self.h)
        self.kernels = L.oslmi_batch_kernels(self.h)
        self._keep = None

    def __del__(self):
        try:
            lib().oslmi_batch_destroy(self.h)
        except Exception:
            pass

    def bind(self, items):
        arr = (BatchItem * len(items))()
        for k, (n, glob, outputs, wo, wi, ev) in enumerate(items):
            arr[k].globals = batched_globals(n, glob)
            arr[k].d_outputs = outputs.data_ptr() if outputs is not None else None
            arr[k].d_wo = wo.data_ptr() if wo is not None else None
            arr[k].d_wi = wi.data_ptr() if wi is not None else None
            arr[k].d_eval = ev.data_ptr() if ev is not None else None
        check(lib().oslmi_batch_bind(self.h, len(items), arr))
        self._keep = (arr, items)

    @staticmethod
    def _stream(stream):
        import torch
        s = stream or torch.cuda.current_stream()
        return ctypes.c_void_p(s.cuda_stream)

    def set_config(self, **kw):
        c = self.config()
        for k, v in kw.items():
            setattr(c, k, int(v))
        check(lib().oslmi_batch_set_config(self.h, ctypes.byref(c)))

    def config(self):
        c = BatchConfig()
        check(lib().oslmi_batch_get_config(self.h, ctypes.byref(c)))
        return c

    def launch(self, stream=None):
        check(lib().oslmi_batch_launch(self.h, self._stream(stream)))

    def launch_per_group(self, stream=None):
        check(lib().oslmi_batch_launch_per_group(self.h, self._stream(stream)))

    def capture(self, stream):
        check(lib().oslmi_batch_capture(self.h, self._stream(stream)))

    def autotune(self, stream=None, reps=3, graph=False):
        """Tunes the launch shape; with graph=True every candidate is timed
        as a replayed hipGraph on `stream` and the winner stays captured."""
        c = BatchConfig()
        ms = ctypes.c_double()
        check(lib().oslmi_batch_autotune(self.h, self._stream(stream), reps, 1 if graph else 0,
                                         ctypes.byref(c), ctypes.byref(ms)))
        return c.as_dict(), ms.value
