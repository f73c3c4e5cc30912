"""OSO reader, group building and the HIP emitter (no GPU needed: emission
and the hiprtc compile run on the host)."""
import os

import pytest

from openshadinglanguage_amd import OSLMIError
from openshadinglanguage_amd import shadingsys as S

pytest.importorskip("ctypes")


def test_emit_fbm_cell():
    ss, g = S.build_fbm_cell()
    src = ss.emit(g)
    assert "oslmi_group_main" in src
    assert "perlin3<float>" in src and "intnoise3<true>" in src
    # no derivatives anywhere in this group
    assert "Df " not in src and "Dv " not in src


def test_emit_layered3_derivative_analysis():
    ss, g = S.build_layered3()
    src = ss.emit(g)
    # P of the pattern layer carries derivatives (gabor + filterwidth need them)
    assert "Dv(ld3(A, 0, i), ld3(A, 1, i), ld3(A, 2, i))" in src
    assert "simplex3_d(" in src and "vperlin3<Df>" in src and "gabor(gu_" in src
    # the surface layer's noise needs no derivatives
    assert "perlin3<float>" in src
    assert "blackbody_rgb(s_bb" in src and "composite_eval" in src


def test_occupancy_attribute():
    ss, g = S.build_fbm_cell()
    assert "amdgpu_waves_per_eu(6)" in ss.emit(g)  # the default target
    ss.attribute("gpu_waves_per_eu", 0)
    assert "amdgpu_waves_per_eu" not in ss.emit(g)
    ss.attribute("gpu_waves_per_eu", 4)
    assert "amdgpu_waves_per_eu(4)" in ss.emit(g)
    for bad in ("9", "-1", "six", ""):
        with pytest.raises(OSLMIError):
            ss.attribute("gpu_waves_per_eu", bad)


def test_hiprtc_compile_both_groups(tmp_path):
    for build in (S.build_fbm_cell, S.build_layered3):
        ss, g = build(S.ShadingSystem(jit_cache=str(tmp_path)))
        ss.compile_group(g)
    assert len([f for f in os.listdir(tmp_path) if f.endswith(".hsaco")]) == 2


def test_bad_oso_is_rejected():
    ss = S.ShadingSystem(jit_cache=None)
    with pytest.raises(OSLMIError):
        ss.LoadMemoryCompiledShader("bad", "OpenShadingLanguage 1.00\nshader bad\ncode ___main___\n\tadd x y z\n")
    with pytest.raises(OSLMIError):
        ss.LoadMemoryCompiledShader("bad2", "not an oso file\n")


def test_unknown_param_and_layer_errors():
    ss, _ = S.build_fbm_cell()
    g = ss.ShaderGroupBegin("g2")
    ss.Parameter(g, "nosuchparam", "float", 1.0)
    with pytest.raises(OSLMIError):
        ss.Shader(g, "surface", "fbm_cell", "a")
    g3 = ss.ShaderGroupBegin("g3")
    ss.Shader(g3, "surface", "fbm_cell", "a")
    with pytest.raises(OSLMIError):
        ss.ConnectShaders(g3, "a", "Cout", "missing", "c0")


def test_parameter_override_reaches_code():
    ss = S.ShadingSystem(jit_cache=None)
    ss.load_shader_file("fbm_cell")
    g = ss.ShaderGroupBegin("g")
    ss.Parameter(g, "octaves", "int", 7)
    ss.Parameter(g, "c0", "color", (0.5, 0.25, 1.0))
    ss.Shader(g, "surface", "fbm_cell", "fbm")
    ss.ShaderGroupEnd(g)
    ss.set_outputs(g, ["Cout"])
    src = ss.emit(g)
    assert "int v0_0 = 7;" in src
    assert "V3(0x1p-1f, 0x1p-2f, 0x1p+0f)" in src


def test_unsupported_op_reports_error():
    oso = ("OpenShadingLanguage 1.00\nshader s\nglobal point P\nlocal float x\n"
           "code ___main___\n\tspline x P\n\tend\n")
    ss = S.ShadingSystem(jit_cache=None)
    ss.LoadMemoryCompiledShader("s", oso)
    g = ss.ShaderGroupBegin("g")
    ss.Shader(g, "surface", "s", "a")
    ss.ShaderGroupEnd(g)
    with pytest.raises(OSLMIError, match="unsupported op"):
        ss.emit(g)


def test_batch_kernel_emits_and_compiles(tmp_path):
    """A batch over two benchmark groups and two generated ones compiles
    for gfx950 without a device; bad configs are rejected on the host."""
    from openshadinglanguage_amd import groupgen as G
    ss = S.ShadingSystem(jit_cache=str(tmp_path))
    _, g1 = S.build_fbm_cell(ss)
    _, g2 = S.build_layered3(ss)
    specs, _ = G.make_groups(4, 3)
    g3 = S.build_generated_group(ss, specs[1], "gen1")
    g4 = S.build_generated_group(ss, specs[3], "gen3")
    b = S.GroupBatch(ss, [g1, g2, g3, g4], load=False)
    assert b.jit_seconds >= 0
    # each group's own kernel (its register need decides the packing) and
    # one batch kernel holding all four
    assert len([f for f in os.listdir(tmp_path) if f.endswith(".hsaco")]) == 5
    assert b.kernels == 1
    assert b.config().as_dict() == dict(block=256, ppt=1, blocks_per_cu=8, bb_lds=1, order=0,
                                        streams=1)
    for bad in (dict(block=96), dict(ppt=0), dict(ppt=9), dict(blocks_per_cu=0), dict(order=2),
                dict(bb_lds=3), dict(streams=0), dict(streams=5)):
        with pytest.raises(OSLMIError):
            b.set_config(**bad)
    b.set_config(block=64, ppt=4)
    assert b.config().block == 64 and b.config().ppt == 4


def test_batch_rejects_open_group():
    ss = S.ShadingSystem(jit_cache=None)
    ss.load_shader_file("fbm_cell")
    g = ss.ShaderGroupBegin("open")
    ss.Shader(g, "surface", "fbm_cell", "a")
    with pytest.raises(RuntimeError):
        S.GroupBatch(ss, [g], load=False)


def _microfacet_shader(xa, ya, u):
    from openshadinglanguage_amd import groupgen as G
    o = G.Oso("surface", "mf_%s_%s_%s" % (xa, ya, u))
    o.sym("global", "normal", "N")
    o.sym("global", "closure color", "Ci")
    o.sym("oparam", "float", "fout", "0")
    o.sym("oparam", "color", "cout", "0 0 0")
    cl = o.tmp("closure color")
    o.op("closure", cl, o.const("string", "microfacet"), o.const("string", "ggx"), "N",
         o.const("vector", *u), o.const("float", xa), o.const("float", ya),
         o.const("float", 1.5), o.const("int", 0))
    o.op("assign", "Ci", cl)
    return o


def test_anisotropic_microfacet_builds_the_tangent_frame_lobe():
    """Equal roughnesses about N keep the short isotropic lobe; two
    roughnesses or a tangent build the anisotropic lobe (the kernel then
    defines OSLMI_FULL_MICROFACET), which isotropic kernels never carry."""
    ss = S.ShadingSystem()
    ok = _microfacet_shader(0.3, 0.3, (0, 0, 0))
    g = S.build_generated_group(ss, ([(ok.name, "l0", ok.text())], []), "mf_ok")
    src = ss.emit(g)
    assert "closure_microfacet_ggx" in src and "OSLMI_FULL_MICROFACET" not in src
    for k, (xa, ya, u) in enumerate(((0.3, 0.2, (0, 0, 0)), (0.3, 0.3, (1, 0, 0)))):
        o = _microfacet_shader(xa, ya, u)
        g = S.build_generated_group(ss, ([(o.name, "l0", o.text())], []), "mf_aniso%d" % k)
        src = ss.emit(g)
        assert "closure_microfacet_full(kClosureMicrofacetGGX" in src
        assert "#define OSLMI_FULL_MICROFACET 1" in src


def _beckmann_shader():
    from openshadinglanguage_amd import groupgen as G
    o = G.Oso("surface", "beckmann_mix")
    o.sym("global", "normal", "N")
    o.sym("global", "closure color", "Ci")
    o.sym("oparam", "float", "fout", "0")
    o.sym("oparam", "color", "cout", "0 0 0")
    d, m, dw, mw = (o.tmp("closure color") for _ in range(4))
    o.op("closure", d, o.const("string", "diffuse"), "N")
    a = o.const("float", 0.25)
    o.op("closure", m, o.const("string", "microfacet"), o.const("string", "beckmann"), "N",
         o.const("vector", 0, 0, 0), a, a, o.const("float", 1.5), o.const("int", 0))
    o.op("mul", dw, d, o.const("color", 0.4, 0.3, 0.2))
    o.op("mul", mw, m, o.const("color", 0.6, 0.6, 0.6))
    o.op("add", "Ci", dw, mw)
    return o


def test_beckmann_emits_for_eval_and_sampling():
    o = _beckmann_shader()
    ss = S.ShadingSystem()
    g = S.build_generated_group(ss, ([(o.name, "l0", o.text())], []), "bk")
    assert "closure_microfacet_beckmann" in ss.emit(g)
    ss2 = S.ShadingSystem()
    ss2.attribute("closure_sample", 1)
    g2 = S.build_generated_group(ss2, ([(o.name, "l0", o.text())], []), "bk_s")
    assert "closure_microfacet_beckmann" in ss2.emit(g2)


def test_closure_capacity_follows_dataflow():
    """The lobe capacity the emitter gives closure values counts lobes
    through add/mul/assign and connections: a closure reused in an add
    needs two slots, one grown in a loop the composite's maximum of 8, and
    Ci written by an upstream layer is the group's surface closure."""
    import _cfprogs as C
    for prog, cap in ((C.closure_reuse, 2), (C.closure_loop, 8)):
        o, _ = prog()
        ss = S.ShadingSystem()
        g = S.build_generated_group(ss, ([(o.name, "l0", o.text())], []), o.name)
        assert "using Closure = LobeList<%d>;" % cap in ss.emit(g)
    osos, conns, _ = C.closure_upstream_layers()
    ss = S.ShadingSystem()
    layers = [(o.name, "l%d" % k, o.text()) for k, o in enumerate(osos)]
    src = ss.emit(S.build_generated_group(ss, (layers, conns), "up"))
    assert "using Closure = LobeList<2>;" in src
    assert src.count("ci_ = ") == 2  # both layers hand Ci on


def _mx_shader(name, kind, extra=()):
    """A surface shader with one MaterialX closure; extra: keyword (name,
    Oso-constant args) pairs appended to the closure call."""
    import _mxprogs as X
    w = X._W(name)
    x, y, z = w.f
    if kind == "dielectric":
        args = ["N", w.zero_v(), w.cconst(1, 1, 1), w.zero_c(), x, x, w.fconst(1.5),
                w.o.const("string", "ggx")]
    elif kind == "dielectric_refr":
        args = ["N", w.zero_v(), w.cconst(1, 1, 1), w.cconst(0.5, 0.5, 0.5), x, x, w.fconst(1.5),
                w.o.const("string", "ggx")]
        kind = "dielectric"
    else:
        args = ["N", w.cconst(1, 1, 1), z]
    cname = "dielectric_bsdf" if kind == "dielectric" else "sheen_bsdf"
    for key, val in extra:
        args += [w.o.const("string", key), val(w)]
    return w.done(w.closure(cname, *args))


def test_materialx_closures_emit_and_reject_what_they_do_not_model():
    """dielectric_bsdf / sheen_bsdf / layer compile to libbsdl lobes (the
    kernel defines OSLMI_MX_CLOSURES); a transmission tint, an absorption
    or a sheen mode other than a constant zero, and unknown keywords, are
    rejected at emit time; labels and thin-film keywords pass."""
    import _mxprogs as X
    ss = S.ShadingSystem()
    for name, o in X.programs():
        g = S.build_generated_group(ss, ([(o.name, "l0", o.text())], []), name, outputs=())
        src = ss.emit(g)
        assert "#define OSLMI_MX_CLOSURES 1" in src and "mxwo_" in src, name
    ok = [_mx_shader("ok_d", "dielectric", (("label", lambda w: w.o.const("string", "coat")),
                                            ("thinfilm_thickness", lambda w: w.fconst(0.0)))),
          _mx_shader("ok_s", "sheen", (("mode", lambda w: w.o.const("int", 0)),))]
    for o in ok:
        g = S.build_generated_group(ss, ([(o.name, "l0", o.text())], []), o.name, outputs=())
        ss.emit(g)
    bad = [(_mx_shader("b_refr", "dielectric_refr"), "transmission"),
           (_mx_shader("b_abs", "dielectric", (("absorption", lambda w: w.cconst(0.2, 0, 0)),)),
            "absorption"),
           (_mx_shader("b_mode", "sheen", (("mode", lambda w: w.o.const("int", 1)),)), "mode"),
           (_mx_shader("b_kw", "sheen", (("shininess", lambda w: w.fconst(1.0)),)), "shininess")]
    for o, word in bad:
        g = S.build_generated_group(ss, ([(o.name, "l0", o.text())], []), o.name, outputs=())
        with pytest.raises(OSLMIError, match=word):
            ss.emit(g)
    # classic groups are compiled without the MaterialX lobe state
    ss2, g2 = S.build_layered3()
    assert "OSLMI_MX_CLOSURES" not in ss2.emit(g2)


def test_periodic_noise_emits_lattice_wraps_and_rejects_simplex():
    """pnoise/psnoise lower to the wrapped lattice (make_wrap for perlin,
    wrap_period for cell/hash and gabor's cell seeds), the vector gabor to
    gabor3; periodic simplex and a period count that does not match the
    coordinates are rejected, as the reference rejects them."""
    import _pnprogs as PN
    ss = S.ShadingSystem(jit_cache=None)
    want = {"pn_perlin": "make_wrap(", "pn_uperlin": "make_wrap(", "pn_default": "make_wrap(",
            "ps_default": "make_wrap(", "pn_cell": "wrap_period(", "pn_hash": "wrap_period(",
            "pn_gabor": "gabor3(gu_", "n_gabor3": "gabor3(gu_"}
    for name, text in PN.programs():
        g = S.build_generated_group(ss, ([(name, "surface", text)], []), name, outputs=PN.OUTPUTS)
        src = ss.emit(g)
        assert want[name] in src, name
        if name == "pn_gabor":
            assert ", true, V3(" in src
    hdr = ("OpenShadingLanguage 1.00\nshader s\nglobal point P\noparam float f 0\n"
           "const string $c1 \"simplex\"\nconst float $c2 3\n")
    for body, word in (("\tpnoise f $c1 P P\n", "periodic noise type"),
                       ("\tpnoise f P\n", "Pperiod"),
                       ("\tpnoise f P $c2\n", "periods do not match")):
        ss.LoadMemoryCompiledShader("s", hdr + "code ___main___\n" + body + "\tend\n")
        g = ss.ShaderGroupBegin("g")
        ss.Shader(g, "surface", "s", "a")
        ss.ShaderGroupEnd(g)
        with pytest.raises(OSLMIError, match=word):
            ss.emit(g)
