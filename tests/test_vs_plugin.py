"""The VapourSynth API-3 plugin entry (vsplugin/bm3d_vsplugin.cpp) under a
mock host (tests/vs_host/mock_vs_host.cpp) that loads it with dlopen, calls
VapourSynthPluginInit, invokes the registered functions with argument maps
and pulls frames through the filter graph from 4 threads.

CPU: the identifiers and argument strings of source/source.cpp:991-1027
and rtc_source/source.cpp:1358-1397, the creation errors in the reference's
order and wording (source/source.cpp:457-611, 880-909), and BM3Dv2's
passthrough when no plane is processed (source/source.cpp:932-943).
GPU: frames produced through the plugin equal the CPU oracle bit for bit
(extractor_exp set), including a final pass fed by a basic-pass node,
subsampled planes with one plane skipped, the NVRTC variant's string
options, and BM3D + VAggregate as separate filters with the BM3D_V_radius /
BM3D_V_process frame properties (source/source.cpp:428-435).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from bm3d_amd import synth
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vapoursynth-bm3dcuda_amd")
LIB = os.path.join(PKG, "lib")
PLUGIN = os.path.join(LIB, "libbm3dcuda_mi355x.so")
PLUGIN_RTC = os.path.join(LIB, "libbm3dcuda_rtc_mi355x.so")
HOST_SRC = os.path.join(ROOT, "tests", "vs_host", "mock_vs_host.cpp")

# The reference's registrations (source/source.cpp:1001-1024,
# rtc_source/source.cpp:1368-1394); the plugin appends one optional key.
REF_ARGS = ("clip:clip;ref:clip:opt;sigma:float[]:opt;block_step:int[]:opt;bm_range:int[]:opt;"
            "radius:int:opt;ps_num:int[]:opt;ps_range:int[]:opt;chroma:int:opt;device_id:int:opt;"
            "fast:int:opt;extractor_exp:int:opt;zero_init:int:opt;")
REF_RTC_ARGS = ("clip:clip;ref:clip:opt;sigma:float[]:opt;block_step:int[]:opt;bm_range:int[]:opt;"
                "radius:int:opt;ps_num:int[]:opt;ps_range:int[]:opt;chroma:int:opt;device_id:int:opt;"
                "fast:int:opt;extractor_exp:int:opt;bm_error_s:data[]:opt;transform_2d_s:data[]:opt;"
                "transform_1d_s:data[]:opt;zero_init:int:opt;")
VAGG_ARGS = "clip:clip;src:clip;planes:int[];"
API_VERSION = (3 << 16) | 6


class Build:
    """A mock host with the plugins it loads (and their library)."""

    def __init__(self, host, lib_dir, env=None):
        self.host = host
        self.plugin = os.path.join(lib_dir, "libbm3dcuda_mi355x.so")
        self.plugin_rtc = os.path.join(lib_dir, "libbm3dcuda_rtc_mi355x.so")
        self.lib = os.path.join(lib_dir, "libbm3d_mi355x.so")
        self.env = env or {}


def _plain(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    if not os.path.exists(PLUGIN):
        pytest.fail("plugin not built: run __graft_entry__.build()")
    exe = str(tmp_path_factory.mktemp("vs_host") / "mock_vs_host")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", f"-I{os.path.join(PKG, 'vsplugin')}", HOST_SRC, "-o", exe,
                    "-ldl", "-pthread"], check=True)
    return Build(exe, LIB)


def _sanitized(tmp_path_factory):
    """AddressSanitizer + UBSan build of the library's host code, the plugins
    and the mock host (host code only; kernels as built)."""
    import build_lib
    if not os.path.exists(build_lib.CLANGXX):
        pytest.skip("ROCm clang++ not available")
    out = str(tmp_path_factory.mktemp("asan"))
    exe = build_lib.build_sanitized(out)
    return Build(exe, out, env={"ASAN_OPTIONS": "detect_leaks=1:exitcode=23",
                                "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"})


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    return _plain(tmp_path_factory)


@pytest.fixture(scope="module", params=["plain", "asan+ubsan"])
def cpu_build(request, tmp_path_factory):
    """The CPU tests run on the plain build and on the sanitizer build."""
    if request.param == "plain":
        return _plain(tmp_path_factory)
    if not os.environ.get("BM3D_SKIP_SANITIZED"):
        return _sanitized(tmp_path_factory)
    pytest.skip("sanitizer build skipped (BM3D_SKIP_SANITIZED)")


def run(host, script, env=None, timeout=240):
    if isinstance(host, Build):
        env = dict(host.env, **(env or {}))
        host = host.host
    r = subprocess.run([host], input=script, capture_output=True, text=True, timeout=timeout,
                       env=dict(os.environ, **(env or {})))
    assert r.returncode == 0, f"host exited {r.returncode}\n{r.stdout}\n{r.stderr}"
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    assert r.stdout.rstrip().endswith("done"), r.stdout
    return r.stdout.splitlines()


def write_raw(path, clip):
    np.ascontiguousarray(clip, np.float32).tofile(path)
    return str(path)


def read_raw(path, N, shapes):
    flat = np.fromfile(path, np.float32)
    per = sum(h * w for h, w in shapes)
    assert flat.size == N * per
    frames = flat.reshape(N, per)
    out, off = [], 0
    for h, w in shapes:
        out.append(frames[:, off:off + h * w].reshape(N, h, w))
        off += h * w
    return out


def errors_of(lines):
    return [l.split(" ", 2)[2] for l in lines if l.startswith("error ")]


# ---------------------------------------------------------------- CPU
def test_registration(cpu_build):
    b = cpu_build
    own = os.path.join(os.path.dirname(b.plugin), "libbm3dmi355x.so")
    for lib, ident, ns, args in ((b.plugin, "com.wolframrhodium.bm3dcuda", "bm3dcuda", REF_ARGS),
                                 (b.plugin_rtc, "com.wolframrhodium.bm3dcuda_rtc", "bm3dcuda_rtc", REF_RTC_ARGS),
                                 (own, "com.mi355x.bm3d", "bm3dmi355x", REF_RTC_ARGS)):
        lines = run(b, f"plugin {lib}\n")
        cfg = lines[0].split(" ", 5)
        assert cfg[:5] == ["config", ident, ns, str(API_VERSION), "1"]
        reg = {l.split(" ")[1]: l.split(" ")[2] for l in lines if l.startswith("register ")}
        assert list(reg) == ["BM3D", "VAggregate", "BM3Dv2"]
        assert reg["BM3D"] == reg["BM3Dv2"] == args + "precision:data:opt;"
        assert reg["VAggregate"] == VAGG_ARGS


def test_creation_errors(cpu_build, tmp_path):
    """The reference's checks, in its order, before any device is touched."""
    g = write_raw(tmp_path / "g.raw", np.zeros((3, 1, 40, 48), np.float32))
    g2 = write_raw(tmp_path / "g2.raw", np.zeros((2, 1, 40, 48), np.float32))
    yuv = write_raw(tmp_path / "y.raw", np.zeros((3, 3, 40, 48), np.float32))
    y420 = write_raw(tmp_path / "y420.raw", np.zeros((3 * (40 * 48 + 2 * 20 * 24)), np.float32))
    (tmp_path / "g8.raw").write_bytes(bytes(3 * 40 * 48))
    cases = [
        ("clip=n:g8", "BM3D: only constant format 32bit float input supported"),
        ("clip=n:gh", "BM3D: only constant format 32bit float input supported"),
        ("clip=n:g ref=n:y", 'BM3D: "ref" must be of the same format as "clip"'),
        ("clip=n:g ref=n:gw", 'BM3D: "ref" must be of the same dimensions as "clip"'),
        ("clip=n:g ref=n:g2", 'BM3D: "ref" must be of the same number of frames as "clip"'),
        ("clip=n:g sigma=f:3,-1", 'BM3D: "sigma" must be non-negative'),
        ("clip=n:g block_step=i:0", 'BM3D: "block_step" must be in range [1, 8]'),
        ("clip=n:g block_step=i:4,9", 'BM3D: "block_step" must be in range [1, 8]'),
        ("clip=n:g bm_range=i:0", 'BM3D: "bm_range" must be positive'),
        ("clip=n:g radius=i:-1", 'BM3D: "radius" must be non-negative'),
        ("clip=n:g ps_num=i:9", 'BM3D: "ps_num" must be in range [1, 8]'),
        ("clip=n:g ps_range=i:0", 'BM3D: "ps_range" must be positive'),
        ("clip=n:y420 chroma=i:1", 'BM3D: clip format must be YUV444 when "chroma" is true'),
        ("clip=n:g sigma=f:-1 block_step=i:0", 'BM3D: "sigma" must be non-negative'),
        ("clip=n:g precision=s:fp16", 'BM3D: "precision" must be "fp32" or "mixed"'),
        ("clip=n:g device_id=i:-1", "BM3D: invalid device ID (-1)"),
    ]
    script = [f"plugin {cpu_build.plugin}",
              f"source g GRAYS 48 40 3 {g}", f"source g2 GRAYS 48 40 2 {g2}",
              f"source gw GRAYS 40 40 3 {write_raw(tmp_path / 'gw.raw', np.zeros((3, 1, 40, 40)))}",
              f"source y YUV444PS 48 40 3 {yuv}", f"source y420 YUV420PS 48 40 3 {y420}",
              f"source g8 GRAY8 48 40 3 {tmp_path / 'g8.raw'}",
              f"source gh GRAYH 48 40 3 {write_raw(tmp_path / 'gh.raw', np.zeros(3 * 40 * 48 // 2))}"]
    # BM3Dv2 checks the same, except that with no plane left to process
    # (sigma -1 on a GRAY clip) it returns the clip before BM3D's checks run
    v2 = [(a, m) for a, m in cases if a != "clip=n:g sigma=f:-1 block_step=i:0"]
    script += [f"invoke out BM3D {args}" for args, _ in cases]
    script += [f"invoke out BM3Dv2 {args}" for args, _ in v2]
    script += ["invoke v VAggregate clip=n:g src=n:g planes=i:0",
               "invoke v VAggregate clip=n:g src=n:g planes=i:3"]
    got = errors_of(run(cpu_build, "\n".join(script) + "\n"))
    want = [m for _, m in cases] + [m for _, m in v2]
    want += ['VAggregate: "clip" is not a BM3D output with radius > 0 for "src"'] * 2
    assert got == want


def test_bm3dv2_passthrough(cpu_build, tmp_path):
    """BM3Dv2 with every sigma below FLT_EPSILON returns the clip itself, and
    its frames are the source's."""
    g = write_raw(tmp_path / "g.raw", np.arange(2 * 16 * 16, dtype=np.float32).reshape(2, 1, 16, 16))
    lines = run(cpu_build, f"plugin {cpu_build.plugin}\nsource s GRAYS 16 16 2 {g}\n"
                           f"invoke o BM3Dv2 clip=n:s sigma=f:0\nfetch o {tmp_path / 'o.raw'} 2\n")
    assert "ok o 16 16 2 source" in lines
    (o,) = read_raw(tmp_path / "o.raw", 2, [(16, 16)])
    np.testing.assert_array_equal(o, np.fromfile(g, np.float32).reshape(2, 16, 16))


# ---------------------------------------------------------------- GPU
def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.gpu
def test_two_pass_through_plugin_equals_oracle(host, tmp_path):
    _gpu()
    H, W, N = 48, 64, 6
    clean, noisy = synth.make_clip(H, W, N, 1, noise_sd=12.0, seed=300, motion=True)
    src = write_raw(tmp_path / "in.raw", noisy)
    kw = "sigma=f:12 block_step=i:4 bm_range=i:5 radius=i:1 ps_num=i:2 ps_range=i:3 extractor_exp=i:10"
    lines = run(host, "\n".join([
        f"plugin {PLUGIN}", f"source s GRAYS {W} {H} {N} {src}",
        f"invoke basic BM3Dv2 clip=n:s {kw}",
        f"invoke final BM3Dv2 clip=n:s ref=n:basic {kw}",
        f"fetch basic {tmp_path / 'basic.raw'} 4",
        f"fetch final {tmp_path / 'final.raw'} 4",
        f"invoke final1 BM3Dv2 clip=n:s ref=n:basic fast=i:0 {kw}",
        f"fetch final1 {tmp_path / 'final1.raw'} 4"]) + "\n")
    assert not errors_of(lines), lines
    assert [l for l in lines if l.startswith("ok final")] == [f"ok final {W} {H} {N} BM3Dv2",
                                                             f"ok final1 {W} {H} {N} BM3Dv2"]
    p = O.Params(sigma=(12.0,), block_step=4, bm_range=5, radius=1, ps_num=2, ps_range=3, extractor_exp=10)
    ob, of = O.two_pass(noisy, p)
    (gb,) = read_raw(tmp_path / "basic.raw", N, [(H, W)])
    (gf,) = read_raw(tmp_path / "final.raw", N, [(H, W)])
    (gf1,) = read_raw(tmp_path / "final1.raw", N, [(H, W)])
    np.testing.assert_array_equal(gb, ob[:, 0])
    np.testing.assert_array_equal(gf, of[:, 0])
    np.testing.assert_array_equal(gf1, of[:, 0])


@pytest.mark.gpu
def test_subsampled_planes_and_skipped_plane(host, tmp_path):
    """YUV420 with sigma [10, 0, 8]: planes 0 and 2 filtered independently,
    plane 1 the source's (source/source.cpp:284-294, 838-846)."""
    _gpu()
    H, W, N = 48, 64, 4
    _, luma = synth.make_clip(H, W, N, 1, noise_sd=10.0, seed=301, motion=True)
    _, chroma = synth.make_clip(H // 2, W // 2, N, 2, noise_sd=8.0, seed=302, motion=True)
    raw = np.concatenate([luma.reshape(N, -1), chroma.reshape(N, -1)], axis=1)
    src = write_raw(tmp_path / "in.raw", raw)
    kw = "sigma=f:10,0,8 block_step=i:4 bm_range=i:4 radius=i:1 ps_num=i:2 ps_range=i:2 extractor_exp=i:10"
    lines = run(host, "\n".join([
        f"plugin {PLUGIN}", f"source s YUV420PS {W} {H} {N} {src}",
        f"invoke o BM3Dv2 clip=n:s {kw}", f"fetch o {tmp_path / 'o.raw'} 4"]) + "\n")
    assert not errors_of(lines), lines
    y, u, v = read_raw(tmp_path / "o.raw", N, [(H, W), (H // 2, W // 2), (H // 2, W // 2)])
    prm = dict(block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=2, extractor_exp=10)
    np.testing.assert_array_equal(y, O.bm3dv2(luma, None, O.Params(sigma=(10.0,), **prm))[:, 0])
    np.testing.assert_array_equal(u, chroma[:, 0])
    np.testing.assert_array_equal(v, O.bm3dv2(chroma[:, 1:2], None, O.Params(sigma=(8.0,), **prm))[:, 0])


@pytest.mark.gpu
def test_bm3d_then_vaggregate_and_props(host, tmp_path):
    """BM3D radius 1 stacks carry BM3D_V_radius / BM3D_V_process; VAggregate
    over them equals BM3Dv2; radius 0 BM3D equals the oracle's frame."""
    _gpu()
    H, W, N = 40, 56, 5
    _, noisy = synth.make_clip(H, W, N, 1, noise_sd=15.0, seed=303, motion=True)
    src = write_raw(tmp_path / "in.raw", noisy)
    kw = "sigma=f:15 block_step=i:4 bm_range=i:4 ps_num=i:2 ps_range=i:2 extractor_exp=i:10"
    lines = run(host, "\n".join([
        f"plugin {PLUGIN}", f"source s GRAYS {W} {H} {N} {src}",
        f"invoke st BM3D clip=n:s radius=i:1 {kw}",
        "props st 2",
        "invoke va VAggregate clip=n:st src=n:s planes=i:0",
        f"invoke v2 BM3Dv2 clip=n:s radius=i:1 {kw}",
        f"invoke sp BM3D clip=n:s {kw}",
        f"fetch va {tmp_path / 'va.raw'} 4", f"fetch v2 {tmp_path / 'v2.raw'} 4",
        f"fetch sp {tmp_path / 'sp.raw'} 4"]) + "\n")
    assert not errors_of(lines), lines
    assert f"ok st {W} {H * 2 * 3} {N} BM3D" in lines
    props = {l.split()[1]: l.split()[2:] for l in lines if l.startswith("prop ")}
    assert props["BM3D_V_radius"] == ["i", "1"]
    assert props["BM3D_V_process"] == ["i", "1", "1", "1"]
    assert props["_SourceFrame"] == ["i", "2"]  # the source frame's properties are kept
    (va,) = read_raw(tmp_path / "va.raw", N, [(H, W)])
    (v2,) = read_raw(tmp_path / "v2.raw", N, [(H, W)])
    np.testing.assert_array_equal(va, v2)
    prm = dict(block_step=4, bm_range=4, ps_num=2, ps_range=2, extractor_exp=10)
    np.testing.assert_array_equal(v2, O.bm3dv2(noisy, None, O.Params(sigma=(15.0,), radius=1, **prm))[:, 0])
    (sp,) = read_raw(tmp_path / "sp.raw", N, [(H, W)])
    np.testing.assert_array_equal(sp, O.bm3dv2(noisy, None, O.Params(sigma=(15.0,), radius=0, **prm))[:, 0])


@pytest.mark.gpu
def test_rtc_options_and_mixed_precision(host, tmp_path):
    _gpu()
    H, W, N = 48, 64, 5
    clean, noisy = synth.make_clip(H, W, N, 1, noise_sd=15.0, seed=304, motion=True)
    src = write_raw(tmp_path / "in.raw", noisy)
    kw = "sigma=f:15 block_step=i:4 bm_range=i:5 radius=i:1 ps_num=i:2 ps_range=i:3"
    opts = "bm_error_s=s:ZSSD transform_2d_s=s:haar transform_1d_s=s:bior1.5"
    lines = run(host, "\n".join([
        f"plugin {PLUGIN_RTC}", f"source s GRAYS {W} {H} {N} {src}",
        f"invoke o BM3Dv2 clip=n:s {kw} extractor_exp=i:10 {opts}",
        f"invoke m BM3Dv2 clip=n:s {kw} precision=s:mixed",
        f"invoke f BM3Dv2 clip=n:s {kw}",
        "invoke bad BM3Dv2 clip=n:s bm_error_s=s:mse",
        f"fetch o {tmp_path / 'o.raw'} 4", f"fetch m {tmp_path / 'm.raw'} 4",
        f"fetch f {tmp_path / 'f.raw'} 4"]) + "\n")
    assert errors_of(lines) == ["BM3D: invalid 'bm_error_s': mse"], lines
    (o,) = read_raw(tmp_path / "o.raw", N, [(H, W)])
    p = O.Params(sigma=(15.0,), block_step=4, bm_range=5, radius=1, ps_num=2, ps_range=3, extractor_exp=10,
                 bm_error=O.BM_ERRORS["zssd"], transform_2d=O.TRANSFORMS["haar"],
                 transform_1d=O.TRANSFORMS["bior1.5"])
    np.testing.assert_array_equal(o, O.bm3dv2(noisy, None, p)[:, 0])
    (m,) = read_raw(tmp_path / "m.raw", N, [(H, W)])
    (f,) = read_raw(tmp_path / "f.raw", N, [(H, W)])
    assert abs(synth.psnr(m, clean[:, 0]) - synth.psnr(f, clean[:, 0])) <= 0.05


@pytest.mark.gpu
@pytest.mark.parametrize("site,message", [(1, "out of memory"), (2, "out of memory"), (3, "launch failure")])
def test_runtime_failure_reaches_the_frame_context(host, tmp_path, site, message):
    """A runtime failure inside GetFrame becomes setFilterError (source/
    source.cpp:225-231, 305-314) for that frame; the instance then serves the
    same frames correctly."""
    _gpu()
    H, W, N = 40, 48, 4
    _, noisy = synth.make_clip(H, W, N, 1, noise_sd=10.0, seed=305, motion=True)
    src = write_raw(tmp_path / "in.raw", noisy)
    kw = "sigma=f:10 block_step=i:4 bm_range=i:4 radius=i:1 ps_num=i:2 ps_range=i:2 extractor_exp=i:10"
    lib = os.path.join(LIB, "libbm3d_mi355x.so")
    lines = run(host, "\n".join([
        f"plugin {PLUGIN}", f"source s GRAYS {W} {H} {N} {src}",
        f"invoke o BM3Dv2 clip=n:s {kw}",
        f"inject {lib} {site} 0",
        f"fetch o {tmp_path / 'bad.raw'} 1",
        f"fetch o {tmp_path / 'o.raw'} 4"]) + "\n")
    errs = [l for l in lines if l.startswith("fetch-error")]
    assert len(errs) == 1 and errs[0].startswith("fetch-error 0 BM3D: ") and message in errs[0], lines
    (o,) = read_raw(tmp_path / "o.raw", N, [(H, W)])
    prm = dict(block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=2, extractor_exp=10)
    np.testing.assert_array_equal(o, O.bm3dv2(noisy, None, O.Params(sigma=(10.0,), **prm))[:, 0])
