"""The plugin boundary: bbvh-hip.so exports the four symbols of the plugin
contract (librta/plugins.cpp:82-85), the rta-hip host loads it through
them, and on a GPU the frame loop reproduces the oracle's closest hits."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
from rta_amd import abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLUGIN = os.path.join(ROOT, "rta_amd", "lib", "bbvh-hip.so")
DRIVER = os.path.join(ROOT, "rta_amd", "bin", "rta-hip")
F32MAX = np.float32(np.finfo(np.float32).max)


def run_driver(*args, timeout=300):
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=timeout)


def test_plugin_exports_the_four_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", PLUGIN], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert {"description", "initialize", "parse_cmdline", "create_rt_set"} <= exported
    lib = C.CDLL(PLUGIN)
    lib.description.restype = C.c_char_p
    assert b"bbvh-hip" in lib.description()


def test_plugin_option_parsing():
    lib = C.CDLL(PLUGIN)
    lib.parse_cmdline.argtypes = [C.c_int, C.POINTER(C.c_char_p)]

    def parse(*args):
        argv = (C.c_char_p * (len(args) + 1))(b"x", *[a.encode() for a in args])
        return lib.parse_cmdline(len(args) + 1, argv)

    assert parse("-b", "lbvh", "-t", "cis", "-l", "2f4", "-v") == 0
    assert parse("--trav", "packed", "--block", "64", "--refill", "8", "--steps", "4",
                 "--leaf-min", "12") == 0
    assert parse("-b", "median") == 0        # the host-built median tree (the default)
    assert parse("-b", "sah") == 0           # any value but lbvh selects it (main.cpp:88)
    assert parse("-b") != 0                  # a value is required
    assert parse("-s") == 0                  # shadow tracer (persistent layout)
    assert parse("-s", "--trav", "plain") != 0   # no shadow form of the plain kernel
    assert parse("-A") != 0                  # Aila's box variant is not provided
    assert parse("--bogus") != 0


def test_host_reports_missing_plugin():
    r = run_driver("no-such-plugin", timeout=60)
    assert r.returncode != 0
    assert "no library for plugin 'no-such-plugin'" in r.stderr


def test_host_reports_plugin_missing_a_symbol(tmp_path):
    """A library that loads but lacks the contract's symbols (the dlsym error
    path of librta/plugins.cpp:20-28): the host names the failure and stops
    before any GPU work."""
    src = tmp_path / "partial.c"
    src.write_text('char *description(void) { return "partial"; }\n')
    lib = tmp_path / "partial.so"
    subprocess.run(["gcc", "-shared", "-fPIC", "-o", str(lib), str(src)], check=True)
    r = run_driver(str(lib), timeout=60)
    assert r.returncode != 0
    assert "initialize" in r.stderr or "parse_cmdline" in r.stderr or "create_rt_set" in r.stderr


def test_host_rejects_bad_usage():
    r = run_driver("-r", "bogus", timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("trav", ["plain", "packed", "persistent"])
def test_driver_frame_matches_oracle(gpu, tmp_path, trav):
    dump = str(tmp_path / "hits.bin")
    r = run_driver("-r", "256x256", "--frames", "2", "--dump-hits", dump, "bbvh-hip", "--",
                   "-b", "lbvh", "-t", trav)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "average rps per frame:" in r.stdout
    got = np.fromfile(dump, dtype=abi.HIT_DTYPE)
    tris = scenes.sphere_scene()
    ref_b = oracle.lbvh_build(tris)
    o, d, t = scenes.sphere_camera(256, 256).host_rays()
    ref, _, _ = oracle.trace_cis(ref_b["tris"], ref_b["nodes"], o, d, t, 256, 256)
    hit = ref["t"] != F32MAX
    assert hit.mean() > 0.2
    assert np.array_equal(got["t"] != F32MAX, hit)
    assert np.array_equal(got["t"][hit], ref["t"][hit])
    assert np.array_equal(got["ref"][hit], ref["ref"][hit])


@pytest.mark.gpu
def test_driver_rebuild_and_dump(gpu, tmp_path):
    dump = str(tmp_path / "tree.bvh")
    r = run_driver("-r", "320x180", "--scene", "terrain:64:500", "--frames", "2", "--rebuild",
                   "--dump-bvh", dump, "bbvh-hip", "--", "-b", "lbvh", "-v")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "LBVH build of" in r.stdout
    lines = open(dump).read().splitlines()
    n = 2 * 64 * 64 + 500
    assert lines[0] == "bvh" and int(lines[1]) == 2 * n - 1
    assert lines[2].split()[1] == "I"


@pytest.mark.gpu
@pytest.mark.parametrize("trav", ["packed", "persistent"])
def test_driver_shadow_tracer(gpu, tmp_path, trav):
    """`-s` selects the shadow tracer: its records are the oracle's
    restatement of trace_shadow_cis on the same tree and rays."""
    dump = str(tmp_path / "hits.bin")
    r = run_driver("-r", "256x256", "--frames", "1", "--dump-hits", dump, "bbvh-hip", "--",
                   "-b", "lbvh", "-s", "--trav", trav)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(dump, dtype=abi.HIT_DTYPE)
    tris = scenes.sphere_scene()
    ref_b = oracle.lbvh_build(tris)
    o, d, t = scenes.sphere_camera(256, 256).host_rays()
    ref, _ = oracle.trace_kernel(ref_b["tris"], ref_b["nodes"], o, d, t, any_hit=True)
    hit = ref["t"] != F32MAX
    assert hit.mean() > 0.2
    assert np.array_equal(got["t"] != F32MAX, hit)
    for f in ("t", "beta", "gamma", "ref"):
        assert np.array_equal(got[f][hit], ref[f][hit])


@pytest.mark.gpu
@pytest.mark.parametrize("builder", [(), ("-b", "median")])
def test_driver_median_builder(gpu, tmp_path, builder):
    """The plugin's default builder, as the reference's (bbvh-cuda/main.cpp:
    28, 158-163): the spatial-median tree built on the host and uploaded.
    Two frames with a rebuild from host triangles between them; the dumped
    .bvh/.tri bodies are the oracle's median tree and leaf-ordered triangles
    in the reference's text format, the hits the CPU child-is tracer's."""
    from helpers import bvh_dump_lines, tri_dump_lines
    hits, prefix = str(tmp_path / "hits.bin"), str(tmp_path / "m")
    r = run_driver("-r", "256x256", "--frames", "2", "--rebuild", "-M", prefix, "--dump-hits",
                   hits, "bbvh-hip", "--", *builder, "-v")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "median build of" in r.stdout and "LBVH build of" not in r.stdout
    tris = scenes.sphere_scene()
    mb = oracle.median_build(tris)
    assert open(prefix + ".bvh").read().splitlines() == bvh_dump_lines(mb["nodes"])
    assert open(prefix + ".tri").read().splitlines() == tri_dump_lines(mb["tris"])
    got = np.fromfile(hits, dtype=abi.HIT_DTYPE)
    o, d, t = scenes.sphere_camera(256, 256).host_rays()
    ref, _, _ = oracle.trace_cis(mb["tris"], mb["nodes"], o, d, t, 256, 256)
    hit = ref["t"] != F32MAX
    assert hit.mean() > 0.2
    assert np.array_equal(got["t"] != F32MAX, hit)
    for f in ("t", "beta", "gamma", "ref"):
        assert np.array_equal(got[f][hit], ref[f][hit])
