"""The `marmel` command-line renderer (marmel-trace_amd/bin/marmel) with the
reference's flags: Whitted through render_image_ex's drop-in (adaptive by
default, --no-adaptive, --adaptive-max, --adaptive-tau) and the streak flare
(--lens-flare, --flare-intensity, --flare-threshold). The written BMP equals
the oracle's adaptive render + flare (which equal the reference's,
tests/test_oracle.py and tests/test_flare.py) within the Whitted bar, and the
library's own result byte for byte; bad flag values exit with code 1."""
import os
import subprocess

import numpy as np
import pytest

import checkers as ck
import flare_cases as fc
from marmel_amd import GpuScene, HostScene, encode_bmp, flare_apply, native

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(native.LIB_DIR), "bin", "marmel")


def run(args, tmp_path):
    out = str(tmp_path / "img.bmp")
    r = subprocess.run([BIN, "--out", out] + args, capture_output=True, text=True, timeout=120)
    return r, out


def test_cli_whitted_adaptive_with_flare(tmp_path):
    w, h, spp, depth = 96, 72, 4, 5
    args = ["--scene-gen", "spheres", "--seed", "7", "--width", str(w), "--height", str(h), "--samples", str(spp),
            "--depth", str(depth), "--no-pathtrace", "--adaptive-max", "12", "--adaptive-tau", "0.03",
            "--lens-flare", "--flare-intensity", "1.5", "--flare-threshold", "0.7"]
    r, out = run(args, tmp_path)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    data = open(out, "rb").read()
    hs = HostScene()
    hs.generate(native.GEN_SPHERES, seed=7)
    hs.build_bvh()
    cam = hs.camera(w / h)
    gs = GpuScene(hs.soa())
    try:
        rgb, _, ns, tot, mx = gs.render_adaptive(cam, w, h, spp, depth, 12, 0.03)
    finally:
        gs.close()
    p = fc.params(intensity=1.5, threshold=0.7)
    lib = np.ascontiguousarray(rgb.copy())
    assert flare_apply(lib, p) == 0
    assert data == encode_bmp(lib)
    assert f"samples: {tot} total, {mx} max per pixel" in r.stdout
    prims, attrs, mats, sky, wl, wm = hs.arrays()
    o = ck.OracleScene(ck.SoaHolder(prims, attrs, mats, sky, wl, wm).soa)
    orgb, _, ons, otot, omx = o.render_adaptive(cam, w, h, spp, depth, 12, 0.03)
    rc, oref = __import__("test_flare").oracle_flare(orgb, p)
    d = np.abs(lib.astype(int) - oref.astype(int)).max(-1)
    print("cli vs oracle: within 1/255", (d <= 1).mean(), "sample totals", tot, otot)
    assert (d <= 1).mean() >= 0.999 and tot == otot


def test_cli_fixed_whitted_and_bad_flags(tmp_path):
    r, out = run(["--no-pathtrace", "--no-adaptive", "--width", "40", "--height", "30", "--samples", "2"], tmp_path)
    assert r.returncode == 0 and "samples: 2400 total, 2 max per pixel" in r.stdout
    for bad in (["--flare-threshold", "1.5"], ["--flare-intensity", "-1"], ["--samples", "0"], ["--bogus"]):
        r, _ = run(bad, tmp_path)
        assert r.returncode == 1, bad
