"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and its host-side logic (Philox, hex geometry) matches published vectors and
the oracle.  No GPU compute is called here."""
import os
import re

import numpy as np

import oracle
from csim_amd import _lib
from csim_amd import ops

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "csx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(csx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [s for s in declared_symbols() if getattr(L, s, None) is None]
    assert not missing, missing
    assert set(declared_symbols()) == set(_lib.SIGNATURES), set(declared_symbols()) ^ set(_lib.SIGNATURES)


def test_version_and_error_channel():
    L = _lib.lib()
    assert b"gfx950" in L.csx_version()
    # a bad argument returns a negative code and a message, it does not abort
    rc = L.csx_hex_layout(5, 3, 1.0, None, None, 0)
    assert rc < 0 and "layout" in _lib.last_error()


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, want in kat:
        assert list(ops.philox4x32_10(ctr, key)) == want


def test_philox_streams_uniform_and_normal():
    u = ops.philox_fill(20261203, 0, 100001)
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.005
    g = ops.philox_fill(20261203, 0, 100000, normal=True)
    assert abs(g.mean()) < 0.02 and abs(g.std() - 1) < 0.02
    assert np.array_equal(g, ops.philox_fill(20261203, 0, 100000, normal=True))
    assert not np.array_equal(g[:10], ops.philox_fill(20261203, 1, 10, normal=True))


def test_hex_layout_matches_oracle():
    for layout, name, rings in ((0, "csim", 20), (0, "csim", 32), (1, "standard", 8), (1, "standard", 20)):
        hx, hy = ops.hex_layout(layout, rings)
        ox, oy, _ = oracle.hex_centers(rings, name)
        assert np.array_equal(hx, ox) and np.array_equal(hy, oy)


def test_hex_rasterization_matches_oracle_exactly():
    import math
    for layout, name, rings, lr in ((1, "standard", 4, 4 * math.sqrt(3) + 1), (0, "csim", 6, 6)):
        nsub, hexnum, nhex = ops.hex_rasterize(layout, rings, 1.0, 512, 0.9, 0.02, 40, 8, 1.0 / 30, 0.5)
        hx, hy, _ = oracle.hex_centers(rings, name)
        design = oracle.hex_design_array(hx, hy, lr, 512, 0.9, 0.02)
        on, oh = oracle.hex_interp_data(design, 40, 8, 1.0 / 30, 0.5, 0.9)
        assert nhex == hx.size
        assert np.array_equal(nsub, on) and np.array_equal(hexnum, oh)


def _tune_report_with(path):
    import json
    import subprocess
    import sys
    code = "import json; from csim_amd import _lib; print(json.dumps(_lib.tune_report()))"
    env = dict(os.environ, CSX_TUNE_CACHE=str(path))
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True, check=True)
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_tuner_cache_file_is_read_for_its_kernel_set(tmp_path):
    """The tuner's persisted choices are read back when the file names the library's
    kernel set, and ignored when it names another (a changed variant table)."""
    digest = _tune_report_with(tmp_path / "none.json")["kernels"]
    assert re.fullmatch(r"[0-9a-f]{16}", digest)
    good = tmp_path / "good.json"
    good.write_text('{"format": 1, "kernels": "%s",\n "choices": {\n  "2:27*1024x256x256": [3, 1],\n'
                    '  "3:9*128x128x1024r12": [7, 1]}}\n' % digest)
    assert _tune_report_with(good)["entries"] == 2
    other = tmp_path / "other.json"
    other.write_text(good.read_text().replace(digest, "0123456789abcdef"))
    assert _tune_report_with(other)["entries"] == 0


def test_committed_tuner_cache_matches_the_kernel_set():
    """csim_amd/tune_cache.json (the choices timed on an MI355X) names this build's
    kernel set, so GPU runs launch the kernels it lists instead of re-timing."""
    path = os.path.join(ROOT, "csim_amd", "tune_cache.json")
    assert os.path.exists(path)
    assert _tune_report_with(path)["entries"] > 0
