"""The Node host: the JS shim over the N-API addon (CPU part).

tests/js/shim_host.js checks, with the real addon's scene builder and a recorder in
place of the GPU-side field calls: the reference's argument normalisation
(js/gi_probes.js:4548-4844), scene-change routing (js/gi_probes.js:3869-3937), that a
plain description and a three-like scene graph build the same buffers, the calls one
tick() makes (grid placement, solve range, clear flag, in-flight guard), transform and
deform edits, layout-error routing to a rebuild, when the classification pass runs, the
install handle's guards, and that the addon refuses scene arrays of the wrong element type
or inconsistent lengths.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "speedball_amd", "napi", "sbgi.node")


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_js_shim_host_logic():
    if not os.path.exists(ADDON):
        pytest.fail("the N-API addon is not built (run __graft_entry__.build())")
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "shim_host.js")], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    assert "passed 11" in r.stdout
