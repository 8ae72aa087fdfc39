"""Lifetime of library objects against their context, in fresh interpreters.

Regression test for an exit-time segmentation fault of a test process (rc 139 after every test had
passed). Cause: Python finalizers of Rasterizer / Adam / Trainer objects ran after the context
that created them had been destroyed (a module-scoped context fixture closed first, or interpreter
teardown finalised the context first); u3d_raster_destroy / u3d_trainer_destroy then read the freed
u3d_ctx and synchronised its destroyed stream. Scene-set members add a second path: the set
destroys its member trainers, so a member's own finalizer must not destroy it again. The context
now owns every object created on it (splat.Context._children): closing it releases them first,
and a released object's finalizer does nothing.

Each scenario runs in its own interpreter and must exit with status 0 and print its marker.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PRELUDE = """
import gc, sys
sys.path[:0] = [{root!r}, {tests!r}]
import numpy as np
from scenes import config0_scene
from univeral_3d_reconstruction_system_amd import splat as U
from univeral_3d_reconstruction_system_amd import synthetic as S

def objects(ctx, scenes=2):
    g, cams = config0_scene(500, 3, 64, 48, 2)
    imgs = [np.full((48, 64, 3), 128, np.uint8) for _ in cams]
    opt = U.options_for_iterations(20)
    arr = U.DeviceArray.from_numpy(ctx, np.arange(1000, dtype=np.float32))
    r = U.Rasterizer(ctx)
    out = r.forward(U.DeviceGaussians.from_host(ctx, g), cams[0], U.make_settings())
    adam = U.Adam(ctx)
    t = U.Trainer(ctx, g, cams, imgs, 1.0, opt)
    t.step()
    s = U.SceneSet(ctx, use_graph=False)
    members = [U.Trainer(ctx, g, cams, imgs, 1.0, opt) for _ in range(scenes)]
    for m in members:
        assert s.add(m)
    s.step()
    return [arr, r, out, adam, t, s] + members
""".format(root=ROOT, tests=os.path.join(ROOT, "tests"))

SCENARIOS = {
    # the context is closed while every object is alive; the objects die afterwards
    "close_first": """
ctx = U.Context(0)
objs = objects(ctx)
ctx.close()
del objs
gc.collect()
print("OK close_first", flush=True)
""",
    # the scene set is released before its member trainers' own finalizers run
    "set_then_members": """
ctx = U.Context(0)
objs = objects(ctx)
s = objs[5]
s._release()
del s, objs
gc.collect()
ctx.close()
print("OK set_then_members", flush=True)
""",
    # nothing is closed: interpreter teardown finalises context and objects in its own order
    "teardown": """
ctx = U.Context(0)
KEEP = objects(ctx)
KEEP.append(ctx)
print("OK teardown", flush=True)
""",
}


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_objects_outliving_their_context_exit_cleanly(name):
    print(f"teardown scenario {name}", flush=True)
    p = subprocess.run([sys.executable, "-c", PRELUDE + SCENARIOS[name]], capture_output=True,
                       text=True, timeout=180)
    assert p.returncode == 0, (p.returncode, p.stdout[-2000:], p.stderr[-2000:])
    assert f"OK {name}" in p.stdout
