"""The JavaScript host (Node N-API addon + crashcat_amd/node/index.js) over the same C ABI.

CPU: the struct layouts the addon reads from the C header equal the ctypes mirror, packing
round-trips, unknown fields and a missing GPU fail loudly.  GPU: the cube heap stepped from
JavaScript with updateWorld is bit-identical to the FP64 oracle (bodies and pair set).
"""
import ctypes as C
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from crashcat_amd import _abi
from crashcat_amd import build as B
from crashcat_amd import scene as SC
from crashcat_amd._lib import BODY_STATE_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
ADDON = os.path.join(HERE, "..", "crashcat_amd", "node", "crashcat_amd.node")
STEP = os.path.join(HERE, "node", "step_world.js")
EVENTS = os.path.join(HERE, "node", "events_world.js")

needs_node = pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")


def _addon():
    if not os.path.exists(ADDON):
        if not B.node_available():
            pytest.skip("Node N-API headers are not installed")
        B.build_node()
    return ADDON


def _node(script):
    _addon()
    r = subprocess.run(["node", "-e", script], capture_output=True, text=True, timeout=60,
                       cwd=os.path.join(HERE, "..", "crashcat_amd", "node"))
    return r


@needs_node
def test_addon_layouts_match_the_ctypes_mirror():
    r = _node("process.stdout.write(JSON.stringify(require('./index.js').layout))")
    assert r.returncode == 0, r.stderr
    L = json.loads(r.stdout)
    mirror = {"cc_settings": _abi.Settings, "cc_shape_desc": _abi.ShapeDesc, "cc_body_desc": _abi.BodyDesc,
              "cc_body_state": _abi.BodyState, "cc_point_constraint_desc": _abi.PointConstraintDesc,
              "cc_step_counters": _abi.StepCounters, "cc_contact_out": _abi.ContactOut,
              "cc_contact_removed": _abi.ContactRemoved}
    for name, st in mirror.items():
        assert L[name]["size"] == C.sizeof(st), name
        names = [f for f, _ in st._fields_ if not f.endswith("_")]
        assert sorted(L[name]["fields"]) == sorted(names), name
        for f in names:
            assert L[name]["fields"][f][0] == getattr(st, f).offset, (name, f)


@needs_node
def test_pack_round_trip_and_unknown_field():
    r = _node("""
        const cc = require('./index.js');
        const b = cc.pack('cc_body_desc', [{shape: 3, position: [1.5, -2, 3], quaternion: [0, 0, 0.6, 0.8],
                                            inv_mass: 0.25, collision_mask: 4294967295}]);
        const r = cc.unpack('cc_body_desc', b)[0];
        let threw = false;
        try { cc.pack('cc_body_desc', [{no_such_field: 1}]); } catch (e) { threw = /no field/.test(e.message); }
        process.stdout.write(JSON.stringify({shape: r.shape, position: r.position, q: r.quaternion,
                                             m: r.inv_mass, mask: r.collision_mask, threw: threw}));
    """)
    assert r.returncode == 0, r.stderr
    v = json.loads(r.stdout)
    assert v == {"shape": 3, "position": [1.5, -2, 3], "q": [0, 0, 0.6, 0.8], "m": 0.25, "mask": 4294967295,
                 "threw": True}


@needs_node
@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_gpu_throws():
    r = _node("""
        const cc = require('./index.js');
        try { cc.createWorld({settings: {}, bodies: []}); process.stdout.write('no error'); }
        catch (e) { process.stdout.write(e.message); }
    """)
    assert r.returncode == 0, r.stderr
    assert "no usable HIP device" in r.stdout


@needs_node
def test_update_world_refuses_validate_callbacks():
    """onBodyPairValidate / onContactValidate decide what a step does and cannot run inside the
    GPU step: updateWorld refuses them before stepping (the contact callbacks are delivered
    after the step instead)."""
    r = _node("""
const cc = require('./index.js');
const w = Object.create(cc.GpuWorld.prototype);
for (const k of ['onBodyPairValidate', 'onContactValidate']) {
  try { cc.updateWorld(w, { [k]() { return true; } }, 1 / 60); console.log('stepped'); }
  catch (e) { console.log(e.message); }
}""")
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 2
    assert "not available on the GPU step: onBodyPairValidate" in lines[0]
    assert "not available on the GPU step: onContactValidate" in lines[1]


@needs_node
@pytest.mark.gpu
def test_node_listener_receives_the_oracle_contact_events(tmp_path):
    """updateWorld with onContactAdded / onContactPersisted / onContactRemoved from JavaScript:
    per step, the same events as the FP64 oracle (keys and point counts), on a heap that lands,
    settles and sleeps."""
    from oracle import refcpu as O
    sc = SC.cube_heap(seed=1001, nx=3, ny=3, nz=3)
    desc = tmp_path / "world.json"
    desc.write_text(json.dumps(sc.desc()))
    steps = 200
    _addon()
    r = subprocess.run(["node", EVENTS, str(desc), "fp64", str(steps)], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    got = [json.loads(x) for x in r.stdout.strip().splitlines()]
    assert len(got) == steps
    orc = O.OracleWorld(sc)
    totals = [0, 0, 0]
    for s in range(steps):
        orc.step(1.0 / 60.0)
        c, rm = orc.contacts(), orc.removed_contacts()
        rows = [[int(k["body_a"]), int(k["body_b"]), int(k["sub_shape_a"]), int(k["sub_shape_b"]),
                 int(k["num_points"])] for k in c]
        pers = [x for x, k in zip(rows, c) if k["persisted"]]
        added = [x for x, k in zip(rows, c) if not k["persisted"]]
        rem = [[int(k["body_a"]), int(k["body_b"]), int(k["sub_shape_a"]), int(k["sub_shape_b"])] for k in rm]
        assert got[s]["added"] == added, f"step {s + 1}: added"
        assert got[s]["persisted"] == pers, f"step {s + 1}: persisted"
        assert got[s]["removed"] == rem, f"step {s + 1}: removed"
        totals = [totals[0] + len(added), totals[1] + len(pers), totals[2] + len(rem)]
    assert min(totals) > 0, totals


@needs_node
@pytest.mark.gpu
def test_node_snapshot_restore_repeats_the_run(tmp_path):
    """world.snapshot() after 20 steps, restored into a second world made from the same
    description: the next 20 steps of both are bit-identical (cc_snapshot / cc_restore)."""
    sc = SC.terrain_drop(vertices=16, n=24, drop=(0.5, 2.5), spheres=True)
    desc = tmp_path / "world.json"
    desc.write_text(json.dumps(sc.desc()))
    _addon()
    script = """
const fs = require('fs');
const cc = require('./index.js');
const desc = JSON.parse(fs.readFileSync(process.argv[1], 'utf8'));
const a = cc.createWorld(desc, { precision: 'fp32', maxPairsPerBody: 16 });
for (let i = 0; i < 20; i++) cc.updateWorld(a, undefined, 1 / 60);
const snap = a.snapshot();
for (let i = 0; i < 20; i++) cc.updateWorld(a, undefined, 1 / 60);
const b = cc.createWorld(desc, { precision: 'fp32', maxPairsPerBody: 16 });
b.restore(snap);
const resumed = b.previousTimeStep;
for (let i = 0; i < 20; i++) cc.updateWorld(b, undefined, 1 / 60);
const x = new Uint8Array(a.bodyStateBuffer()), y = new Uint8Array(b.bodyStateBuffer());
let same = x.length === y.length;
for (let i = 0; same && i < x.length; i++) same = x[i] === y[i];
console.log(JSON.stringify({ same, bytes: snap.byteLength, resumed }));
"""
    r = subprocess.run(["node", "-e", script, str(desc)], capture_output=True, text=True, timeout=120,
                       cwd=os.path.join(HERE, "..", "crashcat_amd", "node"))
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["same"] and out["bytes"] > 0
    assert out["resumed"] == pytest.approx(1 / 60)


@needs_node
@pytest.mark.gpu
@pytest.mark.parametrize("which", ["cube_heap", "terrain"])
def test_node_update_world_cube_heap_bitwise(tmp_path, which):
    from oracle import refcpu as O
    sc = SC.cube_heap(seed=1001) if which == "cube_heap" else SC.terrain_drop(vertices=24, n=40, drop=(0.5, 3.0))
    desc = tmp_path / "world.json"
    desc.write_text(json.dumps(sc.desc()))
    out, pairs = tmp_path / "bodies.bin", tmp_path / "pairs.bin"
    steps = 30
    _addon()
    r = subprocess.run(["node", STEP, str(desc), "fp64", str(steps), str(out), str(pairs)], capture_output=True,
                       text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, dtype=BODY_STATE_DTYPE)
    orc = O.OracleWorld(sc)
    for _ in range(steps):
        orc.step(1.0 / 60.0)
    ref = orc.bodies()
    assert len(got) == len(ref)
    for f in ["position", "quaternion", "center_of_mass_position", "linear_velocity", "angular_velocity", "sleeping"]:
        assert np.array_equal(got[f], ref[f]), f
    gp = np.fromfile(pairs, dtype=np.int32).reshape(-1, 2)
    op = np.asarray(orc.pairs(), dtype=np.int32).reshape(-1, 2)
    assert np.array_equal(gp, op)
    assert json.loads(r.stdout.strip().splitlines()[-1])["overflow"] == 0
