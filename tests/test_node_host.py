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


# ------------------------------------------------------------------ crashcat World adapter
RUN_CRASHCAT = os.path.join(HERE, "node", "crashcat_world_run.js")


def _crashcat_scenes():
    from crashcat_amd import shapes as SH
    chains = SC.Scene()
    chains.add_body(SH.box((20, 0.5, 20)), (0, -0.5, 0), motion_type=_abi.CC_MOTION_STATIC, friction=0.5)
    for k in range(6):
        chains.add_body(SH.box((0.3, 0.3, 0.3)), (0.7 * k - 2.0, 0.4 + 0.1 * k, 0.2 * k), friction=0.5)
    cap = SH.capsule(0.1, 0.3)
    prev = None
    for k in range(5):
        i = chains.add_body(cap, (4.0, 5.0 - 0.8 * k, 0.0), quaternion=(0.0, 0.0, 0.0, 1.0))
        if prev is not None:
            chains.add_point_constraint(prev, i, (0, -0.4, 0), (0, 0.4, 0))
        prev = i
    return {"cube_heap": SC.cube_heap(seed=1001, nx=3, ny=3, nz=3),
            "mixed": SC.mixed_pile(seed=1002, n=60, extent=(6.0, 6.0, 6.0), cylinders=True),
            "chains": chains,
            "terrain": SC.terrain_drop(vertices=16, n=24, drop=(0.5, 2.5)),
            "decorated": SC.decorated_pile(seed=1011, n=60, extent=(6.0, 4.0, 6.0)),
            "compound": SC.compound_pile(seed=1014, n=50, extent=(6.0, 4.0, 6.0)),
            "ccd": SC.ccd_gallery(seed=1015, bullets=10, targets=20)}


@needs_node
@pytest.mark.parametrize("which", ["cube_heap", "mixed", "chains", "terrain", "decorated", "compound", "ccd"])
def test_crashcat_world_upload_equals_the_scene_description(tmp_path, which):
    """The adapter reads crashcat's own objects (World settings, RigidBody pool with
    motionProperties, shape records, point-constraint pool) and produces exactly the upload the
    scene's C-struct description holds, hull support points (computed in JavaScript from the
    hull's points, planes and per-point faces) bit for bit included."""
    sc = _crashcat_scenes()[which]
    f = tmp_path / "world.json"
    f.write_text(json.dumps(sc.crashcat_world()))
    r = _node(f"""
const fs = require('fs');
const cw = require('./crashcat_world.js');
const d = JSON.parse(fs.readFileSync({json.dumps(str(f))}, 'utf8'));
for (const b of d.bodies.pool) b.shape = d.shapes[b.shape];
const world = {{ settings: d.settings, bodies: d.bodies, constraints: d.constraints, previousTimeStep: 0 }};
process.stdout.write(JSON.stringify(cw.describe(world)));
""")
    assert r.returncode == 0, r.stderr
    got, ref = json.loads(r.stdout), sc.desc()
    assert got["worldBodyOffsets"] == ref["worldBodyOffsets"]
    for k in ref["hull"]:
        assert got["hull"][k] == ref["hull"][k], k
    for k in ("positions", "triangles", "nodes"):
        assert got["mesh"][k] == ref["mesh"][k], k
    assert len(got["shapes"]) == len(ref["shapes"])
    for a, b in zip(got["shapes"], ref["shapes"]):
        for k in b:
            assert a[k] == b[k], k
    assert len(got["bodies"]) == len(ref["bodies"])
    for a, b in zip(got["bodies"], ref["bodies"]):
        for k in b:
            assert a[k] == b[k], k
    for k in ref["settings"]:
        assert got["settings"][k] == pytest.approx(ref["settings"][k], rel=0, abs=0), k
    assert got["pointConstraints"] == ref["pointConstraints"]


@needs_node
@pytest.mark.parametrize("what,needle", [
    ("sensor", "is a sensor"), ("compound", "compound children must be convex"),
    ("empty", "EMPTY"), ("hinge", "constraint type 2"),
    ("layers", "object layers"), ("reduction", "manifold reduction"), ("ids", "not in index order"),
    ("moving_mesh", "triangle meshes must be static")])
def test_crashcat_world_unsupported_features_are_refused(tmp_path, what, needle):
    """Features the GPU path does not have are refused with an Error before anything runs
    (so the host never steps such a world differently without knowing)."""
    sc = SC.terrain_drop(vertices=8, n=4, drop=(0.5, 2.5)) if what == "moving_mesh" else SC.cube_heap(seed=1001, nx=2, ny=2, nz=2)
    d = sc.crashcat_world()
    b = d["bodies"]["pool"]
    if what == "sensor":
        b[1]["sensor"] = True
    elif what == "compound":  # a compound inside a compound
        inner = {"type": 5, "aabb": [[0, 0, 0], [1, 1, 1]], "centerOfMass": [0, 0, 0],
                 "children": [{"position": [0, 0, 0], "quaternion": [0, 0, 0, 1], "shape": d["shapes"][b[1]["shape"]]}]}
        d["shapes"].append({"type": 5, "aabb": [[0, 0, 0], [1, 1, 1]], "centerOfMass": [0, 0, 0],
                            "children": [{"position": [0, 0, 0], "quaternion": [0, 0, 0, 1], "shape": inner}]})
        b[1]["shape"] = len(d["shapes"]) - 1
    elif what == "empty":
        d["shapes"].append({"type": 8, "aabb": [[0, 0, 0], [0, 0, 0]], "centerOfMass": [0, 0, 0]})
        b[1]["shape"] = len(d["shapes"]) - 1
    elif what == "hinge":
        d["constraints"]["pools"]["2"] = {"type": 2, "constraints": [{"_pooled": False}]}
    elif what == "layers":
        d["settings"]["layers"]["objectLayers"] = 9
    elif what == "reduction":
        b[2]["useManifoldReduction"] = False
    elif what == "ids":
        b[1]["id"], b[2]["id"] = b[2]["id"], b[1]["id"]
    elif what == "moving_mesh":
        mesh = next(i for i, s in enumerate(d["shapes"]) if s["type"] == 4)
        k = next(i for i, x in enumerate(b) if x["shape"] == mesh)
        b[k]["motionType"] = 2
    f = tmp_path / "world.json"
    f.write_text(json.dumps(d))
    r = _node(f"""
const fs = require('fs');
const cc = require('./index.js');
const d = JSON.parse(fs.readFileSync({json.dumps(str(f))}, 'utf8'));
for (const b of d.bodies.pool) b.shape = d.shapes[b.shape];
const world = {{ settings: d.settings, bodies: d.bodies, constraints: d.constraints, previousTimeStep: 0 }};
try {{ cc.fromCrashcatWorld(world); process.stdout.write('accepted'); }} catch (e) {{ process.stdout.write(e.message); }}
""")
    assert r.returncode == 0, r.stderr
    assert needle in r.stdout, r.stdout


@needs_node
@pytest.mark.gpu
@pytest.mark.parametrize("which", ["cube_heap", "mixed", "chains", "terrain", "decorated", "compound", "ccd"])
def test_crashcat_world_steps_bitwise_with_events_and_forces(tmp_path, which):
    """updateWorld on a crashcat-shaped World: FP64, bit for bit against the oracle in the state
    written back into the RigidBody objects; per step the listener's events (with the body
    objects) are the oracle's; a force and torque put on a body's motionProperties before step 10
    are integrated and cleared as crashcat does."""
    from oracle import refcpu as O
    sc = _crashcat_scenes()[which]
    f = tmp_path / "world.json"
    f.write_text(json.dumps(sc.crashcat_world()))
    steps, fstep, fbody = 40, 10, sc.num_bodies - 1
    _addon()
    r = subprocess.run(["node", RUN_CRASHCAT, str(f), "fp64", str(steps), str(fstep), str(fbody)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = [json.loads(x) for x in r.stdout.strip().splitlines()]
    events, final = lines[:-1], lines[-1]
    assert len(events) == steps and final["previousTimeStep"] == pytest.approx(1 / 60)
    orc = O.OracleWorld(sc)
    n = sc.num_bodies
    for s in range(steps):
        if s == fstep:
            fo, to = np.zeros((n, 3)), np.zeros((n, 3))
            fo[fbody] = (400, 900, 0)
            to[fbody] = (0, 0, 30)
            orc.add_forces(fo, to, wake=False)
        orc.step(1.0 / 60.0)
        c, rm = orc.contacts(), orc.removed_contacts()
        rows = [[int(k["body_a"]), int(k["body_b"]), int(k["sub_shape_a"]), int(k["sub_shape_b"]), int(k["num_points"])]
                for k in c]
        assert events[s]["added"] == [x for x, k in zip(rows, c) if not k["persisted"]], f"step {s + 1}"
        assert events[s]["persisted"] == [x for x, k in zip(rows, c) if k["persisted"]], f"step {s + 1}"
        assert events[s]["removed"] == [[int(k["body_a"]), int(k["body_b"]), int(k["sub_shape_a"]),
                                         int(k["sub_shape_b"])] for k in rm], f"step {s + 1}"
        cc = orc.ccd_bodies()  # world.ccd.ccdBodies written back by the adapter
        assert events[s]["ccd"] == [[int(k["body_index"]), int(k["hit_body_index"]), float(k["fraction"]),
                                     float(k["fraction_plus_slop"])] + [float(x) for x in k["contact_point"]]
                                    for k in cc], f"step {s + 1}"
    assert events[fstep]["forceCleared"]
    ref = orc.bodies()
    got = final["bodies"]
    for i in range(n):
        assert got[i]["position"] == ref["position"][i].tolist(), i
        assert got[i]["quaternion"] == ref["quaternion"][i].tolist(), i
        assert got[i]["com"] == ref["center_of_mass_position"][i].tolist(), i
        assert got[i]["v"] == ref["linear_velocity"][i].tolist(), i
        assert got[i]["w"] == ref["angular_velocity"][i].tolist(), i
        assert got[i]["sleeping"] == bool(ref["sleeping"][i]), i
