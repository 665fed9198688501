'use strict';
// A crashcat World stepped on the GPU.  The upload is read from the world's own objects:
// world.settings (WorldSettings, src/world-settings.ts:4-200), world.bodies.pool (RigidBody,
// src/body/rigid-body.ts:142-252, with its shape and motionProperties), and the point constraints
// of world.constraints.pools (src/constraints/point-constraint.ts:21-25).  Each step writes the
// results back into the same body objects (position, quaternion, centre of mass, velocities,
// sleeping, aabb), clears the force and torque accumulators as crashcat does (src/update.ts:
// 234-238), advances world.previousTimeStep, and calls the listener's contact callbacks with the
// RigidBody objects.  Body indices on the GPU are the pool's indices (a removed body's slot
// becomes an inert placeholder), so the sort keys, which crashcat computes from body indices
// (contact-constraints.ts:1222-1231), are the same.  Features this path does not have are
// refused with an Error before anything runs: sensors, CCD (MotionQuality.LINEAR_CAST) in a world
// it cannot cast in (the library refuses it), empty/user shapes,
// compounds holding compounds, meshes or planes, moving triangle meshes,
// constraint types other than points, enhanced internal-edge removal, manifold reduction turned
// off, kinematic-vs-non-dynamic collision, returnDeepestPoint, a warm-start impulse ratio other
// than 1, more than 8 object layers, and live bodies whose id order differs from their index
// order (ids order contact pairs and break sort-key ties, contact-constraints.ts:448-453,
// 1240-1264).  Node 12 syntax.

const cc = require('./index.js');

// ShapeType (src/shapes/shapes.ts:37-62) -> the C ABI's shape codes
const SHAPE_TO_CC = { 0: 0, 1: 1, 2: 2, 3: 3, 4: 4, 5: 7, 9: 5, 11: 6 };
const SHAPE_NAMES = { 8: 'EMPTY' };
// decorators (ShapeType.TRANSFORMED, SCALED, OFFSET_CENTER_OF_MASS) and the C ABI's CC_DECO_* bits
const DECO_BIT = { 7: 1, 6: 2, 10: 4 };
const MOTION_STATIC = 0;
const POINT_CONSTRAINT = 0;  // ConstraintType.POINT (src/constraints/constraint-id.ts:25-33)
const TRIANGLE_STRIDE = 8;   // src/shapes/utils/triangle-mesh-data.ts:4-14
const MAX_LAYERS = 8;

function isCrashcatWorld(w) {
  return !!(w && w.bodies && Array.isArray(w.bodies.pool) && w.settings && w.settings.layers);
}

function fail(msg) { throw new Error(`crashcat_amd: ${msg}`); }

// combineMaterial (src/constraints/combine-material.ts:20-40): the higher-priority mode of the two
const MODE_PRIORITY = { 3: 4, 2: 3, 4: 2, 1: 1, 0: 0 };
function combine(a, b, ma, mb) {
  const mode = MODE_PRIORITY[ma] >= MODE_PRIORITY[mb] ? ma : mb;
  if (mode === 1) return a * b;
  if (mode === 2) return Math.min(a, b);
  if (mode === 3) return Math.max(a, b);
  if (mode === 4) return Math.sqrt(a * b);
  return (a + b) / 2;
}

function vec(v, n) {
  const out = new Array(n);
  for (let i = 0; i < n; i++) out[i] = Number(v[i]);
  return out;
}
function box3(b) { return { min: vec(b[0], 3), max: vec(b[1], 3) }; }

// The convex-radius-excluded support points of a hull (ConvexHullNoConvexSupport,
// src/shapes/convex-hull.ts:732-786), component by component in the same order as the scene
// builder's shrunk_points (crashcat_amd/shapes.py), so both give the same bits.
function shrunkPoints(shape) {
  const cr = shape.convexRadius;
  const out = [];
  const dot = (a, b) => (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
  const cross = (a, b) => [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]];
  for (const pt of shape.points) {
    const p = pt.position;
    const nf = pt.numFaces;
    const c1 = shape.planes[pt.faces[0]].normal;
    const c2 = shape.planes[pt.faces[nf > 1 ? 1 : 0]].normal;
    let c3;
    if (nf === 3) {
      c3 = shape.planes[pt.faces[2]].normal;
    } else {
      const pp = cross(c1, c2);
      const l = Math.sqrt(dot(pp, pp));
      c3 = l > 0 ? [pp[0] / l, pp[1] / l, pp[2] / l] : [0, 0, 0];
    }
    const w1 = -dot(c1, p) + cr, w2 = -dot(c2, p) + cr, w3 = -dot(c3, p) + (nf === 3 ? cr : 0);
    const c23 = cross(c2, c3), c31 = cross(c3, c1), c12 = cross(c1, c2);
    const den = dot(c1, c23);
    if (nf === 1 || den === 0) {
      out.push(p[0] - c1[0] * cr, p[1] - c1[1] * cr, p[2] - c1[2] * cr);
    } else {
      for (let k = 0; k < 3; k++) out.push(-((w1 * c23[k] + w2 * c31[k]) + w3 * c12[k]) / den);
    }
  }
  return out;
}

// The shrunk support points of a scaled hull (setConvexHullNoConvexSupportScaled,
// src/shapes/convex-hull.ts:842-935), in the order of the scene builder's scaled_shrunk_points
// (crashcat_amd/shapes.py): points scaled, plane normals times 1 / scale and normalised, the
// radius times the smallest |scale|.
function scaledShrunkPoints(shape, sc) {
  const inv = [1 / sc[0], 1 / sc[1], 1 / sc[2]];
  const r = shape.convexRadius * Math.min(Math.abs(sc[0]), Math.abs(sc[1]), Math.abs(sc[2]));
  const out = [];
  const dot = (a, b) => (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
  const cross = (a, b) => [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]];
  const unit = (v) => {
    const l2 = dot(v, v);
    const f = l2 > 0 ? 1 / Math.sqrt(l2) : 0;
    return [v[0] * f, v[1] * f, v[2] * f];
  };
  const nrm = (f) => { const n = shape.planes[f].normal; return unit([n[0] * inv[0], n[1] * inv[1], n[2] * inv[2]]); };
  for (const pt of shape.points) {
    const q = pt.position;
    const p = [q[0] * sc[0], q[1] * sc[1], q[2] * sc[2]];
    const nf = pt.numFaces;
    const c1 = nrm(pt.faces[0]);
    const c2 = nrm(pt.faces[nf > 1 ? 1 : 0]);
    const c3 = nf === 3 ? nrm(pt.faces[2]) : unit(cross(c1, c2));
    const w1 = -dot(c1, p) + r, w2 = -dot(c2, p) + r, w3 = -dot(c3, p) + (nf === 3 ? r : 0);
    const c23 = cross(c2, c3), c31 = cross(c3, c1), c12 = cross(c1, c2);
    const den = dot(c1, c23);
    if (nf === 1 || den === 0) {
      out.push(p[0] + c1[0] * -r, p[1] + c1[1] * -r, p[2] + c1[2] * -r);
    } else {
      for (let k = 0; k < 3; k++) out.push(-((w1 * c23[k] + w2 * c31[k]) + w3 * c12[k]) / den);
    }
  }
  return out;
}

// A chain of decorators over a leaf, as one scale and one local transform (the product of the
// chain's scales; at most one transformed decorator).
function flattenDecorators(sh) {
  let deco = 0, transforms = 0;
  let scale = [1, 1, 1], pos = [0, 0, 0], quat = [0, 0, 0, 1];
  let s = sh;
  while (DECO_BIT[s.type] !== undefined) {
    deco |= DECO_BIT[s.type];
    if (s.type === 7) scale = [scale[0] * s.scale[0], scale[1] * s.scale[1], scale[2] * s.scale[2]];
    else if (s.type === 6) {
      if (transforms++) fail('nested transformed shapes are not on the GPU path');
      pos = vec(s.position, 3);
      quat = vec(s.quaternion, 4);
    }
    s = s.shape;
  }
  return { leaf: s, deco, scale, pos, quat };
}

// Settings in the C ABI's cc_settings field names.
function settingsDesc(s) {
  const L = s.layers;
  const n = L.objectLayers;
  if (n > MAX_LAYERS) fail(`${n} object layers; the GPU path has at most ${MAX_LAYERS}`);
  if (s.narrowphase.returnDeepestPoint) fail('narrowphase.returnDeepestPoint is not on the GPU path');
  if (s.solver.warmStartImpulseRatio !== undefined && s.solver.warmStartImpulseRatio !== 1) {
    fail('solver.warmStartImpulseRatio other than 1 is not on the GPU path');
  }
  if (s.narrowphase.useManifoldReduction === false) fail('narrowphase.useManifoldReduction = false is not on the GPU path');
  // two object layers collide when their pair is enabled and so is their broadphase layers'
  // pair (layers.ts:73-112)
  const pairs = new Array(MAX_LAYERS * MAX_LAYERS).fill(0);
  const nb = L.broadphaseLayers;
  for (let a = 0; a < n; a++) {
    for (let b = 0; b < n; b++) {
      const ba = L.objectLayerToBroadphaseLayer[a], bb = L.objectLayerToBroadphaseLayer[b];
      const bp = ba === undefined || bb === undefined || nb === 0 ? 1 : L.broadphasePairs[ba * nb + bb];
      pairs[a * MAX_LAYERS + b] = L.objectLayerPairs[a * n + b] && bp ? 1 : 0;
    }
  }
  return {
    gravity: vec(s.gravity, 3),
    speculative_contact_distance: s.narrowphase.speculativeContactDistance,
    manifold_tolerance: s.narrowphase.manifoldTolerance,
    collision_tolerance: s.narrowphase.collisionTolerance,
    penetration_tolerance: s.narrowphase.penetrationTolerance,
    normal_cos_max_delta_rotation: s.narrowphase.normalCosMaxDeltaRotation,
    contact_point_preserve_lambda_max_dist_sq: s.contacts.contactPointPreserveLambdaMaxDistSq,
    min_velocity_for_restitution: s.solver.minVelocityForRestitution,
    penetration_slop: s.solver.penetrationSlop,
    baumgarte_factor: s.solver.baumgarteFactor,
    max_penetration_distance: s.solver.maxPenetrationDistance,
    time_before_sleep: s.sleeping.timeBeforeSleep,
    point_velocity_sleep_threshold: s.sleeping.pointVelocitySleepThreshold,
    velocity_iterations: s.solver.velocityIterations,
    position_iterations: s.solver.positionIterations,
    gravity_enabled: s.gravityEnabled ? 1 : 0,
    warm_starting: s.solver.warmStarting ? 1 : 0,
    allow_sleeping: s.sleeping.allowSleeping ? 1 : 0,
    num_object_layers: n,
    object_layer_pairs: pairs,
    collide_with_backfaces: s.narrowphase.collideWithBackfaces ? 1 : 0,
    collide_only_with_active_edges: s.narrowphase.collideOnlyWithActiveEdges ? 1 : 0,
    ccd_linear_cast_threshold: s.ccd.linearCastThreshold,
    ccd_linear_cast_max_penetration: s.ccd.linearCastMaxPenetration,
    ccd_min_velocity_for_restitution: s.ccd.minVelocityForRestitution,
  };
}

// Shapes (deduplicated by object), hull and mesh buffers, in the C ABI's layout.
function shapeTables(pool) {
  const index = new Map();
  const shapes = [];
  const hull = { points: [], shrunkPoints: [], faceFirstVertex: [], faceNumVertices: [], faceVertexIndices: [],
                 planeNormals: [], planeConstants: [] };
  const mesh = { positions: [], triangles: [], nodes: [] };
  let np = 0, nf = 0, nfv = 0, nmv = 0, nmt = 0, nmn = 0;
  // compounds whose children are still to be appended: the children follow every body's shape,
  // each a leaf desc with the child's position and quaternion as its local transform
  const pending = [];
  const makeDesc = (outer, child) => {
    const F = flattenDecorators(outer);
    const sh = F.leaf;
    const t = SHAPE_TO_CC[sh.type];
    if (t === undefined) fail(`shape type ${SHAPE_NAMES[sh.type] || sh.type} is not on the GPU path`);
    if (F.deco && (t === 4 || t === 6 || t === 7)) fail('decorated triangle meshes, planes and compounds are not on the GPU path');
    if (child) {
      if (t === 4 || t === 6 || t === 7 || (F.deco & 2)) {
        fail('compound children must be convex shapes (scaled allowed), not transformed, compound, meshes or planes');
      }
      F.deco |= 2;
      F.pos = vec(child.position, 3);
      F.quat = vec(child.quaternion, 4);
    }
    const aabb = box3(outer.aabb);
    const d = { type: t, radius: 0, half_height: 0, half_extents: [0, 0, 0], convex_radius: 0,
                center_of_mass: vec(sh.centerOfMass, 3), aabb_min: aabb.min, aabb_max: aabb.max,
                hull_first_point: 0, hull_num_points: 0, hull_first_face: 0, hull_num_faces: 0,
                mesh_first_triangle: 0, mesh_num_triangles: 0, mesh_first_node: 0, mesh_num_nodes: 0,
                decorated: F.deco, scale: F.scale, local_position: F.pos, local_quaternion: F.quat,
                compound_first_child: 0, compound_num_children: 0 };
    d.center_of_mass = vec(outer.centerOfMass, 3);
    if (t === 7) pending.push({ d, children: sh.children });
    if (sh.type === 0) {
      d.radius = sh.radius;
    } else if (sh.type === 1) {
      d.half_extents = vec(sh.halfExtents, 3);
      d.convex_radius = sh.convexRadius;
    } else if (sh.type === 2) {
      d.radius = sh.radius;
      d.half_height = sh.halfHeightOfCylinder;
    } else if (sh.type === 9) {
      d.radius = sh.radius;
      d.half_height = sh.halfHeight;
      d.convex_radius = sh.convexRadius;
    } else if (sh.type === 11) {
      d.half_extents = vec(sh.plane.normal, 3);
      d.radius = sh.plane.constant;
      d.half_height = sh.halfExtent;
    } else if (sh.type === 3) {
      d.convex_radius = sh.convexRadius;
      d.hull_first_point = np;
      d.hull_num_points = sh.points.length;
      d.hull_first_face = nf;
      d.hull_num_faces = sh.faces.length;
      for (const p of sh.points) hull.points.push(p.position[0], p.position[1], p.position[2]);
      for (const v of (F.deco & 1 ? scaledShrunkPoints(sh, F.scale) : shrunkPoints(sh))) hull.shrunkPoints.push(v);
      sh.faces.forEach((f, i) => {
        hull.faceFirstVertex.push(nfv);
        hull.faceNumVertices.push(f.numVertices);
        for (let k = 0; k < f.numVertices; k++) hull.faceVertexIndices.push(sh.vertexIndices[f.firstVertex + k]);
        nfv += f.numVertices;
        const pl = sh.planes[i];
        hull.planeNormals.push(pl.normal[0], pl.normal[1], pl.normal[2]);
        hull.planeConstants.push(pl.constant);
      });
      np += sh.points.length;
      nf += sh.faces.length;
    } else if (sh.type === 4) {
      const tb = sh.data.triangleBuffer, tc = sh.data.triangleCount;
      const nodes = sh.bvh.buffer;
      d.mesh_first_triangle = nmt;
      d.mesh_num_triangles = tc;
      d.mesh_first_node = nmn;
      d.mesh_num_nodes = nodes.length / 8;
      for (const v of sh.data.positions) mesh.positions.push(v);
      for (let t = 0; t < tc; t++) {
        const o = t * TRIANGLE_STRIDE;  // a, b, c, normal, active edges, material
        mesh.triangles.push(tb[o] + nmv, tb[o + 1] + nmv, tb[o + 2] + nmv, tb[o + 6]);
      }
      for (const v of nodes) mesh.nodes.push(v);
      nmv += sh.data.positions.length / 3;
      nmt += tc;
      nmn += nodes.length / 8;
    }
    return d;
  };
  const add = (outer) => {
    if (index.has(outer)) return index.get(outer);
    const d = makeDesc(outer, null);
    index.set(outer, shapes.length);
    shapes.push(d);
    return shapes.length - 1;
  };
  const finish = () => {
    for (const { d, children } of pending) {
      if (!children || !children.length) fail('a compound has no children');
      d.compound_first_child = shapes.length;
      d.compound_num_children = children.length;
      for (const c of children) shapes.push(makeDesc(c.shape, c));
    }
    pending.length = 0;
  };
  return { add, finish, shapes, hull, mesh };
}

function bodyDesc(b, shapeIndex) {
  const m = b.motionProperties;
  if (b.sensor) fail(`body ${b.index} is a sensor; sensors are not on the GPU path`);
  if (b.enhancedInternalEdgeRemoval) fail(`body ${b.index} asks for enhanced internal-edge removal, which is not on the GPU path`);
  if (b.useManifoldReduction === false) fail(`body ${b.index} turns manifold reduction off, which is not on the GPU path`);
  if (b.collideKinematicVsNonDynamic) fail(`body ${b.index} collides kinematic vs non-dynamic, which is not on the GPU path`);
  if (b.shape.type === 4 && b.motionType !== MOTION_STATIC) fail(`body ${b.index}: triangle meshes must be static`);
  return {
    shape: shapeIndex, motion_type: b.motionType, object_layer: b.objectLayer,
    collision_groups: b.collisionGroups >>> 0, collision_mask: b.collisionMask >>> 0,
    allow_sleeping: m.allowSleeping ? 1 : 0, allowed_dofs: m.allowedDegreesOfFreedom,
    friction_combine: b.frictionCombineMode, restitution_combine: b.restitutionCombineMode,
    sleeping: b.sleeping ? 1 : 0,
    position: vec(b.position, 3), quaternion: vec(b.quaternion, 4),
    linear_velocity: vec(m.linearVelocity, 3), angular_velocity: vec(m.angularVelocity, 3),
    inv_mass: m.invMass, inv_inertia_diagonal: vec(m.invInertiaDiagonal, 3), inertia_rotation: vec(m.inertiaRotation, 4),
    friction: b.friction, restitution: b.restitution, gravity_factor: m.gravityFactor,
    linear_damping: m.linearDamping, angular_damping: m.angularDamping,
    max_linear_velocity: m.maxLinearVelocity, max_angular_velocity: m.maxAngularVelocity,
    force: vec(m.force, 3), torque: vec(m.torque, 3),
    num_velocity_steps_override: m.numVelocityStepsOverride || 0,
    num_position_steps_override: m.numPositionStepsOverride || 0,
    sensor: 0, motion_quality: m.motionQuality === 1 ? 1 : 0,
  };
}

// an inert stand-in for a removed body's pool slot: static, colliding with nothing
function placeholder(shapeIndex) {
  return { shape: shapeIndex, motion_type: MOTION_STATIC, object_layer: 0, collision_groups: 0, collision_mask: 0,
           allow_sleeping: 1, allowed_dofs: 63, friction_combine: 0, restitution_combine: 0, sleeping: 0,
           position: [0, 0, 0], quaternion: [0, 0, 0, 1], linear_velocity: [0, 0, 0], angular_velocity: [0, 0, 0],
           inv_mass: 0, inv_inertia_diagonal: [0, 0, 0], inertia_rotation: [0, 0, 0, 1], friction: 0, restitution: 0,
           gravity_factor: 0, linear_damping: 0, angular_damping: 0, max_linear_velocity: 0, max_angular_velocity: 0,
           force: [0, 0, 0], torque: [0, 0, 0], num_velocity_steps_override: 0, num_position_steps_override: 0,
           sensor: 0, motion_quality: 0 };
}

// The world description of the C ABI for a crashcat World.
function describe(world) {
  const pool = world.bodies.pool;
  const T = shapeTables(pool);
  const live = pool.filter((b) => b && !b._pooled);
  if (!live.length) fail('the world has no bodies');
  for (let k = 1; k < live.length; k++) {
    if (live[k].id < live[k - 1].id) fail('body ids are not in index order (a body was re-created in a lower slot)');
  }
  let inert = -1;
  const bodies = pool.map((b) => {
    if (b && !b._pooled) return bodyDesc(b, T.add(b.shape));
    if (inert < 0) inert = T.add({ type: 0, radius: 1e-3, aabb: [[-1e-3, -1e-3, -1e-3], [1e-3, 1e-3, 1e-3]], centerOfMass: [0, 0, 0] });
    return placeholder(inert);
  });
  T.finish();
  const pointConstraints = [];
  const pools = (world.constraints && world.constraints.pools) || {};
  for (const key of Object.keys(pools)) {
    const p = pools[key];
    const live = (p.constraints || []).filter((c) => c && !c._pooled);
    if (!live.length) continue;
    if (Number(p.type) !== POINT_CONSTRAINT) fail(`constraint type ${p.type} is not on the GPU path (point constraints only)`);
    for (const c of live) {
      pointConstraints.push({
        body_a: c.bodyIndexA, body_b: c.bodyIndexB, enabled: c.enabled ? 1 : 0,
        local_point_a: vec(c.localSpacePosition1, 3), local_point_b: vec(c.localSpacePosition2, 3),
        num_velocity_steps_override: c.numVelocityStepsOverride || 0,
        num_position_steps_override: c.numPositionStepsOverride || 0,
      });
    }
  }
  return { settings: settingsDesc(world.settings), worldBodyOffsets: [0, bodies.length], shapes: T.shapes, hull: T.hull,
           mesh: T.mesh, bodies, pointConstraints };
}

const FIELD = ['position', 'quaternion', 'center_of_mass_position', 'linear_velocity', 'angular_velocity'];

// A crashcat World bound to one GPU context.
class CrashcatGpuWorld {
  constructor(world, opts) {
    this.world = world;
    this.gpu = cc.createWorld(describe(world), opts);
    this.poolLength = world.bodies.pool.length;
    this.written = null;  // the states last written into the body objects
  }

  // bodies whose pose, velocities or sleep state the host changed since the last step go back
  // to the GPU (cc_set_bodies), as crashcat's own setters act on the next step
  pushChangedBodies() {
    if (!this.written) return;
    const pool = this.world.bodies.pool;
    let changed = false;
    const states = this.written;
    for (let i = 0; i < pool.length && !changed; i++) {
      const b = pool[i], s = states[i];
      if (!b || b._pooled) continue;
      const m = b.motionProperties;
      const now = [b.position, b.quaternion, b.centerOfMassPosition, m.linearVelocity, m.angularVelocity];
      for (let f = 0; f < FIELD.length && !changed; f++) {
        const a = s[FIELD[f]], v = now[f];
        for (let k = 0; k < a.length; k++) if (a[k] !== v[k]) { changed = true; break; }
      }
      if ((b.sleeping ? 1 : 0) !== s.sleeping) changed = true;
    }
    if (!changed) return;
    for (let i = 0; i < pool.length; i++) {
      const b = pool[i], s = states[i];
      if (!b || b._pooled) continue;
      const m = b.motionProperties;
      s.position = vec(b.position, 3);
      s.quaternion = vec(b.quaternion, 4);
      s.center_of_mass_position = vec(b.centerOfMassPosition, 3);
      s.linear_velocity = vec(m.linearVelocity, 3);
      s.angular_velocity = vec(m.angularVelocity, 3);
      s.sleeping = b.sleeping ? 1 : 0;
    }
    this.gpu.setBodies(states);
  }

  pushForces() {
    const pool = this.world.bodies.pool;
    const n = pool.length;
    let any = false;
    const f = new Float64Array(3 * n), t = new Float64Array(3 * n);
    for (let i = 0; i < n; i++) {
      const b = pool[i];
      if (!b || b._pooled) continue;
      const m = b.motionProperties;
      for (let k = 0; k < 3; k++) {
        f[3 * i + k] = m.force[k];
        t[3 * i + k] = m.torque[k];
        if (m.force[k] !== 0 || m.torque[k] !== 0) any = true;
      }
    }
    if (any) this.gpu.addForces(f, t, false);
  }

  writeBack() {
    const states = this.gpu.bodies();
    const pool = this.world.bodies.pool;
    for (let i = 0; i < pool.length; i++) {
      const b = pool[i];
      if (!b || b._pooled) continue;
      const s = states[i], m = b.motionProperties;
      for (let k = 0; k < 3; k++) {
        b.position[k] = s.position[k];
        b.centerOfMassPosition[k] = s.center_of_mass_position[k];
        m.linearVelocity[k] = s.linear_velocity[k];
        m.angularVelocity[k] = s.angular_velocity[k];
        m.force[k] = 0;
        m.torque[k] = 0;
        b.aabb[0][k] = s.aabb_min[k];
        b.aabb[1][k] = s.aabb_max[k];
      }
      for (let k = 0; k < 4; k++) b.quaternion[k] = s.quaternion[k];
      b.sleeping = s.sleeping !== 0;
    }
    this.written = states;
  }

  // world.ccd.ccdBodies as updateWorld leaves it (src/ccd.ts:19-60, src/update.ts:131-140): the
  // step's CCD records in time-of-impact order, and each CCD body's ccdBodyIndex, its place in
  // body order (before the sort).  A record without a hit keeps the reset sub-shape ids (0).
  writeCcd() {
    if (!this.world.ccd) this.world.ccd = { ccdBodies: [] };
    const pool = this.world.bodies.pool;
    for (const b of pool) if (b && !b._pooled) b.ccdBodyIndex = -1;
    const recs = this.gpu.ccdBodies();
    const byBody = recs.map((r) => r.body_index).sort((x, y) => x - y);
    byBody.forEach((bi, k) => { if (pool[bi]) pool[bi].ccdBodyIndex = k; });
    this.world.ccd.ccdBodies = recs.map((r) => ({
      bodyIndex: r.body_index, hitBodyIndex: r.hit_body_index,
      deltaPosition: r.delta_position.slice(), contactNormal: r.contact_normal.slice(),
      contactPoint: r.contact_point.slice(), fraction: r.fraction, fractionPlusSlop: r.fraction_plus_slop,
      linearCastThresholdSq: r.linear_cast_threshold_sq, maxPenetration: r.max_penetration,
      subShapeId1: r.sub_shape_id_1, subShapeId2: r.sub_shape_id_2,
      contactSettings: {
        combinedFriction: r.combined_friction, combinedRestitution: r.combined_restitution, isSensor: false,
        invMassScale1: 1, invMassScale2: 1, invInertiaScale1: 1, invInertiaScale2: 1,
      },
    }));
  }

  // The listener's contact callbacks with the RigidBody objects (src/listener.ts:23-99): the
  // manifold in crashcat's form (base offset at the origin, so the relative points are world
  // points), the contact settings with the combined friction and restitution.
  dispatch(listener) {
    const pool = this.world.bodies.pool;
    const added = listener.onContactAdded, persisted = listener.onContactPersisted;
    if (typeof added === 'function' || typeof persisted === 'function') {
      for (const c of this.gpu.contacts()) {
        const cb = c.persisted ? persisted : added;
        if (typeof cb !== 'function') continue;
        const n = c.num_points;
        const manifold = {
          numContactPoints: n, baseOffset: [0, 0, 0],
          relativeContactPointsOnA: c.position_a.slice(0, 3 * n), relativeContactPointsOnB: c.position_b.slice(0, 3 * n),
          worldSpaceNormal: c.normal.slice(), penetrationDepth: 0,
          subShapeIdA: c.sub_shape_a, subShapeIdB: c.sub_shape_b, materialIdA: -1, materialIdB: -1,
        };
        const A = pool[c.body_a], B = pool[c.body_b];
        const settings = {
          combinedFriction: combine(A.friction, B.friction, A.frictionCombineMode, B.frictionCombineMode),
          combinedRestitution: combine(A.restitution, B.restitution, A.restitutionCombineMode, B.restitutionCombineMode),
          isSensor: false, invMassScale1: 1, invMassScale2: 1, invInertiaScale1: 1, invInertiaScale2: 1,
        };
        cb(A, B, manifold, settings);
      }
    }
    if (typeof listener.onContactRemoved === 'function') {
      for (const r of this.gpu.removedContacts()) {
        listener.onContactRemoved(pool[r.body_a].id, pool[r.body_b].id, r.sub_shape_a, r.sub_shape_b);
      }
    }
  }

  step(listener, timeStep) {
    if (this.world.bodies.pool.length !== this.poolLength) {
      fail('bodies were added or removed since the world was bound to the GPU: bind it again (fromCrashcatWorld)');
    }
    this.pushChangedBodies();
    this.pushForces();
    this.gpu.step(timeStep, 1);
    this.writeBack();
    this.writeCcd();
    this.world.previousTimeStep = timeStep;
    if (listener) this.dispatch(listener);
  }
}

const bound = new WeakMap();

// Bind a crashcat World to a GPU context (opts: as createWorld's).  Refuses what the GPU path
// does not have, before any GPU work.
function fromCrashcatWorld(world, opts) {
  if (!isCrashcatWorld(world)) fail('fromCrashcatWorld needs a crashcat World (settings, bodies.pool)');
  const g = new CrashcatGpuWorld(world, opts);
  bound.set(world, g);
  return g;
}

// updateWorld(world, listener, timeStep) for a crashcat World: bound on first use
function updateCrashcatWorld(world, listener, timeStep) {
  let g = bound.get(world);
  if (!g) g = fromCrashcatWorld(world);
  g.step(listener, timeStep);
}

module.exports = { fromCrashcatWorld, updateCrashcatWorld, isCrashcatWorld, describe, shrunkPoints, CrashcatGpuWorld };
