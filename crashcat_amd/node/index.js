'use strict';
// JavaScript host of the GPU step.  A world description (the fields of the C ABI's structs,
// by name) is packed into the structs whose layouts the addon reads from the C header, uploaded
// once, and stepped by updateWorld, which takes the place of crashcat's
// updateWorld(world, listener, timeStep) (src/update.ts:45) for the worlds this path supports.
// Node 12 syntax (no ?. or ??).
const path = require('path');

const addon = require(path.join(__dirname, 'crashcat_amd.node'));
const LAYOUT = addon.layout();
const PRECISION = { fp32: 0, fp64: 1 };

function writeField(dv, base, spec, value) {
  const off = base + spec[0];
  const type = spec[1];
  const count = spec[2];
  for (let i = 0; i < count; i++) {
    let v = count === 1 ? value : value[i];
    if (v === undefined || v === null) v = 0;
    switch (type) {
      case 'f64': dv.setFloat64(off + 8 * i, v, true); break;
      case 'i32': dv.setInt32(off + 4 * i, v, true); break;
      case 'u32': dv.setUint32(off + 4 * i, v >>> 0, true); break;
      case 'u8': dv.setUint8(off + i, v); break;
      case 'i64': dv.setBigInt64(off + 8 * i, BigInt(v), true); break;
      default: throw new Error('crashcat_amd: unknown field type ' + type);
    }
  }
}

function readField(dv, base, spec) {
  const off = base + spec[0];
  const type = spec[1];
  const count = spec[2];
  const out = [];
  for (let i = 0; i < count; i++) {
    switch (type) {
      case 'f64': out.push(dv.getFloat64(off + 8 * i, true)); break;
      case 'i32': out.push(dv.getInt32(off + 4 * i, true)); break;
      case 'u32': out.push(dv.getUint32(off + 4 * i, true)); break;
      case 'u8': out.push(dv.getUint8(off + i)); break;
      case 'i64': out.push(Number(dv.getBigInt64(off + 8 * i, true))); break;
      default: throw new Error('crashcat_amd: unknown field type ' + type);
    }
  }
  return count === 1 ? out[0] : out;
}

// records (objects keyed by the C field names) -> ArrayBuffer of `structName`; a field a
// record does not name is zero
function pack(structName, records) {
  const st = LAYOUT[structName];
  if (!st) throw new Error('crashcat_amd: no struct ' + structName);
  const buf = new ArrayBuffer(st.size * records.length);
  const dv = new DataView(buf);
  const names = Object.keys(st.fields);
  for (let k = 0; k < records.length; k++) {
    const r = records[k];
    for (const name of Object.keys(r)) {
      if (names.indexOf(name) < 0) throw new Error('crashcat_amd: ' + structName + ' has no field ' + name);
    }
    for (const name of names) {
      if (r[name] !== undefined) writeField(dv, k * st.size, st.fields[name], r[name]);
    }
  }
  return buf;
}

function unpack(structName, buf) {
  const st = LAYOUT[structName];
  const dv = new DataView(buf);
  const n = buf.byteLength / st.size;
  const out = [];
  for (let k = 0; k < n; k++) {
    const r = {};
    for (const name of Object.keys(st.fields)) r[name] = readField(dv, k * st.size, st.fields[name]);
    out.push(r);
  }
  return out;
}

// Listener callbacks (src/listener.ts:23-99).  The validate callbacks decide what the step
// does and cannot run inside a GPU step; the contact callbacks are delivered after the step
// from the event records the step leaves (cc_download_contacts, cc_download_removed_contacts).
const STEP_CALLBACKS = ['onBodyPairValidate', 'onContactValidate'];
const EVENT_CALLBACKS = ['onContactAdded', 'onContactPersisted', 'onContactRemoved'];

function callbackNames(listener) {
  if (listener === undefined || listener === null) return [];
  return Object.keys(listener).filter((k) => typeof listener[k] === 'function');
}

// The manifold handed to onContactAdded / onContactPersisted: world-space normal (A -> B),
// contact points on A and B in world space, sub-shape ids; bodies ordered id A < id B.
function eventManifold(c) {
  const n = c.num_points;
  return {
    worldSpaceNormal: c.normal.slice(),
    numContactPoints: n,
    contactPointsOnA: c.position_a.slice(0, 3 * n),
    contactPointsOnB: c.position_b.slice(0, 3 * n),
    subShapeIdA: c.sub_shape_a,
    subShapeIdB: c.sub_shape_b,
  };
}

// One GPU context holding a batch of independent worlds.
class GpuWorld {
  // desc: { settings, worldBodyOffsets, shapes, hull, mesh, bodies, pointConstraints }
  // mesh (optional): { positions, triangles, nodes } as a TriangleMeshShape holds them after
  // buildTriangleMesh: data.positions, the triangle buffer as 4 ints per triangle (a, b, c,
  // active-edge bits) and bvh.buffer, concatenated over the scene's meshes
  // opts: { precision: 'fp32' | 'fp64', maxPairsPerBody, maxContactsPerBody, device,
  //         solverOrder: 'reference' (crashcat's order) | 'colored' (graph-coloured) }
  constructor(desc, opts) {
    const o = opts || {};
    const prec = PRECISION[o.precision || 'fp32'];
    if (prec === undefined) throw new Error('crashcat_amd: precision must be fp32 or fp64');
    const order = { reference: 0, colored: 1 }[o.solverOrder || 'reference'];
    if (order === undefined) throw new Error('crashcat_amd: solverOrder must be reference or colored');
    this.handle = addon.create(prec, o.maxPairsPerBody || 24, o.maxContactsPerBody || 8, o.device || 0, order);
    const hull = desc.hull || {};
    const hullBuffers = {
      points: new Float64Array(hull.points || []),
      shrunkPoints: new Float64Array(hull.shrunkPoints || []),
      faceFirstVertex: new Int32Array(hull.faceFirstVertex || []),
      faceNumVertices: new Int32Array(hull.faceNumVertices || []),
      faceVertexIndices: new Int32Array(hull.faceVertexIndices || []),
      planeNormals: new Float64Array(hull.planeNormals || []),
      planeConstants: new Float64Array(hull.planeConstants || []),
      totalPoints: (hull.points || []).length / 3,
      totalFaces: (hull.faceNumVertices || []).length,
      totalFaceVertices: (hull.faceVertexIndices || []).length,
    };
    if (desc.mesh && desc.mesh.triangles && desc.mesh.triangles.length) {
      hullBuffers.meshPositions = new Float64Array(desc.mesh.positions);
      hullBuffers.meshTriangles = new Int32Array(desc.mesh.triangles);
      hullBuffers.meshNodes = new Float64Array(desc.mesh.nodes);
    }
    const bodies = desc.bodies || [];
    const offsets = desc.worldBodyOffsets || [0, bodies.length];
    addon.upload(this.handle, pack('cc_settings', [desc.settings]), new Int32Array(offsets),
                 pack('cc_body_desc', bodies), pack('cc_shape_desc', desc.shapes || []), hullBuffers,
                 pack('cc_point_constraint_desc', desc.pointConstraints || []));
    this.numBodies = bodies.length;
    this.previousTimeStep = 0;
  }

  step(timeStep, n) {
    addon.step(this.handle, timeStep, n === undefined ? 1 : n);
    this.previousTimeStep = timeStep;
  }

  sync() { addon.sync(this.handle); }

  // body states in the cc_body_state layout (ArrayBuffer), or as objects
  bodyStateBuffer() {
    const buf = new ArrayBuffer(LAYOUT.cc_body_state.size * this.numBodies);
    addon.downloadBodies(this.handle, buf);
    return buf;
  }
  bodies() { return unpack('cc_body_state', this.bodyStateBuffer()); }
  setBodies(states) {
    const buf = states instanceof ArrayBuffer ? states : pack('cc_body_state', states);
    addon.setBodies(this.handle, buf);
  }
  pairs() { return addon.downloadPairs(this.handle); }
  // contacts of the last step (cc_contact_out records, `persisted` set for those the step
  // before also had) and the contacts the last step removed (cc_contact_removed records)
  contacts() { return unpack('cc_contact_out', addon.downloadContacts(this.handle)); }
  removedContacts() { return unpack('cc_contact_removed', addon.downloadRemovedContacts(this.handle)); }
  // the last step's CCD records in time-of-impact order (world.ccd.ccdBodies after updateWorld)
  ccdBodies() { return unpack('cc_ccd_body', addon.downloadCcdBodies(this.handle)); }

  // the state one step hands to the next as an ArrayBuffer (cc_snapshot), and back into a
  // world made from the same description (cc_restore): the steps after it repeat bit for bit
  snapshot() { return addon.snapshot(this.handle); }
  restore(buf) {
    addon.restore(this.handle, buf);
    const dv = buf instanceof ArrayBuffer ? new DataView(buf) : new DataView(buf.buffer, buf.byteOffset, buf.byteLength);
    this.previousTimeStep = dv.getFloat64(40, true);  // the blob header's previous time step
  }

  // Deliver the last step's contact events to a listener: onContactAdded / onContactPersisted
  // (src/constraints/contact-constraints.ts:493-499) per contact in (body a, body b, sub-shape)
  // order, then onContactRemoved (src/contacts.ts:305-307) per removed contact.  Bodies are
  // passed as their indices; the step has already run, so the settings object is informative.
  dispatchContactEvents(listener) {
    const added = listener.onContactAdded, persisted = listener.onContactPersisted;
    if (typeof added === 'function' || typeof persisted === 'function') {
      for (const c of this.contacts()) {
        const cb = c.persisted ? persisted : added;
        if (typeof cb === 'function') cb(c.body_a, c.body_b, eventManifold(c), { isSensor: false });
      }
    }
    if (typeof listener.onContactRemoved === 'function') {
      for (const r of this.removedContacts()) listener.onContactRemoved(r.body_a, r.body_b, r.sub_shape_a, r.sub_shape_b);
    }
  }
  counters() {
    const buf = new ArrayBuffer(LAYOUT.cc_step_counters.size);
    addon.counters(this.handle, buf);
    return unpack('cc_step_counters', buf)[0];
  }

  // addForce / addTorque (src/body/rigid-body.ts:840-869) on every body: Float64Arrays of 3 per
  // body (either may be null); they act on the next step, which then clears them
  addForces(force, torque, wake) {
    addon.addForces(this.handle, force || null, torque || null, wake === undefined ? 1 : (wake ? 1 : 0));
  }

  // differentiable stepping (cc_fwd / cc_bwd): nSteps steps with a checkpoint every
  // checkpointEvery, then the adjoint given dL/d(final centres of mass) (Float64Array, 3 per
  // body) and optionally dL/d(final linear velocities)
  fwd(timeStep, nSteps, checkpointEvery) {
    addon.fwd(this.handle, timeStep, nSteps, checkpointEvery || 10);
    this.previousTimeStep = timeStep;
  }
  bwd(dLdCom, dLdVel) {
    const n = this.numBodies;
    const all = new Float64Array(addon.bwd(this.handle, dLdCom, dLdVel || null));
    return {
      linearVelocity: all.subarray(0, 3 * n),
      angularVelocity: all.subarray(3 * n, 6 * n),
      centerOfMass: all.subarray(6 * n, 9 * n),
      friction: all.subarray(9 * n, 10 * n),
    };
  }
}

// A run of steps as an autograd Function: forward(ctx, world, timeStep, nSteps, checkpointEvery)
// steps the world from its uploaded state and returns the final centres of mass (Float64Array,
// 3 per body); backward(ctx, dLdCom) returns the gradients with respect to the initial linear
// and angular velocities, centres of mass and each body's friction coefficient.
class WorldRun {
  static forward(ctx, world, timeStep, nSteps, checkpointEvery) {
    world.fwd(timeStep, nSteps, checkpointEvery);
    ctx.world = world;
    const com = new Float64Array(3 * world.numBodies);
    world.bodies().forEach((b, i) => com.set(b.center_of_mass_position, 3 * i));
    return com;
  }
  static backward(ctx, dLdCom) { return ctx.world.bwd(dLdCom); }
}

function createWorld(desc, opts) { return new GpuWorld(desc, opts); }

// the listener's callbacks, refusing those that would have to run inside the GPU step
function checkListener(listener) {
  const names = callbackNames(listener);
  const refused = names.filter((k) => STEP_CALLBACKS.includes(k) || !EVENT_CALLBACKS.includes(k));
  if (refused.length) throw new Error(`crashcat_amd: listener callbacks not available on the GPU step: ${refused.join(', ')}`);
  return names;
}

// bind a crashcat World to the GPU (crashcat_world.js); updateWorld binds on first use
function fromCrashcatWorld(world, opts) { return require('./crashcat_world.js').fromCrashcatWorld(world, opts); }

// updateWorld(world, listener, timeStep) (src/update.ts:45): one step of every world of the
// context.  The contact callbacks of the listener are called after the step from its event
// records; a validate callback, which would have to run inside the GPU step, is refused.
function updateWorld(world, listener, timeStep) {
  if (!(world instanceof GpuWorld)) {
    const cw = require('./crashcat_world.js');
    if (!cw.isCrashcatWorld(world)) throw new Error('crashcat_amd: updateWorld needs a world made by createWorld or a crashcat World');
    checkListener(listener);
    cw.updateCrashcatWorld(world, listener, timeStep);
    return;
  }
  const names = checkListener(listener);
  world.step(timeStep, 1);
  if (names.length) world.dispatchContactEvents(listener);
}

module.exports = { createWorld, updateWorld, fromCrashcatWorld, GpuWorld, WorldRun, pack, unpack, layout: LAYOUT, addon };
