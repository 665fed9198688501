'use strict';
// Steps a crashcat-shaped World (the JSON of Scene.crashcat_world(), shapes put back in place as
// shared objects) with crashcat_amd's updateWorld and prints, per step, the listener's events
// (body indices from the RigidBody objects it receives), then the bodies' final state read from
// the body objects.  Arguments: world.json precision steps [forceStep forceBody] [order]
const fs = require('fs');
const path = require('path');
const cc = require(path.join(__dirname, '..', '..', 'crashcat_amd', 'node', 'index.js'));

const [file, precision, stepsArg, forceStepArg, forceBodyArg, order] = process.argv.slice(2);
const d = JSON.parse(fs.readFileSync(file, 'utf8'));
const shapes = d.shapes;
for (const b of d.bodies.pool) b.shape = shapes[b.shape];
const world = { settings: d.settings, bodies: d.bodies, constraints: d.constraints, previousTimeStep: 0 };
cc.fromCrashcatWorld(world, { precision, maxPairsPerBody: 24, solverOrder: order || 'reference' });
const steps = Number(stepsArg), forceStep = Number(forceStepArg || -1), forceBody = Number(forceBodyArg || -1);
const idx = (id) => id % 4294967296;
for (let s = 0; s < steps; s++) {
  const ev = { added: [], persisted: [], removed: [] };
  if (s === forceStep) {
    const m = world.bodies.pool[forceBody].motionProperties;
    m.force[0] += 400; m.force[1] += 900; m.torque[2] += 30;
  }
  const listener = {
    onContactAdded(a, b, m) { ev.added.push([a.index, b.index, m.subShapeIdA, m.subShapeIdB, m.numContactPoints]); },
    onContactPersisted(a, b, m) { ev.persisted.push([a.index, b.index, m.subShapeIdA, m.subShapeIdB, m.numContactPoints]); },
    onContactRemoved(ia, ib, sa, sb) { ev.removed.push([idx(ia), idx(ib), sa, sb]); },
  };
  cc.updateWorld(world, listener, 1 / 60);
  if (s === forceStep) {
    const m = world.bodies.pool[forceBody].motionProperties;
    ev.forceCleared = m.force.every((x) => x === 0) && m.torque.every((x) => x === 0);
  }
  ev.ccd = world.ccd.ccdBodies.map((c) => [c.bodyIndex, c.hitBodyIndex, c.fraction, c.fractionPlusSlop,
                                           c.contactPoint[0], c.contactPoint[1], c.contactPoint[2]]);
  console.log(JSON.stringify(ev));
}
const out = world.bodies.pool.map((b) => ({
  position: b.position, quaternion: b.quaternion, com: b.centerOfMassPosition,
  v: b.motionProperties.linearVelocity, w: b.motionProperties.angularVelocity, sleeping: b.sleeping,
}));
console.log(JSON.stringify({ bodies: out, previousTimeStep: world.previousTimeStep }));
