'use strict';
// Steps a world description through the JavaScript host with updateWorld and a listener, and
// prints one JSON line per step with the contact events the listener received:
//   node events_world.js <desc.json> <fp32|fp64> <steps>
// {"added": [[a, b, subA, subB], ...], "persisted": [...], "removed": [...]}
const fs = require('fs');
const path = require('path');
const cc = require(path.join(__dirname, '..', '..', 'crashcat_amd', 'node', 'index.js'));

const args = process.argv.slice(2);
const desc = JSON.parse(fs.readFileSync(args[0], 'utf8'));
const world = cc.createWorld(desc, { precision: args[1], maxPairsPerBody: 16, maxContactsPerBody: 8 });
const steps = Number(args[2]);
let ev;
const listener = {
  onContactAdded(a, b, m) { ev.added.push([a, b, m.subShapeIdA, m.subShapeIdB, m.numContactPoints]); },
  onContactPersisted(a, b, m) { ev.persisted.push([a, b, m.subShapeIdA, m.subShapeIdB, m.numContactPoints]); },
  onContactRemoved(a, b, sa, sb) { ev.removed.push([a, b, sa, sb]); },
};
for (let i = 0; i < steps; i++) {
  ev = { added: [], persisted: [], removed: [] };
  cc.updateWorld(world, listener, 1 / 60);
  console.log(JSON.stringify(ev));
}
