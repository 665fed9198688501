'use strict';
// Steps a world description through the JavaScript host (crashcat_amd/node/index.js) with
// updateWorld and writes the body states (cc_body_state layout) and the last step's pairs:
//   node step_world.js <desc.json> <fp32|fp64> <steps> <bodies.bin> <pairs.bin>
// Prints the last step's counters as JSON.
const fs = require('fs');
const path = require('path');
const cc = require(path.join(__dirname, '..', '..', 'crashcat_amd', 'node', 'index.js'));

const args = process.argv.slice(2);
const desc = JSON.parse(fs.readFileSync(args[0], 'utf8'));
const world = cc.createWorld(desc, { precision: args[1], maxPairsPerBody: 16, maxContactsPerBody: 8 });
const steps = Number(args[2]);
for (let i = 0; i < steps; i++) cc.updateWorld(world, undefined, 1 / 60);
world.sync();
fs.writeFileSync(args[3], Buffer.from(world.bodyStateBuffer()));
const pairs = world.pairs();
fs.writeFileSync(args[4], Buffer.from(pairs.buffer, pairs.byteOffset, pairs.byteLength));
console.log(JSON.stringify(world.counters()));
