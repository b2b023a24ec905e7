'use strict'
// parseObj (bvh/obj.js) against hand-computed expectations of
// parse-obj.js:3-87: y/z swap, fan triangulation, negative indices,
// first-model-only vertices, f64 bounds that exclude the floor.
const assert = require('assert')
const path = require('path')
const { parseObj } = require(path.join(__dirname, '..', '..', 'webgpu-lvbh-demo_amd', 'bvh', 'obj.js'))

const P = 3.1415
// a quad (fan -> 2 triangles) and a triangle using negative indices
const text = [
  '# comment',
  'v 0 0 0', 'v 1 0 0', 'v 1 2 0', 'v 0 2 0',
  'v 0.5 1 3',
  'vt 0 0', 'vn 0 0 1',
  'f 1/1/1 2/1/1 3/1/1 4/1/1',
  'f -1 -4 -3',
  '   f   1   2    5  ',
  ''
].join('\n')
const r = parseObj(text)
assert.strictEqual(r.NUM_TRIS, 4 + 2)
const t = r.TRI_ARR
// triangle 0: fan (1,2,3) with y<->z: (x, z, y)
assert.deepStrictEqual(t.slice(0, 12), [0, 0, 0, P, 1, 0, 0, P, 1, 0, 2, P])
assert.deepStrictEqual(t.slice(12, 24), [0, 0, 0, P, 1, 0, 2, P, 0, 0, 2, P])
// negative indices: -1 -> 5, -4 -> 2, -3 -> 3
assert.deepStrictEqual(t.slice(24, 36), [0.5, 3, 1, P, 1, 0, 0, P, 1, 0, 2, P])
assert.deepStrictEqual(t.slice(36, 48), [0, 0, 0, P, 1, 0, 0, P, 0.5, 3, 1, P])
// bounds over face vertices (swapped), floor excluded
assert.deepStrictEqual(r.BOUNDS, { min: [0, 0, 0], max: [1, 3, 2] })
// floor: two 20,000-unit triangles at z_min + 0.01
assert.deepStrictEqual(t.slice(48, 60), [-1e4, -1e4, 0.01, P, 1e4, -1e4, 0.01, P, -1e4, 1e4, 0.01, P])
assert.deepStrictEqual(t.slice(60, 72), [1e4, -1e4, 0.01, P, 1e4, 1e4, 0.01, P, -1e4, 1e4, 0.01, P])

// vertices after a second 'o' are not part of model 0
const two = parseObj('v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\no second\nv 9 9 9\nf 1 2 3\n')
assert.strictEqual(two.NUM_TRIS, 3)
// a leading 'o' names the first model; it does not start a second one
const named = parseObj('o first\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n')
assert.strictEqual(named.NUM_TRIS, 3)

assert.throws(() => parseObj('v 0 0 0\nf 0 1 1\n'), /invalid vertex index of 0/)
console.log('obj_cpu ok')
