'use strict'
// Node-side boundary checks that need no GPU: the module and addon load,
// every reference entry point is exported, host-side conversions round like
// the reference's Float32Array.set / DataView.setFloat32, and a missing
// device surfaces as an Error with the lbvh_status code.
const assert = require('assert')
const path = require('path')
const bvh = require(path.join(__dirname, '..', '..', 'webgpu-lvbh-demo_amd', 'bvh'))

for (const name of ['initBVHBuild', 'initAABB_ZidxKernel', 'initRadixSortKernel',
  'initRadixTreeKernel', 'initBVHUpPassKernel', 'initRearrangeKernel', 'createDevice',
  'initBVHBatchBuild', 'raycast', 'createServer']) {
  assert.strictEqual(typeof bvh[name], 'function', name)
}
for (const name of ['buildBatched', 'autotune', 'raycast', 'createServer', 'serverSubmit',
  'serverPoll', 'serverDrain', 'serverMesh', 'serverEvict', 'serverStats',
  'serverSetDigest', 'serverCancel', 'serverSetMaxWait']) {
  assert.strictEqual(typeof bvh.addon[name], 'function', 'addon.' + name)
}

const g = bvh.installGlobals({})
assert.strictEqual(g.initBVHBuild, bvh.initBVHBuild)

// f64 -> f32 rounding of TRI_ARRAY and MODEL_BOUNDS
const arr = [0.1, 1 / 3, 3.1415, 1e-8, -2.5e38, 7, 8, 9, 10, 11, 12, 13]
const f = bvh._internal.toFloat32(arr, 1)
const want = new Float32Array(12)
want.set(arr)
assert.deepStrictEqual(Array.from(f), Array.from(want))
assert.strictEqual(f[0], Math.fround(0.1))
const b = bvh._internal.toBounds({ min: [0.1, -1 / 3, 2], max: [1.1, 5.5, 1e10] })
assert.strictEqual(b[1], Math.fround(-1 / 3))
assert.strictEqual(b[5], Math.fround(1e10))
assert.strictEqual(bvh._internal.toBounds(null), null)
const typed = new Float32Array(24)
assert.strictEqual(bvh._internal.toFloat32(typed, 2), typed)

let threw = false
try {
  bvh.createDevice({ ordinal: 4096 })
} catch (e) {
  threw = true
  assert.strictEqual(e.code, 6)
}
assert.ok(threw, 'createDevice on a missing device must throw')
threw = false
try {
  bvh.createServer({ ordinal: 4096 })
} catch (e) {
  threw = true
  assert.strictEqual(e.code, 6)
}
assert.ok(threw, 'createServer on a missing device must throw')
// Uploads follow mappedAtCreation + Float32Array.set (build-bvh.js:14-21):
// data longer than the buffer throws RangeError before any allocation, a
// shorter one leaves the rest zero.  None of these reach the device.
const ue = bvh._internal.uploadExtent
assert.deepStrictEqual(ue(48, new Float32Array(12)), { bytes: 48, zeros: 0 })
assert.deepStrictEqual(ue(48, new Float32Array(6)), { bytes: 24, zeros: 24 })
assert.deepStrictEqual(ue(48, new ArrayBuffer(10)), { bytes: 10, zeros: 38 })
assert.deepStrictEqual(ue(48, undefined), { bytes: 0, zeros: 0 })
assert.deepStrictEqual(ue(0, new Float32Array(0)), { bytes: 0, zeros: 0 })
assert.throws(() => ue(48, new Float32Array(13)), RangeError)
assert.throws(() => ue(0, new Uint8Array(1)), RangeError)
assert.throws(() => ue(48, [1, 2, 3]), TypeError)

// A device whose allocator must never be reached: every call below is
// refused on the host first.
const noDevice = Object.create(bvh.LBVHDevice.prototype)
noDevice.ctx = null
noDevice.queue = { onSubmittedWorkDone: () => Promise.resolve() }
let warned = []
const origWarn = console.warn
console.warn = (msg) => { warned.push(String(msg)) }

async function hostChecks () {
  // build(): a Float32Array of more than 12 * N floats
  await assert.rejects(bvh.initBVHBuild(noDevice).build(new Float32Array(13), 1, null), RangeError)
  // and a JS Array of more than 12 * N numbers (Float32Array.set itself)
  await assert.rejects(bvh.initBVHBuild(noDevice).build(new Array(25).fill(0), 2, null), RangeError)
  // batched: a mesh array longer than its NUM_TRIS, or not 12 floats per triangle
  await assert.rejects(bvh.initBVHBatchBuild(noDevice).build(
    [{ TRI_ARRAY: new Float32Array(24), NUM_TRIS: 1 }], null), RangeError)
  await assert.rejects(bvh.initBVHBatchBuild(noDevice).build([new Float32Array(13)], null),
    RangeError)
  // up pass: each input must hold `size` entries (warning + undefined, as
  // the other shims report a mismatch)
  const up = bvh.initBVHUpPassKernel(noDevice)
  const buf = (size) => ({ size, ptr: 1n })
  for (const [idx, aabb, par] of [[4, 64, 16], [8, 32, 16], [8, 64, 8]]) {
    warned = []
    const r = await up.execute(buf(idx), buf(aabb), buf(par), 2)
    assert.strictEqual(r, undefined)
    assert.strictEqual(warned.length, 1)
    assert.ok(/too small for requested size \[ 2 \]/.test(warned[0]), warned[0])
  }
  // ray caster: BVH / triangle buffers too small for NUM_TRIS, ragged rays
  await assert.rejects(bvh.raycast(noDevice, buf(64), buf(96), 2, new Float32Array(8)), RangeError)
  await assert.rejects(bvh.raycast(noDevice, buf(128), buf(48), 2, new Float32Array(8)), RangeError)
  await assert.rejects(bvh.raycast(noDevice, buf(128), buf(96), 2, new Float32Array(9)), RangeError)
}

hostChecks().then(() => {
  console.warn = origWarn
  console.log('boundary_cpu ok')
}, (e) => {
  console.warn = origWarn
  console.error(e)
  process.exit(1)
})
