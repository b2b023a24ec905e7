'use strict'
// bvh/index.js -- the reference's scripts/bvh interface on MI355X.
//
// Drop-in for AddisonPrairie/WebGPU-LVBH-demo scripts/bvh:
//   initBVHBuild(device)                 build-bvh.js:1-10
//     .build(TRI_ARRAY, NUM_TRIS, MODEL_BOUNDS)
//         -> Promise<{ BVH_BUFFER, O_TRIANGLE_BUFFER }>  build-bvh.js:12-66
//   initAABB_ZidxKernel(device).execute(TRI, N, bounds)  aabb-z-idx.js:57-148
//   initRadixSortKernel(device).execute(KEYS, N)         radix-sort.js:161-500
//   initRadixTreeKernel(device).execute(SORTED_KEYS, N)  radix-tree.js:50-120
//   initBVHUpPassKernel(device).execute(IDX, AABB, PARENT, N)
//                                                        bvh-up-pass.js:70-156
//   initRearrangeKernel(device).execute(I_TRI, IDX, N)   rearrange.js:57-139
//
// `device` comes from createDevice() instead of navigator.gpu; buffers are
// LBVHBuffer handles with the reference's byte sizes, .size and .destroy(),
// plus .read() (the reference never reads back).  All GPU work goes through
// the N-API addon into liblbvh.so (HIP kernels); nothing runs on the CPU.
// Written for Node 12: no optional chaining, no top-level await.

const path = require('path')

const addon = require(path.join(__dirname, '..', 'lib', 'lbvh.node'))

const DIV_MODES = { mul_rcp: 0, ieee: 1 }

class LBVHBuffer {
  constructor (device, size, ptr) {
    this.device = device
    this.size = size
    this.ptr = ptr
  }

  destroy () {
    if (this.ptr !== null) {
      addon.free(this.device.ctx, this.ptr)
      this.ptr = null
    }
  }

  // Device -> host copy after all queued work; resolves to an ArrayBuffer.
  async read () {
    await this.device.queue.onSubmittedWorkDone()
    return addon.download(this.device.ctx, this.ptr, this.size)
  }
}

class LBVHDevice {
  constructor (options) {
    const o = options || {}
    this.ctx = addon.createContext(o.ordinal || 0, o.hbmBudget || 0)
    this.setDivMode(o.divMode || 'mul_rcp')
    const ctx = this.ctx
    this.queue = {
      // The context's stream drains off the JS thread.
      onSubmittedWorkDone: function () { return addon.synchronize(ctx) }
    }
  }

  setDivMode (mode) {
    if (!(mode in DIV_MODES)) throw new TypeError(`unknown division mode ${mode}`)
    addon.setDivMode(this.ctx, DIV_MODES[mode])
    this.divMode = mode
  }

  // GPUDevice.createBuffer({ size }) (build-bvh.js:14-18); `data` optionally
  // fills it like the mappedAtCreation upload (build-bvh.js:20-21): data
  // longer than the buffer throws RangeError before anything is allocated,
  // as Float32Array.set does, and a shorter one leaves the rest zero, as a
  // new mapped buffer is.
  createBuffer (descriptor, data) {
    const size = descriptor.size
    const fill = uploadExtent(size, data)
    const buf = new LBVHBuffer(this, size, addon.alloc(this.ctx, size))
    try {
      if (fill.bytes > 0) addon.upload(this.ctx, buf.ptr, data)
      if (fill.zeros > 0) {
        addon.upload(this.ctx, buf.ptr + BigInt(fill.bytes), new Uint8Array(fill.zeros))
      }
    } catch (e) {
      buf.destroy()
      throw e
    }
    return buf
  }

  bytesInUse () { return addon.bytesInUse(this.ctx) }

  // Times every (tree tile, sort digit width) variant per power-of-two size
  // bucket in [nMin, nMax] on device-synthesised meshes and keeps the
  // fastest for later builds.
  // -> { bucketSize: { '256/8': ms, '256/10': ms, ..., '1024/10': ms } }
  autotune (nMin, nMax, reps) {
    const ms = addon.autotune(this.ctx, nMin, nMax, reps || 3)
    const names = []
    for (const tile of [256, 512, 1024]) for (const bits of [8, 10]) names.push(tile + '/' + bits)
    const out = {}
    let bucket = 1
    while (2 * bucket <= nMin) bucket *= 2
    for (let i = 0; i < ms.length / names.length; i++, bucket *= 2) {
      out[bucket] = {}
      names.forEach((name, v) => { out[bucket][name] = ms[names.length * i + v] })
    }
    return out
  }
}

function createDevice (options) {
  return new LBVHDevice(options)
}

// Bytes of `data` (a typed array or ArrayBuffer) copied into a new buffer of
// `size` bytes and the zero bytes after them: { bytes, zeros }.  Throws
// RangeError, as Float32Array.set does (build-bvh.js:20), when data does
// not fit.  No data: nothing copied, nothing zeroed.
function uploadExtent (size, data) {
  if (data === undefined || data === null) return { bytes: 0, zeros: 0 }
  const bytes = data.byteLength
  if (typeof bytes !== 'number') {
    throw new TypeError('buffer data must be a typed array or an ArrayBuffer')
  }
  if (bytes > size) {
    throw new RangeError(`Source is too large: ${bytes} bytes for a buffer of ${size}`)
  }
  return { bytes: bytes, zeros: size - bytes }
}

// TRI_ARRAY (JS Array of 12*N numbers, parse-obj.js:55-59) -> Float32Array,
// rounding each value like Float32Array.set (build-bvh.js:20).  Typed arrays
// and ArrayBuffers are used as they are.
function toFloat32 (TRI_ARRAY, NUM_TRIS) {
  if (TRI_ARRAY instanceof Float32Array) return TRI_ARRAY
  if (TRI_ARRAY instanceof ArrayBuffer) return new Float32Array(TRI_ARRAY)
  const out = new Float32Array(12 * NUM_TRIS)
  out.set(TRI_ARRAY)
  return out
}

// MODEL_BOUNDS {min:[x,y,z], max:[x,y,z]} -> Float32Array(6), rounded like
// DataView.setFloat32 (aabb-z-idx.js:116-122); null -> GPU reduction.
function toBounds (MODEL_BOUNDS) {
  if (MODEL_BOUNDS === undefined || MODEL_BOUNDS === null) return null
  const b = new Float32Array(6)
  for (let c = 0; c < 3; c++) {
    b[c] = MODEL_BOUNDS.min[c]
    b[3 + c] = MODEL_BOUNDS.max[c]
  }
  return b
}

function initBVHBuild (device) {
  return { build }

  // build-bvh.js:12-66: upload, one fused device pipeline, free the input.
  async function build (TRI_ARRAY, NUM_TRIS, MODEL_BOUNDS) {
    const tris = toFloat32(TRI_ARRAY, NUM_TRIS)
    const I_TRIANGLE_BUFFER = device.createBuffer({ size: NUM_TRIS * 48 }, tris)
    const BVH_BUFFER = device.createBuffer({ size: NUM_TRIS * 64 })
    const O_TRIANGLE_BUFFER = device.createBuffer({ size: NUM_TRIS * 48 })
    if (NUM_TRIS > 0) {
      addon.build(device.ctx, I_TRIANGLE_BUFFER.ptr, NUM_TRIS,
        toBounds(MODEL_BOUNDS), BVH_BUFFER.ptr, O_TRIANGLE_BUFFER.ptr, null, null)
    }
    await device.queue.onSubmittedWorkDone()
    I_TRIANGLE_BUFFER.destroy()
    return { BVH_BUFFER, O_TRIANGLE_BUFFER }
  }
}

// ---- batched builds (no counterpart in the reference) ------------------------

// initBVHBatchBuild(device).build(MESHES, BOUNDS) builds many independent
// meshes in one graph-captured device pipeline.  MESHES: array of
// { TRI_ARRAY, NUM_TRIS } (or Float32Arrays of 12 * n floats); BOUNDS: null
// (per-mesh GPU bounds) or an array of MODEL_BOUNDS, one per mesh.  Mesh k's
// nodes are BVH_BUFFER bytes [64 * OFFSETS[k], 64 * OFFSETS[k + 1]) and its
// triangles O_TRIANGLE_BUFFER bytes [48 * OFFSETS[k], ...), each exactly the
// single-mesh build's output.
function initBVHBatchBuild (device) {
  return { build }

  async function build (MESHES, BOUNDS) {
    const S = MESHES.length
    const OFFSETS = new Uint32Array(S + 1)
    // Each mesh fills 12 * n floats of its range like Float32Array.set: a
    // longer array throws RangeError, a shorter one leaves zeros.
    const parts = MESHES.map(function (m) {
      if (m instanceof Float32Array) {
        if (m.length % 12 !== 0) throw new RangeError('a mesh needs 12 floats per triangle')
        return { tri: m, n: m.length / 12 }
      }
      return { tri: toFloat32(m.TRI_ARRAY, m.NUM_TRIS), n: m.NUM_TRIS }
    })
    for (let k = 0; k < S; k++) {
      if (parts[k].tri.length > 12 * parts[k].n) {
        throw new RangeError(`Source is too large: mesh ${k} has ${parts[k].tri.length} floats for ${parts[k].n} triangles`)
      }
      OFFSETS[k + 1] = OFFSETS[k] + parts[k].n
    }
    const total = OFFSETS[S]
    const tris = new Float32Array(12 * total)
    for (let k = 0; k < S; k++) tris.set(parts[k].tri, 12 * OFFSETS[k])
    let bounds = null
    if (BOUNDS !== undefined && BOUNDS !== null) {
      bounds = new Float32Array(6 * S)
      for (let k = 0; k < S; k++) bounds.set(toBounds(BOUNDS[k]), 6 * k)
    }
    const I_TRIANGLE_BUFFER = device.createBuffer({ size: total * 48 }, tris)
    const BVH_BUFFER = device.createBuffer({ size: total * 64 })
    const O_TRIANGLE_BUFFER = device.createBuffer({ size: total * 48 })
    if (total > 0) {
      addon.buildBatched(device.ctx, I_TRIANGLE_BUFFER.ptr, OFFSETS, bounds,
        BVH_BUFFER.ptr, O_TRIANGLE_BUFFER.ptr, null, null)
    }
    await device.queue.onSubmittedWorkDone()
    I_TRIANGLE_BUFFER.destroy()
    return { BVH_BUFFER, O_TRIANGLE_BUFFER, OFFSETS }
  }
}

// ---- consumer check: the path tracer's traversal (pathtracer.js:495-628) ----

// Closest hits of RAYS (Float32Array, 8 floats per ray: origin.xyz, pad,
// dir.xyz, pad) through a built BVH; resolves to a Float32Array of 4 floats
// per ray (normal.xyz facing the ray, distance; -1 for a miss).
async function raycast (device, BVH_BUFFER, O_TRIANGLE_BUFFER, NUM_TRIS, RAYS) {
  if (BVH_BUFFER.size < 64 * NUM_TRIS || O_TRIANGLE_BUFFER.size < 48 * NUM_TRIS) {
    throw new RangeError(`raycast: BVH [ ${BVH_BUFFER.size} ] or triangle [ ${O_TRIANGLE_BUFFER.size} ] buffer is too small for ${NUM_TRIS} triangles`)
  }
  if (RAYS.length % 8 !== 0) throw new RangeError('raycast: RAYS must hold 8 floats per ray')
  const count = RAYS.length / 8
  const R = device.createBuffer({ size: RAYS.length * 4 }, RAYS)
  const H = device.createBuffer({ size: count * 16 })
  if (count > 0) {
    addon.raycast(device.ctx, BVH_BUFFER.ptr, O_TRIANGLE_BUFFER.ptr, NUM_TRIS,
      R.ptr, count, H.ptr)
  }
  const hits = new Float32Array(await H.read())
  R.destroy()
  H.destroy()
  return hits
}

// ---- serving (no counterpart in the reference) ---------------------------------

const REQUEST_KINDS = { build: 0, refit: 1 }
const SOURCES = { host: 0, device: 1, uniform: 2, jitter: 3 }

// A device-resident build/refit service: requests overlap on the GPU in
// execution slots, are admitted by HBM budget (least recently used idle
// meshes are evicted), and built trees stay resident.
class BVHServer {
  constructor (options) {
    const o = options || {}
    this.ordinal = o.ordinal || 0
    this.srv = addon.createServer(this.ordinal, o.hbmBudget || 0, o.slots || 4)
    if (o.divMode) addon.serverSetDivMode(this.srv, DIV_MODES[o.divMode])
    this.reader = null
  }

  // request: { kind: 'build' | 'refit', meshId, source: 'host' | 'device' |
  // 'uniform' | 'jitter', tri (Float32Array for 'host', LBVHBuffer for
  // 'device'), n, seed, param } -> request id (BigInt)
  submit (request) {
    const r = request
    const kind = REQUEST_KINDS[r.kind]
    const source = SOURCES[r.source || 'host']
    if (kind === undefined || source === undefined) {
      throw new TypeError('unknown request kind or source')
    }
    let tri = null
    if (r.source === 'device') tri = r.tri.ptr
    else if (r.tri !== undefined && r.tri !== null) tri = toFloat32(r.tri, r.n)
    return addon.serverSubmit(this.srv, kind, BigInt(r.meshId), source, tri, r.n,
      BigInt(r.seed || 0), r.param || 0)
  }

  // Completions not yet returned: [{ requestId, meshId, kind, status, n,
  // slot, submitS, startS, doneS, digestBvh, digestTri }]; times in seconds
  // (steady clock); digests (BigInt) only after setDigest(true).
  poll () { return addon.serverPoll(this.srv) }

  // Later requests digest their results, for checkers of the service.
  setDigest (enable) { addon.serverSetDigest(this.srv, !!enable) }

  // Cancels a request still waiting for admission (it completes with status
  // 9); throws with code 7 if it already runs, 8 if it has completed.
  cancel (requestId) { addon.serverCancel(this.srv, BigInt(requestId)) }

  // Later requests not admitted within `seconds` complete with status 10.
  setMaxWait (seconds) { addon.serverSetMaxWait(this.srv, seconds) }

  drain () { return addon.serverDrain(this.srv) }

  evict (meshId) { addon.serverEvict(this.srv, BigInt(meshId)) }

  stats () { return addon.serverStats(this.srv) }

  // The resident mesh's BVH_BUFFER and O_TRIANGLE_BUFFER bytes.
  async readMesh (meshId) {
    const m = addon.serverMesh(this.srv, BigInt(meshId))
    if (this.reader === null) this.reader = createDevice({ ordinal: this.ordinal })
    const ctx = this.reader.ctx
    return {
      n: m.n,
      bvh: addon.download(ctx, m.bvh, 64 * m.n),
      otri: addon.download(ctx, m.tri, 48 * m.n)
    }
  }
}

function createServer (options) {
  return new BVHServer(options)
}

// ---- per-stage shims (the reference's execute() functions) -----------------

function initAABB_ZidxKernel (device) {
  return { execute }

  // aabb-z-idx.js:57-148
  async function execute (TRIANGLE_BUFFER, size, bounds) {
    if (TRIANGLE_BUFFER.size !== 48 * size) {
      console.warn(`in AABB/Z-index: buffer size [ ${TRIANGLE_BUFFER.size} ] does not match requested size [ ${size} ]`)
      return
    }
    const AABB_BUFFER = device.createBuffer({ size: size * 32 })
    const Z_IDX_BUFFER = device.createBuffer({ size: size * 4 })
    if (size > 0) {
      addon.aabbMorton(device.ctx, TRIANGLE_BUFFER.ptr, size, toBounds(bounds),
        AABB_BUFFER.ptr, Z_IDX_BUFFER.ptr)
    }
    await device.queue.onSubmittedWorkDone()
    return { AABB_BUFFER, Z_IDX_BUFFER }
  }
}

function initRadixSortKernel (device) {
  return { execute }

  // radix-sort.js:161-500: destructive, sorted keys replace valBuffer's.
  // (The reference's guard calls console.warning, which does not exist; a
  // warning is printed here instead.)
  async function execute (valBuffer, size) {
    if (valBuffer.size !== size * 4) {
      console.warn(`in radix sort: buffer size [ ${valBuffer.size} ] does not match requested size [ ${size} ]`)
      return
    }
    const IDX_BUFFER = device.createBuffer({ size: valBuffer.size })
    if (size > 0) addon.sort(device.ctx, valBuffer.ptr, IDX_BUFFER.ptr, size)
    await device.queue.onSubmittedWorkDone()
    return { IDX_BUFFER }
  }
}

function initRadixTreeKernel (device) {
  return { execute }

  // radix-tree.js:50-120
  async function execute (KEY_BUFFER, SIZE) {
    if (KEY_BUFFER.size !== 4 * SIZE) {
      console.warn(`in radix tree: buffer size [ ${KEY_BUFFER.size} ] does not match requested size [ ${SIZE} ]`)
      return
    }
    const PARENT_BUFFER = device.createBuffer({ size: SIZE * 8 })
    if (SIZE > 0) addon.radixTree(device.ctx, KEY_BUFFER.ptr, SIZE, PARENT_BUFFER.ptr)
    await device.queue.onSubmittedWorkDone()
    return { PARENT_BUFFER }
  }
}

function initBVHUpPassKernel (device) {
  return { execute }

  // bvh-up-pass.js:70-156.  The reference has no size guard here (WebGPU
  // bounds-checks the shader's reads); buffers too small for `size` are
  // refused the way the other shims refuse a mismatch.
  async function execute (IDX_BUFFER, AABB_BUFFER, PARENT_BUFFER, size) {
    const need = [[IDX_BUFFER, 4, 'index'], [AABB_BUFFER, 32, 'AABB'],
      [PARENT_BUFFER, 8, 'parent']]
    for (const [buf, per, name] of need) {
      if (buf.size < per * size) {
        console.warn(`in BVH up pass: ${name} buffer size [ ${buf.size} ] is too small for requested size [ ${size} ]`)
        return
      }
    }
    const BVH_BUFFER = device.createBuffer({ size: size * 64 })
    if (size > 0) {
      addon.upPass(device.ctx, IDX_BUFFER.ptr, AABB_BUFFER.ptr, PARENT_BUFFER.ptr,
        size, BVH_BUFFER.ptr)
    }
    await device.queue.onSubmittedWorkDone()
    return { BVH_BUFFER }
  }
}

function initRearrangeKernel (device) {
  return { execute }

  // rearrange.js:57-139
  async function execute (I_TRIANGLE_BUFFER, INDEX_BUFFER, size) {
    if (I_TRIANGLE_BUFFER.size !== 48 * size) {
      console.warn(`in rearrange: triangle buffer size [ ${I_TRIANGLE_BUFFER.size} ] does not match requested size [ ${size} ]`)
      return
    }
    if (INDEX_BUFFER.size !== 4 * size) {
      console.warn(`in rearrange: index buffer size [ ${INDEX_BUFFER.size} ] does not match requested size [ ${size} ]`)
      return
    }
    const O_TRIANGLE_BUFFER = device.createBuffer({ size: I_TRIANGLE_BUFFER.size })
    if (size > 0) {
      addon.rearrange(device.ctx, I_TRIANGLE_BUFFER.ptr, INDEX_BUFFER.ptr, size,
        O_TRIANGLE_BUFFER.ptr)
    }
    await device.queue.onSubmittedWorkDone()
    return { O_TRIANGLE_BUFFER }
  }
}

// The reference loads its scripts as browser globals (index.html:3-12);
// installGlobals() defines the same names on the global object.
function installGlobals (target) {
  const g = target || global
  g.initBVHBuild = initBVHBuild
  g.initAABB_ZidxKernel = initAABB_ZidxKernel
  g.initRadixSortKernel = initRadixSortKernel
  g.initRadixTreeKernel = initRadixTreeKernel
  g.initBVHUpPassKernel = initBVHUpPassKernel
  g.initRearrangeKernel = initRearrangeKernel
  return g
}

module.exports = {
  createDevice,
  initBVHBuild,
  initBVHBatchBuild,
  raycast,
  createServer,
  BVHServer,
  initAABB_ZidxKernel,
  initRadixSortKernel,
  initRadixTreeKernel,
  initBVHUpPassKernel,
  initRearrangeKernel,
  installGlobals,
  LBVHBuffer,
  LBVHDevice,
  addon,
  _internal: { toFloat32, toBounds, uploadExtent }
}
