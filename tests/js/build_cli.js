'use strict'
// Runs the reference-shaped JS API on the GPU and writes its outputs.
//   node build_cli.js <tri.f32> <n> <bounds.json|-> <out_prefix> [mode]
// mode "build" (default): initBVHBuild(device).build(TRI_ARRAY, N, BOUNDS)
//   with TRI_ARRAY a plain JS Array, as main.js:88 passes it.
// mode "stages": the five per-stage execute() shims in build-bvh.js order.
// Writes <out_prefix>.bvh and <out_prefix>.otri (and .keys/.idx/.par for
// stages).
const fs = require('fs')
const path = require('path')
const bvh = require(path.join(__dirname, '..', '..', 'webgpu-lvbh-demo_amd', 'bvh'))

async function main () {
  const [triPath, nStr, boundsArg, prefix, modeArg] = process.argv.slice(2)
  const n = parseInt(nStr, 10)
  const mode = modeArg || 'build'
  const raw = fs.readFileSync(triPath)
  const f32 = new Float32Array(raw.buffer, raw.byteOffset, raw.length / 4)
  const bounds = boundsArg === '-' ? null : JSON.parse(fs.readFileSync(boundsArg, 'utf8'))
  const device = bvh.createDevice({ ordinal: 0 })
  const write = async (buf, ext) => {
    fs.writeFileSync(prefix + ext, Buffer.from(await buf.read()))
  }
  if (mode === 'build') {
    const TRI_ARR = Array.from(f32, (x) => x) // a plain Array of numbers
    const { BVH_BUFFER, O_TRIANGLE_BUFFER } = await bvh.initBVHBuild(device).build(TRI_ARR, n, bounds)
    if (BVH_BUFFER.size !== 64 * n || O_TRIANGLE_BUFFER.size !== 48 * n) throw new Error('sizes')
    await write(BVH_BUFFER, '.bvh')
    await write(O_TRIANGLE_BUFFER, '.otri')
  } else {
    const I_TRI = device.createBuffer({ size: n * 48 }, f32)
    const { AABB_BUFFER, Z_IDX_BUFFER } = await bvh.initAABB_ZidxKernel(device).execute(I_TRI, n, bounds)
    const { IDX_BUFFER } = await bvh.initRadixSortKernel(device).execute(Z_IDX_BUFFER, n)
    const { PARENT_BUFFER } = await bvh.initRadixTreeKernel(device).execute(Z_IDX_BUFFER, n)
    const { BVH_BUFFER } = await bvh.initBVHUpPassKernel(device).execute(IDX_BUFFER, AABB_BUFFER, PARENT_BUFFER, n)
    const { O_TRIANGLE_BUFFER } = await bvh.initRearrangeKernel(device).execute(I_TRI, IDX_BUFFER, n)
    // a size mismatch warns and returns undefined, as in the reference
    const bad = await bvh.initRadixTreeKernel(device).execute(Z_IDX_BUFFER, n + 1)
    if (bad !== undefined) throw new Error('size guard')
    await write(BVH_BUFFER, '.bvh')
    await write(O_TRIANGLE_BUFFER, '.otri')
    await write(Z_IDX_BUFFER, '.keys')
    await write(IDX_BUFFER, '.idx')
    await write(PARENT_BUFFER, '.par')
    for (const b of [AABB_BUFFER, Z_IDX_BUFFER, PARENT_BUFFER, IDX_BUFFER, I_TRI]) b.destroy()
  }
  console.log(JSON.stringify({ ok: true, n, mode, bytesInUse: device.bytesInUse() }))
}

main().catch((e) => { console.error(e); process.exit(1) })
